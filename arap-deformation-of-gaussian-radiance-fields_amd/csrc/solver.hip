// Embedded-deformation ("ARAP") Gauss-Newton on gfx950 (Deform.cpp:95-581).
//
// Unknowns: 12 per free node, x[12f + 3c + r] = R(r,c), x[12f + 9 + r] = t_r.
// Residual groups (each a few rows, sqrt(w)-weighted):
//   rot(i)      6 rows  {c0.c1, c0.c2, c1.c2, |c0|^2-1, |c1|^2-1, |c2|^2-1}
//   reg(i,t)    3 rows  s_ik (R_i d + g_i + t_i - g_k - t_k), d = g_k - g_i
//   static(i,t) 3 rows  -t_k   (static i, free k)
//   con(c)      3 rows  sum over members v of
//                          [sum_j w_j (free j ? R_j (v - g_j) + g_j + t_j : v) - centre_c]
//                       plain mode: one member, centre = its aim; centre mode
//                       (Deform.cpp:483-529): the block's key nodes, centre =
//                       the float mean of their aims
// The Jacobian keeps the reference's asymmetry: d reg / d t_i carries no s_ik
// (Deform.cpp:257 vs :261).
//
// J^T J is never assembled. Products are applied matrix-free in the adjoint
// form: for every free node ("owner") a 16-lane group walks the residual groups
// that touch it — rot, its out-edges, its in-edges and the constraints whose
// source vertex uses it — recomputes J_g p and accumulates J_{g,i}^T (J_g p).
// The linear system is solved by PCG (FP32 vectors, FP64 dots) with a two-level
// additive preconditioner: exact 12x12 block-Jacobi plus a coarse space of one
// affine map per node aggregate (R_j = A_a, t_j = A_a (g_j - c_a) + b_a),
// whose 12n_a x 12n_a Galerkin matrix is assembled and inverted once per frame.
// The step-halving line search evaluates all 50 halvings of the reference
// (alpha = 1 .. 2^-49) in one pass and takes the first with E <= E_old.
//
// No FMA contraction: every kernel that shares a phase body rounds the same
// way, so the persistent and the per-phase launch paths agree bit for bit.
#pragma clang fp contract(off)
#include "common.hpp"
#include "solver.hpp"

#include <algorithm>

namespace arapgs {

namespace {

constexpr int LANES = 16;  // lanes per owner node

// Pointers read from device memory are generic (flat) pointers to the
// compiler, and flat loads and stores cost more than global ones. A pointer
// rebuilt as an integer in the global address space lets every access through
// it compile to a global load or store (all solver buffers are device memory).
template <typename T>
__device__ __forceinline__ T* gptr(T* p) {
    typedef __attribute__((address_space(1))) T GT;
    return (T*)(GT*)(uintptr_t)p;
}

// This block's entry, every pointer in the global address space.
__device__ __forceinline__ SolverEntry load_entry(const SolverEntry* __restrict__ B) {
    SolverEntry E = B[blockIdx.z];
    SolverDev& S = E.S;
    S.free_nodes = gptr(S.free_nodes); S.static_nodes = gptr(S.static_nodes); S.free_id = gptr(S.free_id);
    S.static_id = gptr(S.static_id); S.nbr = gptr(S.nbr); S.edge_scale = gptr(S.edge_scale);
    S.src_idx = gptr(S.src_idx); S.src_w = gptr(S.src_w); S.con_ptr = gptr(S.con_ptr); S.con_mem = gptr(S.con_mem);
    S.caim_ptr = gptr(S.caim_ptr); S.caim_node = gptr(S.caim_node); S.in_ptr = gptr(S.in_ptr);
    S.in_edge = gptr(S.in_edge); S.cm_ptr = gptr(S.cm_ptr); S.cm_con = gptr(S.cm_con); S.cm_sbeg = gptr(S.cm_sbeg);
    S.cm_src = gptr(S.cm_src); S.multi_con = gptr(S.multi_con); S.agg_ptr = gptr(S.agg_ptr);
    S.pair_ptr = gptr(S.pair_ptr); S.pair_j = gptr(S.pair_j); S.term_ptr = gptr(S.term_ptr); S.terms = gptr(S.terms);
    S.agg_member = gptr(S.agg_member); S.agg_of = gptr(S.agg_of); S.agg_center = gptr(S.agg_center);
    S.g = gptr(S.g); S.aim = gptr(S.aim);
    S.wt_out_nb = gptr(S.wt_out_nb); S.wt_in_nb = gptr(S.wt_in_nb); S.wt_cm_rng = gptr(S.wt_cm_rng);
    S.wt_con_nb = gptr(S.wt_con_nb); S.wt_out = gptr(S.wt_out); S.wt_in = gptr(S.wt_in); S.wt_cm = gptr(S.wt_cm);
    S.wt_con = gptr(S.wt_con); S.wt_out_ik = gptr(S.wt_out_ik); S.wt_in_is = gptr(S.wt_in_is);
    S.wt_cm_i = gptr(S.wt_cm_i);
    E.x = gptr(E.x); E.hsol = gptr(E.hsol); E.r = gptr(E.r); E.z = gptr(E.z); E.p = gptr(E.p); E.q = gptr(E.q);
    E.binv = gptr(E.binv); E.rc = gptr(E.rc); E.acinv = gptr(E.acinv); E.gj_piv = gptr(E.gj_piv);
    E.dmat = gptr(E.dmat); E.pair_m = gptr(E.pair_m);
    E.gj_col = gptr(E.gj_col); E.hwarm = gptr(E.hwarm); E.ac64 = gptr(E.ac64); E.partial = gptr(E.partial);
    E.partial2 = gptr(E.partial2); E.rzp = gptr(E.rzp); E.e_partial = gptr(E.e_partial); E.cand_e = gptr(E.cand_e);
    E.acc_partial = gptr(E.acc_partial); E.energies = gptr(E.energies); E.cg = gptr(E.cg); E.gs = gptr(E.gs);
    E.gn_done = gptr(E.gn_done); E.cg_done = gptr(E.cg_done); E.gbar = gptr(E.gbar);
    return E;
}

__device__ __forceinline__ float xR(const float* x, int r, int c) { return x[3 * c + r]; }

// ---------------------------------------------------------------- residuals
// Residual of one group at unknowns x (+ alpha*h when h != nullptr). Returns
// the number of rows written to f (<= 6).
// Residual rows of group g at x + alpha h (h == nullptr: at x), evaluated in
// FP64 from the FP32 unknowns (each candidate unknown is first rounded to the
// FP32 value the accepted step would store). FP64 keeps the cancellations of
// the affine parts (R d + g_i + t_i - g_k - t_k, the constraint sums) out of
// the gradient, so near the minimum the step shrinks the way the reference's
// double-precision solve does and its stop rule fires at the same iteration.
__device__ __forceinline__ int group_residual_rows(const SolverDev& S, const float* __restrict__ x,
                                                   const float* __restrict__ h, float alpha, int64_t g, double& f0,
                                                   double& f1, double& f2, double& f3, double& f4, double& f5) {
    auto X = [&](int node_free, int comp) -> double {
        const int64_t o = 12 * (int64_t)node_free + comp;
        return (double)(h ? x[o] + alpha * h[o] : x[o]);
    };
    const double sw_rot = S.sw_rot, sw_reg = S.sw_reg, sw_con = S.sw_con;
    if (g < S.nF) {
        const int fi = (int)g;
        double c[3][3];
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int r = 0; r < 3; r++) c[j][r] = X(fi, 3 * j + r);
        auto dot = [&](int a, int b) { return c[a][0] * c[b][0] + c[a][1] * c[b][1] + c[a][2] * c[b][2]; };
        f0 = sw_rot * dot(0, 1);
        f1 = sw_rot * dot(0, 2);
        f2 = sw_rot * dot(1, 2);
        f3 = sw_rot * (dot(0, 0) - 1.0);
        f4 = sw_rot * (dot(1, 1) - 1.0);
        f5 = sw_rot * (dot(2, 2) - 1.0);
        return 6;
    }
    g -= S.nF;
    if (g < (int64_t)S.nF * S.K) {
        const int fi = (int)(g / S.K), t = (int)(g % S.K);
        // (i, k) and k's free id from the static walk tables (the index
        // chain free_nodes -> nbr -> free_id resolved once)
        const int2 ik = S.wt_out_ik[g];
        const int i = ik.x, k = ik.y;
        const int fk = S.wt_out_nb[g];
        const double s = S.edge_scale[(int64_t)i * S.K + t];
        const float* gi = S.g + 3 * i;
        const float* gk = S.g + 3 * k;
        const double d[3] = {(double)gk[0] - gi[0], (double)gk[1] - gi[1], (double)gk[2] - gi[2]};
        auto reg = [&](int r) {
            const double tk = fk >= 0 ? X(fk, 9 + r) : 0.0;
            const double v = X(fi, r) * d[0] + X(fi, 3 + r) * d[1] + X(fi, 6 + r) * d[2] + (double)gi[r] +
                             X(fi, 9 + r) - (double)gk[r] - tk;
            return sw_reg * v * s;
        };
        f0 = reg(0);
        f1 = reg(1);
        f2 = reg(2);
        f3 = f4 = f5 = 0.0;
        return 3;
    }
    g -= (int64_t)S.nF * S.K;
    if (g < (int64_t)S.nS * S.K) {
        const int si = (int)(g / S.K), t = (int)(g % S.K);
        const int i = S.static_nodes[si];
        const int k = (int)S.nbr[(int64_t)i * S.K + t];
        const int fk = S.free_id[k];
        f0 = sw_reg * -(fk >= 0 ? X(fk, 9) : 0.0);
        f1 = sw_reg * -(fk >= 0 ? X(fk, 10) : 0.0);
        f2 = sw_reg * -(fk >= 0 ? X(fk, 11) : 0.0);
        f3 = f4 = f5 = 0.0;
        return 3;
    }
    g -= (int64_t)S.nS * S.K;
    // constraint group: sum over its members of sw_con (blend(member) - centre)
    const int cid = (int)g;
    float ctr[3] = {0.0f, 0.0f, 0.0f};
    const int ab = S.caim_ptr[cid], ae = S.caim_ptr[cid + 1];
    for (int q = ab; q < ae; q++) {
        const int a = S.caim_node[q];
#pragma unroll
        for (int r = 0; r < 3; r++) ctr[r] += S.aim[3 * a + r];
    }
#pragma unroll
    for (int r = 0; r < 3; r++) ctr[r] = ctr[r] / (float)(ae - ab);
    f0 = f1 = f2 = 0.0;
    for (int q = S.con_ptr[cid]; q < S.con_ptr[cid + 1]; q++) {
        const int c = S.con_mem[q];
        const double v0 = S.g[3 * c], v1 = S.g[3 * c + 1], v2 = S.g[3 * c + 2];
        double n0 = 0.0, n1 = 0.0, n2 = 0.0;
        for (int j = 0; j < S.K; j++) {
            const int nj = (int)S.src_idx[(int64_t)c * S.K + j];
            const double w = S.src_w[(int64_t)c * S.K + j];
            const int fj = S.wt_con_nb[(int64_t)q * S.K + j];  // free_id[nj]
            if (fj < 0) {
                n0 += w * v0;
                n1 += w * v1;
                n2 += w * v2;
            } else {
                const float* gj = S.g + 3 * nj;
                const double d0 = v0 - gj[0], d1 = v1 - gj[1], d2 = v2 - gj[2];
                n0 += w * (X(fj, 0) * d0 + X(fj, 3) * d1 + X(fj, 6) * d2 + (double)gj[0] + X(fj, 9));
                n1 += w * (X(fj, 1) * d0 + X(fj, 4) * d1 + X(fj, 7) * d2 + (double)gj[1] + X(fj, 10));
                n2 += w * (X(fj, 2) * d0 + X(fj, 5) * d1 + X(fj, 8) * d2 + (double)gj[2] + X(fj, 11));
            }
        }
        f0 += sw_con * (n0 - (double)ctr[0]);
        f1 += sw_con * (n1 - (double)ctr[1]);
        f2 += sw_con * (n2 - (double)ctr[2]);
    }
    f3 = f4 = f5 = 0.0;
    return 3;
}

// The rows travel as six scalars (3-row groups set rows 3-5 to zero) so that
// every branch keeps them in registers.
__device__ __forceinline__ int group_residual(const SolverDev& S, const float* __restrict__ x,
                                              const float* __restrict__ h, float alpha, int64_t g, double f[6]) {
    double f0, f1, f2, f3, f4, f5;
    const int nr = group_residual_rows(S, x, h, alpha, g, f0, f1, f2, f3, f4, f5);
    f[0] = f0; f[1] = f1; f[2] = f2; f[3] = f3; f[4] = f4; f[5] = f5;
    return nr;
}

__device__ __forceinline__ double block_reduce_d(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < nw; w++) r += sh[w];
    __syncthreads();
    return r;  // valid in thread 0
}

// The line search's candidates past the full step are evaluated only when the
// full step was rejected: with first_e set, a launch for candidates c0 >= 1
// returns at once when E(x + h) = first_e[0] <= E_old.
__device__ __forceinline__ bool first_accepted(const double* first_e, const GnScalars* gs) {
    return first_e && !(first_e[0] > gs->e_old);
}

// Energies of x + 2^-c h for c in [c0, c0 + gridDim.y) (use_h == 0: the
// energy of x). blockIdx.y = c - c0. Per-block FP64 partials at
// partial[c * e_blocks + block], reduced by reduce_partials_kernel. use_skip:
// nothing once the solve stopped; lazy: nothing when the full step (cand_e[0])
// was accepted. x_ext (single-scene calls) replaces the entry's x.
__global__ void __launch_bounds__(256) energy_kernel(const SolverEntry* __restrict__ B, int use_h, int use_skip, int c0,
                                                     int lazy, const float* __restrict__ x_ext) {
    const SolverEntry E = load_entry(B);
    if ((int)blockIdx.x >= E.e_blocks) return;
    if (use_skip && *E.gn_done) return;
    if (lazy && first_accepted(E.cand_e, E.gs)) return;
    const SolverDev& S = E.S;
    const float* __restrict__ x = x_ext ? x_ext : E.x;
    const float* __restrict__ h = use_h ? E.hsol : nullptr;
    double* __restrict__ partial = E.e_partial;
    __shared__ double sh[8];
    const int c = blockIdx.y + c0;
    const float alpha = ldexpf(1.0f, -c);
    double acc = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < S.n_groups;
         g += (int64_t)E.e_blocks * blockDim.x) {
        double f[6];
        const int nr = group_residual(S, x, h, alpha, g, f);
#pragma unroll
        for (int r = 0; r < 6; r++)
            if (r < nr) acc += f[r] * f[r];
    }
    const double t = block_reduce_d(acc, sh);
    if (threadIdx.x == 0) partial[(int64_t)c * E.e_blocks + blockIdx.x] = t;
}

// out_ext (single-scene calls) replaces the entry's cand_e.
__global__ void reduce_partials_kernel(const SolverEntry* __restrict__ B, int c0, int ncand, int use_skip, int lazy,
                                       double* __restrict__ out_ext) {
    const SolverEntry E = load_entry(B);
    if (use_skip && *E.gn_done) return;
    if (lazy && first_accepted(E.cand_e, E.gs)) return;
    const double* __restrict__ partial = E.e_partial;
    const int nblocks = E.e_blocks;
    double* __restrict__ out = out_ext ? out_ext : E.cand_e;
    __shared__ double sh[8];
    for (int c = c0; c < c0 + ncand; c++) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc += partial[(int64_t)c * nblocks + b];
        const double t = block_reduce_d(acc, sh);
        if (threadIdx.x == 0) out[c] = t;
    }
}

// ---------------------------------------------------------------- owner walk
// Visit every residual group touching free node fi with its Jacobian block
// wrt fi and (WITH_JV) the group's J*v for a given vector v (12 per free node).
// Item 0 is the rotation group: fr(Ji[6][9] (the R columns), Jv[6]). Every
// other group has 3 rows whose block wrt a node is I_3-structured — row r
// holds (a_0, a_1, a_2) at the R columns 3c + r and b at the t column 9 + r —
// so it is passed as f3(a[3], b, Jv[3]).
// Items: 0 = rot, 1..K = out-edges, then in-edges, then constraint memberships.
// KT > 0: the neighbour count is a compile-time constant (== S.K), so the
// constraint item's K-node loop unrolls and its gathers issue together.
template <int KT = 0, bool WITH_JV = true, typename FnRot, typename Fn3>
__device__ __forceinline__ void owner_item(const SolverDev& S, int fi, int item, const float* __restrict__ x,
                                           const float* __restrict__ v, FnRot&& fr, Fn3&& f3) {
    if (item == 0) {
        // rot(i): d(c_a . c_b)/d c_a = c_b, d|c_j|^2 / d c_j = 2 c_j
        float Ji[6][9];
        float Jv[6];
#pragma unroll
        for (int a = 0; a < 6; a++)
#pragma unroll
            for (int b = 0; b < 9; b++) Ji[a][b] = 0.0f;
        const float* xi = x + 12 * (int64_t)fi;
        const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                Ji[p][3 * pa[p] + c] = S.sw_rot * xi[3 * pb[p] + c];
                Ji[p][3 * pb[p] + c] += S.sw_rot * xi[3 * pa[p] + c];
            }
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int c = 0; c < 3; c++) Ji[3 + j][3 * j + c] = 2.0f * S.sw_rot * xi[3 * j + c];
        const int64_t vo = 12 * (int64_t)fi;
#pragma unroll
        for (int a = 0; a < 6; a++) {
            float s = 0.0f;
            if (WITH_JV) {
#pragma unroll
                for (int b = 0; b < 9; b++) s += Ji[a][b] * v[vo + b];
            }
            Jv[a] = s;
        }
        fr(Ji, Jv);
        return;
    }
    float a[3] = {0.0f, 0.0f, 0.0f}, b = 0.0f;
    float Jv[3] = {0.0f, 0.0f, 0.0f};
    item -= 1;
    // the walk tables hold the frame's edge vectors, scales, group blocks and
    // source slots, computed with the same operations as from the node
    // positions directly (walk_table_kernel)
    if (item < S.K) {
        const int64_t o = (int64_t)fi * S.K + item;
        const float4 e = S.wt_out[o];
        const int fk = S.wt_out_nb[o];
        const float s = e.w;
        const float d[3] = {e.x, e.y, e.z};
#pragma unroll
        for (int c = 0; c < 3; c++) a[c] = S.sw_reg * s * d[c];
        b = S.sw_reg;
        if (WITH_JV) {
            const int64_t vo = 12 * (int64_t)fi;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                float jv = S.sw_reg * (s * (v[vo + r] * d[0] + v[vo + 3 + r] * d[1] + v[vo + 6 + r] * d[2]) + v[vo + 9 + r]);
                if (fk >= 0) jv -= S.sw_reg * s * v[12 * (int64_t)fk + 9 + r];
                Jv[r] = jv;
            }
        }
        f3(a, b, Jv);
        return;
    }
    item -= S.K;
    const int ib = S.in_ptr[fi];
    const int n_in = S.in_ptr[fi + 1] - ib;
    if (item < n_in) {
        const int q = ib + item;
        const int fs = S.wt_in_nb[q];
        if (fs < 0) {
            // static(src, t): rows -sw_reg t_i
            b = -S.sw_reg;
            if (WITH_JV) {
#pragma unroll
                for (int r = 0; r < 3; r++) Jv[r] = -S.sw_reg * v[12 * (int64_t)fi + 9 + r];
            }
        } else {
            const float4 e = S.wt_in[q];
            const float s = e.w;
            b = -S.sw_reg * s;
            if (WITH_JV) {
                const float d[3] = {e.x, e.y, e.z};
                const int64_t so = 12 * (int64_t)fs;
#pragma unroll
                for (int r = 0; r < 3; r++)
                    Jv[r] = S.sw_reg * (s * (v[so + r] * d[0] + v[so + 3 + r] * d[1] + v[so + 6 + r] * d[2]) + v[so + 9 + r]) -
                            S.sw_reg * s * v[12 * (int64_t)fi + 9 + r];
            }
        }
        f3(a, b, Jv);
        return;
    }
    item -= n_in;
    const int ent = S.cm_ptr[fi] + item;
    {
        // J of the group wrt node i: sum over the (member, slot) entries that reach i
        const float4 ab = S.wt_cm[ent];
        a[0] = ab.x; a[1] = ab.y; a[2] = ab.z;
        b = ab.w;
    }
    if (WITH_JV) {
        const int KK = KT > 0 ? KT : S.K;
        const int2 rng = S.wt_cm_rng[ent];
        for (int m = rng.x; m < rng.y; m++) {
#pragma unroll
            for (int j = 0; j < (KT > 0 ? KT : KNN_MAX_T); j++) {
                if (KT == 0 && j >= KK) break;
                const int64_t o = (int64_t)m * KK + j;
                const int fj = S.wt_con_nb[o];
                if (fj < 0) continue;
                const float4 e = S.wt_con[o];
                const float w = e.w;
                const float dd[3] = {e.x, e.y, e.z};
                const int64_t jo = 12 * (int64_t)fj;
#pragma unroll
                for (int r = 0; r < 3; r++) Jv[r] += S.sw_con * w * (v[jo + r] * dd[0] + v[jo + 3 + r] * dd[1] + v[jo + 6 + r] * dd[2] + v[jo + 9 + r]);
            }
        }
    }
    f3(a, b, Jv);
}

// The frame's walk tables (node positions fixed for the solve), one record per
// thread: out-edges, in-edges, memberships, then constraint members (K source
// slots each). Every value is formed with the operations the owner walk used
// on the node positions.
__device__ __forceinline__ int64_t walk_records(const SolverDev& S) {
    return (int64_t)S.nF * S.K + S.n_in + S.n_cm + S.n_mem;
}

__global__ void __launch_bounds__(256) walk_table_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    const SolverDev& S = E.S;
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int K = S.K;
    const float* __restrict__ g = S.g;
    if (r < (int64_t)S.nF * K) {
        const int2 ik = S.wt_out_ik[r];
        const float* gi = g + 3 * ik.x;
        const float* gk = g + 3 * ik.y;
        S.wt_out[r] = make_float4(gk[0] - gi[0], gk[1] - gi[1], gk[2] - gi[2],
                                  S.edge_scale[(int64_t)ik.x * K + r % K]);
        return;
    }
    r -= (int64_t)S.nF * K;
    if (r < S.n_in) {
        const int2 is = S.wt_in_is[r];
        const float* gi = g + 3 * is.x;
        const float* gs = g + 3 * is.y;
        S.wt_in[r] = make_float4(gi[0] - gs[0], gi[1] - gs[1], gi[2] - gs[2], S.edge_scale[S.in_edge[r]]);
        return;
    }
    r -= S.n_in;
    if (r < S.n_cm) {
        const int ent = (int)r;
        const float* gi = g + 3 * S.wt_cm_i[ent];
        float a[3] = {0.0f, 0.0f, 0.0f}, b = 0.0f;
        for (int q = S.cm_sbeg[ent]; q < S.cm_sbeg[ent + 1]; q++) {
            const int e = S.cm_src[q];
            const int c = e / K;
            const float wi = S.src_w[e];
            const float dd[3] = {g[3 * c] - gi[0], g[3 * c + 1] - gi[1], g[3 * c + 2] - gi[2]};
#pragma unroll
            for (int cc = 0; cc < 3; cc++) a[cc] += S.sw_con * wi * dd[cc];
            b += S.sw_con * wi;
        }
        S.wt_cm[ent] = make_float4(a[0], a[1], a[2], b);
        return;
    }
    r -= S.n_cm;
    if (r < S.n_mem) {
        const int c = S.con_mem[r];
        const float vv[3] = {g[3 * c], g[3 * c + 1], g[3 * c + 2]};
        for (int j = 0; j < K; j++) {
            const int nj = (int)S.src_idx[(int64_t)c * K + j];
            const float* gj = g + 3 * nj;
            S.wt_con[r * K + j] = make_float4(vv[0] - gj[0], vv[1] - gj[1], vv[2] - gj[2], S.src_w[(int64_t)c * K + j]);
        }
    }
}

__device__ __forceinline__ int owner_items(const SolverDev& S, int fi) {
    return 1 + S.K + (S.in_ptr[fi + 1] - S.in_ptr[fi]) + (S.cm_ptr[fi + 1] - S.cm_ptr[fi]);
}

__device__ __forceinline__ float group_reduce16(float v) {
    v += __shfl_xor(v, 8, 16);
    v += __shfl_xor(v, 4, 16);
    v += __shfl_xor(v, 2, 16);
    v += __shfl_xor(v, 1, 16);
    return v;
}

__device__ __forceinline__ double group_reduce16d(double v) {
    v += __shfl_xor(v, 8, 16);
    v += __shfl_xor(v, 4, 16);
    v += __shfl_xor(v, 2, 16);
    v += __shfl_xor(v, 1, 16);
    return v;
}

// b = -J^T f, diagonal block D_i = sum J_i^T J_i and its inverse; |b|^2 partials.
// D_i = (the rotation group's 9x9 block, lane 0 only) + the I_3-structured sum
// of the 3-row groups, C[p][q] = sum a'_p a'_q with a' = (a_0, a_1, a_2, b),
// placed at (3p + r, 3q + r) (index 3 -> 9 + r): 10 sums per lane to reduce.
__global__ void __launch_bounds__(256) rhs_diag_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    if ((int)blockIdx.x >= E.nb_owner || *E.gn_done) return;
    const SolverDev& S = E.S;
    const float* __restrict__ x = E.x;
    float* __restrict__ b = E.r;
    float* __restrict__ binv = E.binv;
    double* __restrict__ partial = E.partial;
    __shared__ float sD[256 / LANES][144];
    __shared__ double shd[8];
    const int grp = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    const int fi = blockIdx.x * (256 / LANES) + grp;
    double bl[12];
    float C[10];
#pragma unroll
    for (int a = 0; a < 12; a++) bl[a] = 0.0;
#pragma unroll
    for (int a = 0; a < 10; a++) C[a] = 0.0f;
    if (fi < S.nF) {
        const int ni = owner_items(S, fi);
        for (int it = lane; it < ni; it += LANES) {
            // residual value of the group for the gradient: evaluate f via J*x
            // identities is not possible for the affine parts, so recompute f.
            double f[6];
            int64_t gid;
            // map item -> global group id for the residual
            {
                int item = it;
                if (item == 0) gid = fi;
                else if ((item -= 1) < S.K) gid = S.nF + (int64_t)fi * S.K + item;
                else {
                    item -= S.K;
                    const int n_in = S.in_ptr[fi + 1] - S.in_ptr[fi];
                    if (item < n_in) {
                        const int e = S.in_edge[S.in_ptr[fi] + item];
                        const int src = e / S.K, t = e % S.K;
                        const int fs = S.free_id[src];
                        gid = fs >= 0 ? S.nF + (int64_t)fs * S.K + t
                                      : S.nF + (int64_t)S.nF * S.K + (int64_t)S.static_id[src] * S.K + t;
                    } else {
                        item -= n_in;
                        gid = S.nF + (int64_t)(S.nF + S.nS) * S.K + S.cm_con[S.cm_ptr[fi] + item];
                    }
                }
            }
            group_residual(S, x, nullptr, 0.0f, gid, f);
            owner_item<0, false>(
                S, fi, it, x, x,
                [&](float (&Ji)[6][9], float (&)[6]) {
#pragma unroll
                    for (int r = 0; r < 6; r++) {
#pragma unroll
                        for (int a = 0; a < 9; a++) bl[a] -= (double)Ji[r][a] * f[r];
                    }
                },
                [&](float (&a3)[3], float bb, float (&)[3]) {
                    const float v4[4] = {a3[0], a3[1], a3[2], bb};
#pragma unroll
                    for (int r = 0; r < 3; r++) {
#pragma unroll
                        for (int c = 0; c < 3; c++) bl[3 * c + r] -= (double)v4[c] * f[r];
                        bl[9 + r] -= (double)bb * f[r];
                    }
                    int o = 0;
#pragma unroll
                    for (int p = 0; p < 4; p++)
#pragma unroll
                        for (int q = p; q < 4; q++) C[o++] += v4[p] * v4[q];
                });
        }
    }
#pragma unroll
    for (int a = 0; a < 12; a++) bl[a] = group_reduce16d(bl[a]);
#pragma unroll
    for (int a = 0; a < 10; a++) C[a] = group_reduce16(C[a]);
    double nb = 0.0;
    const bool own = fi < S.nF;
    float* Dm = sD[grp];
    if (own && lane == 0) {
        for (int e = 0; e < 144; e++) Dm[e] = 0.0f;
        // the rotation group's 9x9 block J^T J (item 0, the only dense one),
        // summed over its six rows in order after the walk instead of in
        // 45 accumulators live across it
        owner_item<0, false>(
            S, fi, 0, x, x,
            [&](float (&Ji)[6][9], float (&)[6]) {
                for (int a = 0; a < 9; a++)
                    for (int c = a; c < 9; c++) {
                        float d = 0.0f;
#pragma unroll
                        for (int r = 0; r < 6; r++) d += Ji[r][a] * Ji[r][c];
                        Dm[a * 12 + c] = d;
                        Dm[c * 12 + a] = d;
                    }
            },
            [&](float (&)[3], float, float (&)[3]) {});
        int o = 0;
        for (int p = 0; p < 4; p++)
            for (int q = p; q < 4; q++) {
                for (int r = 0; r < 3; r++) {
                    const int u = p < 3 ? 3 * p + r : 9 + r;
                    const int w = q < 3 ? 3 * q + r : 9 + r;
                    Dm[u * 12 + w] += C[o];
                    if (u != w) Dm[w * 12 + u] += C[o];
                }
                o++;
            }
        for (int a = 0; a < 12; a++) {
            const float ba = (float)bl[a];
            b[12 * (int64_t)fi + a] = ba;
            nb += (double)ba * ba;
        }
    }
    __syncthreads();
    if (own) {
        // D^-1 by Gauss-Jordan on the 16 lanes of the owner's group: lane c < 12
        // holds column c; the pivot column is read from lane k by a shuffle.
        // D is symmetric positive definite (a pivot below 1e-30 is floored).
        float a[12];
#pragma unroll
        for (int r = 0; r < 12; r++) a[r] = lane < 12 ? Dm[r * 12 + lane] : 0.0f;
        if (E.cg_asm && lane < 12) {
#pragma unroll
            for (int r = 0; r < 12; r++) E.dmat[144 * (int64_t)fi + r * 12 + lane] = a[r];
        }
#pragma unroll
        for (int k = 0; k < 12; k++) {
            const float d = __shfl(a[k], k, LANES);
            const float iv = 1.0f / fmaxf(d, 1e-30f);
            const float rk = a[k] * iv;
#pragma unroll
            for (int r = 0; r < 12; r++) {
                const float colv = __shfl(a[r], k, LANES);
                if (lane == k) a[r] = (r == k) ? iv : -colv * iv;
                else a[r] = (r == k) ? rk : a[r] - colv * rk;
            }
        }
        if (lane < 12) {
#pragma unroll
            for (int r = 0; r < 12; r++) binv[144 * (int64_t)fi + r * 12 + lane] = a[r];
        }
    }
    const double t = block_reduce_d(nb, shd);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__device__ __forceinline__ double sum_partials(const double* __restrict__ p, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += p[i];
    return s;
}

// Sum of n FP64 partials by the whole workgroup (result in every thread).
__device__ __forceinline__ double block_sum_partials(const double* __restrict__ p, int n, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += p[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; w++) r += sh[w];
    __syncthreads();
    return r;
}

// PCG iteration k in three phases (no host round trips, no scalar kernels).
// The search direction's product is carried recursively (Chronopoulos-Gear):
// with s = A z,  p_k = z + beta_k p_{k-1}  and  q_k = A p_k = s + beta_k q_{k-1},
// so the operator only ever reads z, and p, q are updated node-locally.
//   apply   beta_k = rz_k / rz_{k-1} (0 at k = 0); s = A z; p, q as above
//           (in place); partial p_k.q_k per owner block
//   update  alpha = rz_k / (p_k.q_k); h += alpha p_k; r -= alpha q_k; z = Binv r;
//           rc = P^T r; partial |r|^2               (one workgroup per aggregate)
//   coarse  z += P (Ac^-1 rc); partial rz_{k+1} = r.z
// then |r|^2 <= tol^2 |b|^2 or the iteration cap ends the solve. The phases run
// either as three kernels per iteration (solver_cg_iteration) or inside one
// persistent cooperative kernel per solve with grid barriers between them
// (solver_cg_persistent). Every workgroup reduces the (few dozen) FP64 partials
// it needs itself, in the same order, so all of them take the same decisions.

// apply for owner block ob (16 free nodes, 16 lanes each)
template <int KT, int BS = 256>
__device__ __forceinline__ void apply_body(const SolverDev& S, const float* __restrict__ x,
                                           const float* __restrict__ z, float beta, bool first, float* __restrict__ p,
                                           float* __restrict__ q, double* __restrict__ partial, int ob, double* shd) {
    const int grp = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    const int fi = ob * (BS / LANES) + grp;
    float acc[12];
#pragma unroll
    for (int a = 0; a < 12; a++) acc[a] = 0.0f;
    if (fi < S.nF) {
        const int ni = owner_items(S, fi);
        for (int it = lane; it < ni; it += LANES) {
            owner_item<KT>(
                S, fi, it, x, z,
                [&](float (&Ji)[6][9], float (&Jv)[6]) {
#pragma unroll
                    for (int r = 0; r < 6; r++)
#pragma unroll
                        for (int a = 0; a < 9; a++) acc[a] += Ji[r][a] * Jv[r];
                },
                [&](float (&a3)[3], float bb, float (&Jv)[3]) {
#pragma unroll
                    for (int r = 0; r < 3; r++) {
#pragma unroll
                        for (int c = 0; c < 3; c++) acc[3 * c + r] += a3[c] * Jv[r];
                        acc[9 + r] += bb * Jv[r];
                    }
                });
        }
    }
    float mine = 0.0f;
#pragma unroll
    for (int a = 0; a < 12; a++) {
        const float t = group_reduce16(acc[a]);
        if (lane == a) mine = t;
    }
    double pq = 0.0;
    if (fi < S.nF && lane < 12) {
        const int64_t o = 12 * (int64_t)fi + lane;
        const float pn = first ? z[o] : z[o] + beta * p[o];
        const float qn = first ? mine : mine + beta * q[o];
        p[o] = pn;
        q[o] = qn;
        pq = (double)pn * qn;
    }
    const double t = block_reduce_d(pq, shd);
    if (threadIdx.x == 0) partial[ob] = t;
}

// update + restriction for aggregate a; mode 0 skips the h, r update (z0 = M^-1 r0)
__device__ __forceinline__ void update_body(const SolverDev& S, int mode, float alpha, float* __restrict__ hsol,
                                            float* __restrict__ r, const float* __restrict__ p,
                                            const float* __restrict__ q, const float* __restrict__ binv,
                                            float* __restrict__ z, float* __restrict__ rc,
                                            double* __restrict__ partial, int a, float (*sred)[12], double* shd) {
    const int beg = S.agg_ptr[a], end = S.agg_ptr[a + 1];
    const float* ca = S.agg_center + 3 * a;
    float rcl[12];
#pragma unroll
    for (int c = 0; c < 12; c++) rcl[c] = 0.0f;
    double rr = 0.0;
    // 16 lanes per node: lane c < 12 handles component c
    for (int m = beg + (int)(threadIdx.x / LANES); m < end; m += 256 / LANES) {
        const int fi = S.agg_member[m];
        const int c = threadIdx.x % LANES;
        float rv = 0.0f;
        if (c < 12) {
            const int64_t o = 12 * (int64_t)fi + c;
            if (mode == 1) {
                hsol[o] += alpha * p[o];
                r[o] -= alpha * q[o];
            }
            rv = r[o];
            rr += (double)rv * rv;
        }
        // z = Binv r (row c), needs all 12 r of the node: gather within 16 lanes
        float rn[12];
#pragma unroll
        for (int j = 0; j < 12; j++) rn[j] = __shfl(rv, j, 16);
        if (c < 12) {
            float zz = 0.0f;
#pragma unroll
            for (int j = 0; j < 12; j++) zz += binv[144 * (int64_t)fi + c * 12 + j] * rn[j];
            z[12 * (int64_t)fi + c] = zz;
        }
        // restriction P^T r: R part (9) passes through; t part r_t contributes to
        // b (identity) and to A(r,c') with weight (g_j - c_a)_c'
        const int node = S.free_nodes[fi];
        const float dj[3] = {S.g[3 * node] - ca[0], S.g[3 * node + 1] - ca[1], S.g[3 * node + 2] - ca[2]};
        if (c == 0) {
#pragma unroll
            for (int q9 = 0; q9 < 9; q9++) rcl[q9] += rn[q9];
#pragma unroll
            for (int rr3 = 0; rr3 < 3; rr3++) {
#pragma unroll
                for (int cc = 0; cc < 3; cc++) rcl[3 * cc + rr3] += dj[cc] * rn[9 + rr3];
                rcl[9 + rr3] += rn[9 + rr3];
            }
        }
    }
    if (threadIdx.x % LANES == 0)
#pragma unroll
        for (int c = 0; c < 12; c++) sred[threadIdx.x / LANES][c] = rcl[c];
    __syncthreads();
    if (threadIdx.x < 12) {
        float t = 0.0f;
        for (int g = 0; g < 256 / LANES; g++) t += sred[g][threadIdx.x];
        rc[12 * a + threadIdx.x] = t;
    }
    const double tr = block_reduce_d(rr, shd);
    if (threadIdx.x == 0) partial[a] = tr;
}

// coarse correction for aggregate a: zc_a = (Ac^-1 rc)_a, z += P_a zc_a; partial r.z
__device__ __forceinline__ void prolong_body(const SolverDev& S, const float* __restrict__ acinv,
                                             const float* __restrict__ rc, const float* __restrict__ r,
                                             float* __restrict__ z, double* __restrict__ rz_out, int a, float* zc,
                                             float (*zred)[12], double* shd) {
    const int nc = 12 * S.n_agg;
    const int ld = S.nc_pad;
    // 12 rows of Ac^-1 times rc: every thread takes columns j = tid, tid+256, ...
    // (coalesced row loads, rc read once), then a block reduction of 12 sums
    {
        float part[12];
#pragma unroll
        for (int row = 0; row < 12; row++) part[row] = 0.0f;
        const float* A0 = acinv + (int64_t)(12 * a) * ld;
#pragma unroll 4
        for (int j = threadIdx.x; j < nc; j += 256) {
            const float rj = rc[j];
#pragma unroll
            for (int row = 0; row < 12; row++) part[row] += A0[(int64_t)row * ld + j] * rj;
        }
        const int wl = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int row = 0; row < 12; row++) {
            float v = part[row];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (wl == 0) zred[wv][row] = v;
        }
        __syncthreads();
        if (threadIdx.x < 12) zc[threadIdx.x] = (zred[0][threadIdx.x] + zred[1][threadIdx.x]) +
                                              (zred[2][threadIdx.x] + zred[3][threadIdx.x]);
    }
    __syncthreads();
    const float* ca = S.agg_center + 3 * a;
    double rz = 0.0;
    const int beg = S.agg_ptr[a], end = S.agg_ptr[a + 1];
    for (int m = beg + (int)(threadIdx.x / LANES); m < end; m += 256 / LANES) {
        const int fi = S.agg_member[m];
        const int c = threadIdx.x % LANES;
        if (c >= 12) continue;
        const int node = S.free_nodes[fi];
        float add;
        if (c < 9) add = zc[c];
        else {
            const int rr3 = c - 9;
            const float dj[3] = {S.g[3 * node] - ca[0], S.g[3 * node + 1] - ca[1], S.g[3 * node + 2] - ca[2]};
            add = zc[9 + rr3] + zc[rr3] * dj[0] + zc[3 + rr3] * dj[1] + zc[6 + rr3] * dj[2];
        }
        const int64_t o = 12 * (int64_t)fi + c;
        const float zz = z[o] + add;
        z[o] = zz;
        rz += (double)zz * r[o];
    }
    const double t = block_reduce_d(rz, shd);
    if (threadIdx.x == 0) rz_out[a] = t;
    __syncthreads();  // zc is reused by the caller's next aggregate
}

__device__ __forceinline__ float cg_beta(const SolverDev& S, const double* rz_cur, const double* rz_prev, bool first,
                                         double* shd) {
    if (first) return 0.0f;
    const double den = block_sum_partials(rz_prev, S.n_agg, shd);
    const double num = block_sum_partials(rz_cur, S.n_agg, shd);
    return den != 0.0 ? (float)(num / den) : 0.0f;
}

__device__ __forceinline__ float cg_alpha(const SolverDev& S, const double* pq_part, int n_pq, const double* rz_cur,
                                          double* shd) {
    const double pq = block_sum_partials(pq_part, n_pq, shd);
    const double rz = block_sum_partials(rz_cur, S.n_agg, shd);
    return pq != 0.0 ? (float)(rz / pq) : 0.0f;
}

// ------------------------------------------------------ assembled operator
// Structured Jacobian row vector of free node `node` in one residual group
// (rows r = 0..2: a_c at column 3c + r, b at column 9 + r), as u = (a0, a1, a2, b).
__device__ __forceinline__ void pair_term_u(const SolverDev& S, int4 tm, int node_i, int node_j, float (&ui)[4],
                                            float (&uj)[4]) {
#pragma unroll
    for (int c = 0; c < 4; c++) ui[c] = uj[c] = 0.0f;
    if (tm.x == 0 || tm.x == 1) {
        // edge (src, t) with neighbour dst: src gets (sw s d, sw), d = g_dst - g_src; dst gets (0, -sw s)
        const int src = tm.y, t = tm.z;
        const int dst = (int)S.nbr[(int64_t)src * S.K + t];
        const float sc = S.edge_scale[(int64_t)src * S.K + t];
        const float* gs = S.g + 3 * src;
        const float* gd = S.g + 3 * dst;
        const float d[3] = {gd[0] - gs[0], gd[1] - gs[1], gd[2] - gs[2]};
        float us[4], ud[4];
#pragma unroll
        for (int c = 0; c < 3; c++) us[c] = S.sw_reg * sc * d[c];
        us[3] = S.sw_reg;
        ud[0] = ud[1] = ud[2] = 0.0f;
        ud[3] = -S.sw_reg * sc;
        const bool i_is_src = tm.x == 0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            ui[c] = i_is_src ? us[c] : ud[c];
            uj[c] = i_is_src ? ud[c] : us[c];
        }
        return;
    }
    // constraint group: every (member, slot) whose source is the node
    const int cg = tm.y;
    for (int q = S.con_ptr[cg]; q < S.con_ptr[cg + 1]; q++) {
        const int m = S.con_mem[q];
        const float* gm = S.g + 3 * m;
        for (int sl = 0; sl < S.K; sl++) {
            const int e = m * S.K + sl;
            const int nj = (int)S.src_idx[e];
            if (nj != node_i && nj != node_j) continue;
            const float wv = S.src_w[e];
            const float* gj = S.g + 3 * nj;
            const float dd[3] = {gm[0] - gj[0], gm[1] - gj[1], gm[2] - gj[2]};
            float* u = nj == node_i ? ui : uj;
#pragma unroll
            for (int c = 0; c < 3; c++) u[c] += S.sw_con * wv * dd[c];
            u[3] += S.sw_con * wv;
        }
    }
}

// M_ij = sum over the pair's terms of u_i u_j^T (row-major 4x4), one thread per
// free owner (its pairs in order). Once per frame: the node positions change.
__global__ void pair_coef_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    const SolverDev& S = E.S;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (!E.cg_asm || S.n_pairs == 0 || f >= S.nF) return;
    const int node_i = S.free_nodes[f];
    for (int p = S.pair_ptr[f]; p < S.pair_ptr[f + 1]; p++) {
        const int node_j = S.free_nodes[S.pair_j[p]];
        float m[16];
#pragma unroll
        for (int c = 0; c < 16; c++) m[c] = 0.0f;
        for (int t = S.term_ptr[p]; t < S.term_ptr[p + 1]; t++) {
            float ui[4], uj[4];
            pair_term_u(S, S.terms[t], node_i, node_j, ui, uj);
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) m[4 * a + b] += ui[a] * uj[b];
        }
        float4* o = reinterpret_cast<float4*>(E.pair_m + 16 * (int64_t)p);
#pragma unroll
        for (int a = 0; a < 4; a++) o[a] = make_float4(m[4 * a], m[4 * a + 1], m[4 * a + 2], m[4 * a + 3]);
    }
}

// The CG operator, assembled: s_i = D_i v_i + sum_pairs (M_ij (x) I_3) v_j, with
// the apply's p, q update and p.q partial. 16 lanes per owner: lane l takes
// pairs l, l + 16, ... and lane c < 12 row c of D_i v_i.
template <int BS>
__device__ __forceinline__ void apply_asm_body(const SolverEntry& E, const float* __restrict__ z, float beta,
                                               bool first, float* __restrict__ p, float* __restrict__ q,
                                               double* __restrict__ partial, int ob, double* shd) {
    const SolverDev& S = E.S;
    const int grp = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    const int fi = ob * (BS / LANES) + grp;
    float acc[12];
#pragma unroll
    for (int a = 0; a < 12; a++) acc[a] = 0.0f;
    float mine = 0.0f;
    if (fi < S.nF) {
        for (int pp = S.pair_ptr[fi] + lane; pp < S.pair_ptr[fi + 1]; pp += LANES) {
            const float4* m4 = reinterpret_cast<const float4*>(E.pair_m + 16 * (int64_t)pp);
            const float4* v4 = reinterpret_cast<const float4*>(z + 12 * (int64_t)S.pair_j[pp]);
            const float4 ma = m4[0], mb = m4[1], mc = m4[2], md = m4[3];
            const float4 va = v4[0], vb = v4[1], vc = v4[2];
            const float v[12] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w, vc.x, vc.y, vc.z, vc.w};
            const float M[4][4] = {{ma.x, ma.y, ma.z, ma.w}, {mb.x, mb.y, mb.z, mb.w}, {mc.x, mc.y, mc.z, mc.w},
                                   {md.x, md.y, md.z, md.w}};
            // component (c, r): R column c row r at 3c + r, t row r at 9 + r
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    float sacc = 0.0f;
#pragma unroll
                    for (int c2 = 0; c2 < 4; c2++) sacc += M[c][c2] * v[(c2 < 3 ? 3 * c2 : 9) + r];
                    acc[(c < 3 ? 3 * c : 9) + r] += sacc;
                }
        }
        if (lane < 12) {
            const float4* d4 = reinterpret_cast<const float4*>(E.dmat + 144 * (int64_t)fi + 12 * lane);
            const float4* v4 = reinterpret_cast<const float4*>(z + 12 * (int64_t)fi);
#pragma unroll
            for (int h = 0; h < 3; h++) {
                const float4 dv = d4[h], vv = v4[h];
                mine += dv.x * vv.x + dv.y * vv.y + dv.z * vv.z + dv.w * vv.w;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 12; a++) {
        const float t = group_reduce16(acc[a]);
        if (lane == a) mine += t;
    }
    double pq = 0.0;
    if (fi < S.nF && lane < 12) {
        const int64_t o = 12 * (int64_t)fi + lane;
        const float pn = first ? z[o] : z[o] + beta * p[o];
        const float qn = first ? mine : mine + beta * q[o];
        p[o] = pn;
        q[o] = qn;
        pq = (double)pn * qn;
    }
    const double t = block_reduce_d(pq, shd);
    if (threadIdx.x == 0) partial[ob] = t;
}

template <int BS>
__global__ void __launch_bounds__(BS) apply_asm_kernel(const SolverEntry* __restrict__ B, int k, int first,
                                                       int warm_iter, const float* __restrict__ v_ext) {
    const SolverEntry E = load_entry(B);
    if ((int)blockIdx.x >= E.nb_apply || *E.cg_done) return;
    const SolverDev& S = E.S;
    const double* rz_cur = E.rzp + (k & 1) * S.n_agg;
    const double* rz_prev = E.rzp + ((k + 1) & 1) * S.n_agg;
    const float* z = v_ext ? v_ext : (warm_iter >= 0 ? E.hwarm + (int64_t)warm_iter * E.n : E.z);
    __shared__ double shd[8];
    const float beta = cg_beta(S, rz_cur, rz_prev, first != 0, shd);
    apply_asm_body<BS>(E, z, beta, first != 0, E.p, E.q, E.partial, blockIdx.x, shd);
}

// CG iteration k (r.z partials of iterations k and k-1 in the two halves of
// rzp). The vector is z, or (warm_iter >= 0) the warm-start step of that GN
// iteration, or v_ext (single-scene test calls).
// BS: workgroup size (the CG block size: BS / 16 owners per workgroup, E.nb_apply
// workgroups, one p.q partial each).
template <int KT, int BS = 256>
__global__ void __launch_bounds__(BS) apply_kernel(const SolverEntry* __restrict__ B, int k, int first, int warm_iter,
                                                   const float* __restrict__ v_ext) {
    const SolverEntry E = load_entry(B);
    if ((int)blockIdx.x >= E.nb_apply || *E.cg_done) return;
    const SolverDev& S = E.S;
    const double* rz_cur = E.rzp + (k & 1) * S.n_agg;
    const double* rz_prev = E.rzp + ((k + 1) & 1) * S.n_agg;
    const float* z = v_ext ? v_ext : (warm_iter >= 0 ? E.hwarm + (int64_t)warm_iter * E.n : E.z);
    __shared__ double shd[8];
    const float beta = cg_beta(S, rz_cur, rz_prev, first != 0, shd);
    apply_body<KT, BS>(S, E.x, z, beta, first != 0, E.p, E.q, E.partial, blockIdx.x, shd);
}

// One workgroup per aggregate: (mode 1) h += alpha p, r -= alpha q; then
// z = Binv r and the coarse restriction rc_a = P_a^T r; partial |r|^2.
__global__ void __launch_bounds__(256) update_restrict_kernel(const SolverEntry* __restrict__ B, int mode, int k) {
    const SolverEntry E = load_entry(B);
    const SolverDev& S = E.S;
    if ((int)blockIdx.x >= S.n_agg || *E.cg_done) return;
    __shared__ float sred[256 / LANES][12];
    __shared__ double shd[8];
    const double* rz_cur = E.rzp + (k & 1) * S.n_agg;
    const float alpha = mode == 1 ? cg_alpha(S, E.partial, E.nb_apply, rz_cur, shd) : 0.0f;
    update_body(S, mode, alpha, E.hsol, E.r, mode == 1 ? E.p : nullptr, mode == 1 ? E.q : nullptr, E.binv, E.z, E.rc,
                E.partial2, blockIdx.x, sred, shd);
}

// One workgroup per aggregate: zc_a = (Ac^-1 rc)_a, z += P_a zc_a; partial r.z.
// Workgroup 0 also runs the convergence test on the |r|^2 partials.
// test == 0 (the preconditioned start): rz_0 into the first half of rzp, no
// convergence test; test == 1: CG iteration k, rz_{k+1} into half (k+1) & 1.
__global__ void __launch_bounds__(256) coarse_prolong_kernel(const SolverEntry* __restrict__ B, int k, int test) {
    const SolverEntry E = load_entry(B);
    const SolverDev& S = E.S;
    if ((int)blockIdx.x >= S.n_agg || *E.cg_done) return;
    __shared__ float zc[12];
    __shared__ float zred[4][12];
    __shared__ double shd[8];
    const int a = blockIdx.x;
    double* rz_out = E.rzp + (test ? ((k + 1) & 1) : 0) * S.n_agg;
    prolong_body(S, E.acinv, E.rc, E.r, E.z, rz_out, a, zc, zred, shd);
    const double rr = (a == 0 && test) ? block_sum_partials(E.partial2, S.n_agg, shd) : 0.0;
    if (threadIdx.x == 0 && a == 0 && test) {
        CgScalars* cs = E.cg;
        cs->rr = rr;
        cs->iters += 1;
        if (rr <= cs->rr_stop || cs->iters >= cs->max_iters) *E.cg_done = 1;
    }
}

// ------------------------------------------------ two kernels per iteration
// The same PCG with the coarse restriction carried by a recurrence, so that
// the update and the coarse phases become one kernel. With q_k = A p_k and
// r_{k+1} = r_k - alpha_k q_k, the restricted residual obeys
//   rc_{k+1} = P^T r_{k+1} = rc_k - alpha_k P^T q_k,
// and P^T q_k is formed where q_k is: in the apply, whose workgroups own one
// aggregate each. The coarse product of iteration k then needs only rc_k,
// P^T q_k (both complete after the apply) and alpha_k, so it runs in the
// same launch as the node-local update:
//   apply   [convergence test on |r_k|^2] beta_k; s = A z; p, q; partial p.q
//           per aggregate; (P^T q)_a                    (workgroup a, aggregate a)
//   update  alpha; rc_{k+1} = rc_k - alpha P^T q (every workgroup forms the
//           whole vector for its coarse rows, workgroup a stores its own part);
//           zc_a = (Ac^-1 rc_{k+1})_a; h += alpha p; r -= alpha q; partial |r|^2;
//           z = Binv r + P_a zc_a; partial r.z
// rc ping-pongs between two halves of the rc buffer (iteration k reads half
// k & 1) and P^T q sits in its third part. The convergence test of |r_k|^2
// opens apply k (every workgroup sums the partials itself and takes the same
// decision; workgroup 0 records it) and is repeated by the host-check flag
// kernel, with the same reduction, after a chunk of iterations.

// Sums of NS arrays of n FP64 partials in an order that does not depend on the
// workgroup size: the first wave adds lane-strided runs, then a shuffle tree.
// Every thread gets the results (through sh, NS doubles).
template <int NS>
__device__ __forceinline__ void wave0_sums(const double* const (&p)[NS], int n, double* sh, double (&out)[NS]) {
    __syncthreads();
    if (threadIdx.x < 64) {
#pragma unroll
        for (int j = 0; j < NS; j++) {
            double v = 0.0;
            for (int i = threadIdx.x; i < n; i += 64) v += p[j][i];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (threadIdx.x == 0) sh[j] = v;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NS; j++) out[j] = sh[j];
    __syncthreads();
}

// The convergence test after k CG iterations from |r_k|^2; workgroup 0 (or the
// flag kernel) records it. Returns whether the CG is done.
__device__ __forceinline__ bool fused_cg_test(const SolverEntry& E, double rr, int k, bool record) {
    CgScalars* cs = E.cg;
    const bool stop = rr <= cs->rr_stop || k >= cs->max_iters;
    if (record && threadIdx.x == 0) {
        cs->rr = rr;
        cs->iters = k;
        if (stop) *E.cg_done = 1;
    }
    return stop;
}

// BS: the CG block size (BS / 16 member nodes of the aggregate per pass).
template <int KT, int BS>
__global__ void __launch_bounds__(BS) cg_apply_agg_kernel(const SolverEntry* __restrict__ B, int k) {
    const SolverEntry E = load_entry(B);
    const SolverDev& S = E.S;
    const int a = blockIdx.x;
    if (a >= S.n_agg || *E.cg_done) return;
    __shared__ double shd[BS / 64];
    __shared__ float sred[BS / LANES][12];
    const double* rz_cur = E.rzp + (k & 1) * S.n_agg;
    const double* rz_prev = E.rzp + ((k + 1) & 1) * S.n_agg;
    const bool first = k == 0;
    float beta = 0.0f;
    if (!first) {
        const double* const ps[3] = {E.partial2, rz_cur, rz_prev};
        double sums[3];
        wave0_sums<3>(ps, S.n_agg, shd, sums);
        if (fused_cg_test(E, sums[0], k, a == 0)) return;
        beta = sums[2] != 0.0 ? (float)(sums[1] / sums[2]) : 0.0f;
    }
    const float* __restrict__ x = E.x;
    const float* __restrict__ z = E.z;
    float* __restrict__ p = E.p;
    float* __restrict__ q = E.q;
    const int grp = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    const int beg = S.agg_ptr[a], end = S.agg_ptr[a + 1];
    const float* ca = S.agg_center + 3 * a;
    float rcl[12];
#pragma unroll
    for (int c = 0; c < 12; c++) rcl[c] = 0.0f;
    double pq = 0.0;
    for (int m = beg + grp; m < end; m += BS / LANES) {
        const int fi = S.agg_member[m];
        float acc[12];
#pragma unroll
        for (int c = 0; c < 12; c++) acc[c] = 0.0f;
        const int ni = owner_items(S, fi);
        for (int it = lane; it < ni; it += LANES) {
            owner_item<KT>(
                S, fi, it, x, z,
                [&](float (&Ji)[6][9], float (&Jv)[6]) {
#pragma unroll
                    for (int r = 0; r < 6; r++)
#pragma unroll
                        for (int c = 0; c < 9; c++) acc[c] += Ji[r][c] * Jv[r];
                },
                [&](float (&a3)[3], float bb, float (&Jv)[3]) {
#pragma unroll
                    for (int r = 0; r < 3; r++) {
#pragma unroll
                        for (int c = 0; c < 3; c++) acc[3 * c + r] += a3[c] * Jv[r];
                        acc[9 + r] += bb * Jv[r];
                    }
                });
        }
        float mine = 0.0f;
#pragma unroll
        for (int c = 0; c < 12; c++) {
            const float t = group_reduce16(acc[c]);
            if (lane == c) mine = t;
        }
        float qn = 0.0f;
        if (lane < 12) {
            const int64_t o = 12 * (int64_t)fi + lane;
            const float pn = first ? z[o] : z[o] + beta * p[o];
            qn = first ? mine : mine + beta * q[o];
            p[o] = pn;
            q[o] = qn;
            pq += (double)pn * qn;
        }
        // (P^T q)_a: the R part (9) passes through; q_t reaches b (identity)
        // and A(r, c') with weight (g_j - c_a)_c'
        float qv[12];
#pragma unroll
        for (int j = 0; j < 12; j++) qv[j] = __shfl(qn, j, LANES);
        if (lane == 0) {
            const int node = S.free_nodes[fi];
            const float dj[3] = {S.g[3 * node] - ca[0], S.g[3 * node + 1] - ca[1], S.g[3 * node + 2] - ca[2]};
#pragma unroll
            for (int c = 0; c < 9; c++) rcl[c] += qv[c];
#pragma unroll
            for (int r3 = 0; r3 < 3; r3++) {
#pragma unroll
                for (int cc = 0; cc < 3; cc++) rcl[3 * cc + r3] += dj[cc] * qv[9 + r3];
                rcl[9 + r3] += qv[9 + r3];
            }
        }
    }
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 12; c++) sred[grp][c] = rcl[c];
    __syncthreads();
    if (threadIdx.x < 12) {
        float t = 0.0f;
        for (int g2 = 0; g2 < BS / LANES; g2++) t += sred[g2][threadIdx.x];
        E.rc[2 * (int64_t)S.nc_pad + 12 * a + threadIdx.x] = t;
    }
    const double t = block_reduce_d(pq, shd);
    if (threadIdx.x == 0) E.partial[a] = t;
}

template <int BS>
__global__ void __launch_bounds__(BS) cg_update_coarse_kernel(const SolverEntry* __restrict__ B, int k) {
    const SolverEntry E = load_entry(B);
    const SolverDev& S = E.S;
    const int a = blockIdx.x;
    if (a >= S.n_agg || *E.cg_done) return;
    __shared__ double shd[BS / 64];
    __shared__ float zc[12];
    __shared__ float zred[BS / 64][12];
    const double* rz_cur = E.rzp + (k & 1) * S.n_agg;
    double* rz_next = E.rzp + ((k + 1) & 1) * S.n_agg;
    float alpha;
    {
        const double* const ps[2] = {rz_cur, E.partial};
        double sums[2];
        wave0_sums<2>(ps, S.n_agg, shd, sums);
        alpha = sums[1] != 0.0 ? (float)(sums[0] / sums[1]) : 0.0f;
    }
    const int nc = 12 * S.n_agg;
    const int ld = S.nc_pad;
    const float* rc_cur = E.rc + (k & 1) * (int64_t)ld;
    float* rc_next = E.rc + ((k + 1) & 1) * (int64_t)ld;
    const float* ptq = E.rc + 2 * (int64_t)ld;
    // zc_a = 12 rows of Ac^-1 times rc_{k+1}, formed on the fly
    {
        float part[12];
#pragma unroll
        for (int row = 0; row < 12; row++) part[row] = 0.0f;
        const float* A0 = E.acinv + (int64_t)(12 * a) * ld;
#pragma unroll 4
        for (int j = threadIdx.x; j < nc; j += BS) {
            const float rj = rc_cur[j] - alpha * ptq[j];
#pragma unroll
            for (int row = 0; row < 12; row++) part[row] += A0[(int64_t)row * ld + j] * rj;
        }
        const int wl = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int row = 0; row < 12; row++) {
            float v = part[row];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (wl == 0) zred[wv][row] = v;
        }
        if (threadIdx.x < 12) {
            const int j = 12 * a + threadIdx.x;
            rc_next[j] = rc_cur[j] - alpha * ptq[j];
        }
        __syncthreads();
        if (threadIdx.x < 12) {
            float t = 0.0f;
            for (int w = 0; w < BS / 64; w++) t += zred[w][threadIdx.x];
            zc[threadIdx.x] = t;
        }
        __syncthreads();
    }
    float* __restrict__ hsol = E.hsol;
    float* __restrict__ r = E.r;
    float* __restrict__ z = E.z;
    const float* __restrict__ p = E.p;
    const float* __restrict__ q = E.q;
    const float* __restrict__ binv = E.binv;
    const float* ca = S.agg_center + 3 * a;
    const int grp = threadIdx.x / LANES, c = threadIdx.x % LANES;
    double rr = 0.0, rz = 0.0;
    const int beg = S.agg_ptr[a], end = S.agg_ptr[a + 1];
    for (int m = beg + grp; m < end; m += BS / LANES) {
        const int fi = S.agg_member[m];
        float rv = 0.0f;
        if (c < 12) {
            const int64_t o = 12 * (int64_t)fi + c;
            hsol[o] += alpha * p[o];
            r[o] -= alpha * q[o];
            rv = r[o];
            rr += (double)rv * rv;
        }
        float rn[12];
#pragma unroll
        for (int j = 0; j < 12; j++) rn[j] = __shfl(rv, j, LANES);
        if (c < 12) {
            float zz = 0.0f;
#pragma unroll
            for (int j = 0; j < 12; j++) zz += binv[144 * (int64_t)fi + c * 12 + j] * rn[j];
            float add;
            if (c < 9) {
                add = zc[c];
            } else {
                const int node = S.free_nodes[fi];
                const int r3 = c - 9;
                const float dj[3] = {S.g[3 * node] - ca[0], S.g[3 * node + 1] - ca[1], S.g[3 * node + 2] - ca[2]};
                add = zc[9 + r3] + zc[r3] * dj[0] + zc[3 + r3] * dj[1] + zc[6 + r3] * dj[2];
            }
            const float zn = zz + add;
            z[12 * (int64_t)fi + c] = zn;
            rz += (double)zn * rv;
        }
    }
    const double trr = block_reduce_d(rr, shd);
    if (threadIdx.x == 0) E.partial2[a] = trr;
    const double trz = block_reduce_d(rz, shd);
    if (threadIdx.x == 0) rz_next[a] = trz;
}

// flag = 1 when every scene's CG is done after k iterations of the two-kernel
// form: the convergence test of apply k, run here first so that the host sees
// it at the end of a chunk (apply k repeats it with the same result).
// iters (when not null): the largest CG iteration count over the scenes.
__global__ void fused_flag_kernel(const SolverEntry* __restrict__ B, int n, int k, int* __restrict__ flag,
                                  int* __restrict__ iters, int* __restrict__ gn) {
    __shared__ double shd[4];
    __shared__ int s_all, s_iters;
    if (threadIdx.x == 0) {
        s_all = 1;
        s_iters = 0;
    }
    for (int s = 0; s < n; s++) {
        const SolverEntry E = load_entry(B + s);  // launched with gridDim.z == 1
        if (*E.cg_done) {
            if (threadIdx.x == 0) s_iters = max(s_iters, E.cg->iters);
            continue;
        }
        const double* const ps[1] = {E.partial2};
        double sums[1];
        wave0_sums<1>(ps, E.S.n_agg, shd, sums);
        const bool stop = fused_cg_test(E, sums[0], k, true);
        if (threadIdx.x == 0) {
            if (!stop) s_all = 0;
            s_iters = max(s_iters, k);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (iters) *iters = s_iters;
        if (gn) {
            int gn_all = 1, gn_most = 0;
            for (int s = 0; s < n; s++) {
                gn_all &= *B[s].gn_done != 0 ? 1 : 0;
                gn_most = max(gn_most, B[s].gs->iters_done);
            }
            gn[0] = gn_all;
            gn[1] = gn_most;
        }
        __hip_atomic_store(flag, s_all, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ------------------------------------------------------------ grid barrier
// Generation barrier for a cooperative launch (all workgroups co-resident).
// Every thread first fences its own global writes at device scope (a fence
// only waits for the issuing wave's stores, so one per wave is needed), then
// thread 0 arrives on the monotone counter and waits for the generation to
// reach its epoch + 1; after the workgroup barrier every thread fences again
// (acquire) before reading other workgroups' data. The wait is bounded: after
// GRID_SPIN_LIMIT polls, or when another workgroup has raised `err`, the
// workgroup raises `err` and leaves, so the kernel always drains.
constexpr unsigned GRID_SPIN_LIMIT = 1u << 24;

__device__ __forceinline__ bool grid_sync(GridBarrier* gb, unsigned& epoch) {
    __shared__ int s_ok;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned target = epoch + 1;
        const unsigned arrived = __hip_atomic_fetch_add(&gb->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == target * gridDim.x - 1) {
            __hip_atomic_store(&gb->gen, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned spins = 0;
            while (__hip_atomic_load(&gb->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins >= GRID_SPIN_LIMIT ||
                    __hip_atomic_load(&gb->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    __hip_atomic_store(&gb->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        s_ok = ok;
    }
    __syncthreads();
    __threadfence();
    epoch += 1;
    return s_ok != 0;
}

// The whole PCG solve of one Gauss-Newton iteration in one cooperative launch.
// Preconditions (queued before it): r = b, z0 = M^-1 r0 and rz_0 in rzp[0..n_agg),
// cs->bb, cs->iters = 0; gb zeroed. Owner blocks and aggregates are strided
// over the grid, so any grid size up to the co-resident limit works.
template <int KT>
__global__ void __launch_bounds__(256) pcg_persistent_kernel(SolverDev S, const float* __restrict__ x,
                                                             float* __restrict__ hsol, float* __restrict__ r,
                                                             float* __restrict__ z, float* __restrict__ p,
                                                             float* __restrict__ q, const float* __restrict__ binv,
                                                             float* __restrict__ rc, const float* __restrict__ acinv,
                                                             double* __restrict__ pq_part, double* __restrict__ rr_part,
                                                             double* __restrict__ rzp, CgScalars* __restrict__ cs,
                                                             GridBarrier* __restrict__ gb, int* __restrict__ skip) {
    if (*skip) return;  // read by every workgroup before anything is written
    __shared__ float sred[256 / LANES][12];
    __shared__ float zc[12];
    __shared__ float zred[4][12];
    __shared__ double shd[8];
    const int n_ob = (S.nF + (256 / LANES) - 1) / (256 / LANES);
    const int max_iters = cs->max_iters;
    const double tol2bb = cs->rr_stop;
    unsigned epoch = 0;
    int k = 0;
    double rr = 0.0;
    for (;; k++) {
        const double* rz_cur = rzp + (k & 1) * S.n_agg;
        double* rz_next = rzp + ((k + 1) & 1) * S.n_agg;
        const float beta = cg_beta(S, rz_cur, rz_next, k == 0, shd);
        for (int ob = blockIdx.x; ob < n_ob; ob += gridDim.x)
            apply_body<KT>(S, x, z, beta, k == 0, p, q, pq_part, ob, shd);
        if (!grid_sync(gb, epoch)) return;
        const float alpha = cg_alpha(S, pq_part, n_ob, rz_cur, shd);
        for (int a = blockIdx.x; a < S.n_agg; a += gridDim.x)
            update_body(S, 1, alpha, hsol, r, p, q, binv, z, rc, rr_part, a, sred, shd);
        if (!grid_sync(gb, epoch)) return;
        rr = block_sum_partials(rr_part, S.n_agg, shd);
        if (rr <= tol2bb || k + 1 >= max_iters) break;
        for (int a = blockIdx.x; a < S.n_agg; a += gridDim.x)
            prolong_body(S, acinv, rc, r, z, rz_next, a, zc, zred, shd);
        if (!grid_sync(gb, epoch)) return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cs->rr = rr;
        cs->iters = k + 1;
        *skip = 1;
    }
}

__global__ void grid_barrier_reset_kernel(GridBarrier* gb) {
    gb->count = 0u;
    gb->gen = 0u;
    gb->err = 0;
    gb->pad = 0;
}

// |b|^2 from the rhs partials and the CG stopping level: relative residual
// tol, but never below CG_TOL_FLOOR times the first Gauss-Newton gradient of
// this solve (inexact Newton: once the gradient itself is below that level the
// step is negligible, the solve returns at once and the stop rule fires). An
// all-zero right-hand side ends the solve.
constexpr double CG_TOL_FLOOR = 1e-6;

// Also opens the Gauss-Newton iteration's CG: cg_done starts as gn_done (a
// stopped loop runs no CG), with the entry's cap and tolerance.
__global__ void cg_init_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    if (threadIdx.x != 0) return;
    int* __restrict__ skip = E.cg_done;
    CgScalars* __restrict__ cs = E.cg;
    *skip = *E.gn_done;
    cs->max_iters = E.cg_max_iters;
    cs->tol2 = E.tol2;
    if (*skip) return;
    GnScalars* __restrict__ gs = E.gs;
    const double bb = sum_partials(E.partial, E.nb_owner);
    if (gs->iters_done == 0) gs->bb0 = bb;
    cs->bb = bb;
    cs->rr = bb;
    cs->iters = 0;
    cs->rr_stop = fmax(cs->tol2 * bb, CG_TOL_FLOOR * CG_TOL_FLOOR * gs->bb0);
    if (!(bb > 0.0) || bb <= cs->rr_stop * (1.0 - 1e-12)) *skip = 1;
}

__global__ void zero_kernel(float* __restrict__ v, int64_t n, const int* __restrict__ skip) {
    if (skip && *skip) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = 0.0f;
}

// h = 0 for a cold start (also when the CG is skipped for a negligible gradient:
// the line search then takes a zero step).
__global__ void zero_step_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    if (*E.gn_done) return;
    float* __restrict__ v = E.hsol;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E.n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = 0.0f;
}

// Warm-started CG: h = h0 and r = b - A h0 (q holds A h0); a CG skipped for a
// negligible gradient keeps h = 0 like the cold start.
__global__ void warm_init_kernel(const SolverEntry* __restrict__ B, int iter) {
    const SolverEntry E = load_entry(B);
    if (*E.gn_done) return;
    const int64_t n = E.n;
    float* __restrict__ hsol = E.hsol;
    const float* __restrict__ h0 = E.hwarm + (int64_t)iter * n;
    float* __restrict__ r = E.r;
    const float* __restrict__ q = E.q;
    const bool skip = *E.cg_done != 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (skip) {
            hsol[i] = 0.0f;
        } else {
            hsol[i] = h0[i];
            r[i] = r[i] - q[i];
        }
    }
}

__global__ void copy_kernel(float* __restrict__ dst, const float* __restrict__ src, int64_t n,
                            const int* __restrict__ skip) {
    if (skip && *skip) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// The step of GN iteration iter becomes next solve's warm start for it.
__global__ void save_warm_kernel(const SolverEntry* __restrict__ B, int iter) {
    const SolverEntry E = load_entry(B);
    if (*E.gn_done) return;
    float* __restrict__ dst = E.hwarm + (int64_t)iter * E.n;
    const float* __restrict__ src = E.hsol;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E.n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ---------------------------------------------------------------- coarse space
__global__ void agg_center_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    const SolverDev& S = E.S;
    const int a = blockIdx.x;
    if (a >= S.n_agg) return;
    __shared__ float sx[256], sy[256], sz[256];
    float x = 0.0f, y = 0.0f, z = 0.0f;
    const int beg = S.agg_ptr[a], end = S.agg_ptr[a + 1];
    for (int m = beg + threadIdx.x; m < end; m += blockDim.x) {
        const int node = S.free_nodes[S.agg_member[m]];
        x += S.g[3 * node]; y += S.g[3 * node + 1]; z += S.g[3 * node + 2];
    }
    sx[threadIdx.x] = x; sy[threadIdx.x] = y; sz[threadIdx.x] = z;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sx[threadIdx.x] += sx[threadIdx.x + s];
            sy[threadIdx.x] += sy[threadIdx.x + s];
            sz[threadIdx.x] += sz[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float n = (float)(end - beg > 0 ? end - beg : 1);
        S.agg_center[3 * a] = sx[0] / n; S.agg_center[3 * a + 1] = sy[0] / n; S.agg_center[3 * a + 2] = sz[0] / n;
    }
}

// Galerkin coarse matrix Ac = P^T (J^T J) P accumulated per residual group:
// each group's J projected onto the coarse dofs of the aggregates it touches.
// Coarse Galerkin entries are accumulated as 64-bit fixed point (2^-24 units):
// integer additions commute, so the assembled operator does not depend on the
// order in which the residual groups arrive. Range +-2^39 (5.5e11), resolution
// 6e-8 against diagonal entries >= w_reg K; it is only a preconditioner.
constexpr double COARSE_FX = 16777216.0;

__device__ __forceinline__ void coarse_add(unsigned long long* Ac, int64_t at, float s) {
    atomicAdd(Ac + at, (unsigned long long)__double2ll_rn((double)s * COARSE_FX));
}

__global__ void zero_coarse_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    const int64_t n = 2 * (int64_t)E.S.nc_pad * E.S.nc_pad;  // the 64-bit entries as floats
    float* __restrict__ v = reinterpret_cast<float*>(E.ac64);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = 0.0f;
}

__global__ void coarse_to_float_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    const unsigned long long* __restrict__ Ac64 = E.ac64;
    float* __restrict__ Ac = E.acinv;
    const int64_t n = (int64_t)E.S.nc_pad * E.S.nc_pad;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        Ac[i] = (float)((double)(long long)Ac64[i] / COARSE_FX);
}

// Coarse contribution of the constraint groups with several members (centre
// mode, one per control block). Every (member, slot) entry reaching a free node
// of aggregate a adds to the aggregate's projected row coefficients
//   R part: sw_con w (v_m - g_j) + sw_con w (g_j - c_a),   t part: sw_con w,
// the same per-entry values as the one-member path, summed by one thread in
// entry order (deterministic); then every pair of touched aggregates adds its
// 3-row outer product. One workgroup per group; LDS: 24 bytes per aggregate.
__global__ void __launch_bounds__(256) coarse_multi_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    const SolverDev& S = E.S;
    if ((int)blockIdx.x >= S.n_multi) return;
    unsigned long long* __restrict__ Ac = E.ac64;
    extern __shared__ float lds[];
    float* coef = lds;                                   // 4 per aggregate
    int* list = reinterpret_cast<int*>(coef + 4 * S.n_agg);  // touched aggregates
    int* mark = list + S.n_agg;
    __shared__ int nt;
    for (int i = threadIdx.x; i < 4 * S.n_agg; i += blockDim.x) coef[i] = 0.0f;
    for (int i = threadIdx.x; i < S.n_agg; i += blockDim.x) mark[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int cid = S.multi_con[blockIdx.x];
        int cnt = 0;
        for (int m = S.con_ptr[cid]; m < S.con_ptr[cid + 1]; m++) {
            const int c = S.con_mem[m];
            const float v[3] = {S.g[3 * c], S.g[3 * c + 1], S.g[3 * c + 2]};
            for (int j = 0; j < S.K; j++) {
                const int nj = (int)S.src_idx[(int64_t)c * S.K + j];
                const int fj = S.free_id[nj];
                if (fj < 0) continue;
                const float w = S.src_w[(int64_t)c * S.K + j];
                const int a = S.agg_of[fj];
                const float jt = S.sw_con * w;
                for (int cc = 0; cc < 3; cc++) {
                    const float dd = v[cc] - S.g[3 * nj + cc];
                    const float dj = S.g[3 * nj + cc] - S.agg_center[3 * a + cc];
                    coef[4 * a + cc] += S.sw_con * w * dd + jt * dj;
                }
                coef[4 * a + 3] += jt;
                if (!mark[a]) {
                    mark[a] = 1;
                    list[cnt++] = a;
                }
            }
        }
        nt = cnt;
    }
    __syncthreads();
    const int nc = S.nc_pad;
    const int n = nt;
    for (int p = threadIdx.x; p < n * n; p += blockDim.x) {
        const int a1 = list[p / n], a2 = list[p % n];
        for (int r = 0; r < 3; r++)
            for (int p1 = 0; p1 < 4; p1++)
                for (int p2 = 0; p2 < 4; p2++) {
                    const float val = coef[4 * a1 + p1] * coef[4 * a2 + p2];
                    if (val == 0.0f) continue;
                    const int u = p1 < 3 ? 3 * p1 + r : 9 + r;
                    const int w = p2 < 3 ? 3 * p2 + r : 9 + r;
                    coarse_add(Ac, (int64_t)(12 * a1 + u) * nc + 12 * a2 + w, val);
                }
    }
}

// Every 3-row group's Jacobian block for a node is I_3-structured: row r holds
// (a_0, a_1, a_2) at the R columns 3c + r and b at the t column 9 + r. Its
// projection onto the node's aggregate keeps the structure with
// a_c + b (g_j - c_a)_c, so a group contributes, per pair of touched
// aggregates, the 16 products of their 4-coefficient rows placed in the three
// rows. One thread per group; the (aggregate, coefficients) list of a group
// lives in LDS and nodes of one aggregate are summed in the group's node order.
__global__ void __launch_bounds__(128) coarse_assemble_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    const SolverDev& S = E.S;
    const int64_t nb = (S.n_groups + 127) / 128;
    if ((int64_t)blockIdx.x >= nb) return;
    const float* __restrict__ x = E.x;
    unsigned long long* __restrict__ Ac = E.ac64;
    __shared__ float cv[128][KNN_MAX_T + 1][4];
    __shared__ int ca[128][KNN_MAX_T + 1];
    const int tid = threadIdx.x;
    const int nc = S.nc_pad;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < S.n_groups;
         g += nb * blockDim.x) {
        int64_t gg = g;
        float Jrot[6][9];
        if (gg < S.nF) {
            const int fi = (int)gg;
            const float* xi = x + 12 * (int64_t)fi;
            for (int a = 0; a < 6; a++)
                for (int b = 0; b < 9; b++) Jrot[a][b] = 0.0f;
            const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
            for (int p = 0; p < 3; p++)
                for (int c = 0; c < 3; c++) {
                    Jrot[p][3 * pa[p] + c] = S.sw_rot * xi[3 * pb[p] + c];
                    Jrot[p][3 * pb[p] + c] += S.sw_rot * xi[3 * pa[p] + c];
                }
            for (int j = 0; j < 3; j++)
                for (int c = 0; c < 3; c++) Jrot[3 + j][3 * j + c] = 2.0f * S.sw_rot * xi[3 * j + c];
            // R part maps identically to the aggregate's A: rows x 9 block
            const int a = S.agg_of[fi];
            for (int u = 0; u < 9; u++)
                for (int w = 0; w < 9; w++) {
                    float s = 0.0f;
                    for (int rr = 0; rr < 6; rr++) s += Jrot[rr][u] * Jrot[rr][w];
                    if (s != 0.0f) coarse_add(Ac, (int64_t)(12 * a + u) * nc + 12 * a + w, s);
                }
            continue;
        }
        gg -= S.nF;
        int nn = 0;
        auto push = [&](int fi, float r0, float r1, float r2, float t) {
            const int a = S.agg_of[fi];
            const int node = S.free_nodes[fi];
            const float v[4] = {r0 + t * (S.g[3 * node] - S.agg_center[3 * a]),
                                r1 + t * (S.g[3 * node + 1] - S.agg_center[3 * a + 1]),
                                r2 + t * (S.g[3 * node + 2] - S.agg_center[3 * a + 2]), t};
            for (int q = 0; q < nn; q++)
                if (ca[tid][q] == a) {
                    for (int c = 0; c < 4; c++) cv[tid][q][c] += v[c];
                    return;
                }
            ca[tid][nn] = a;
            for (int c = 0; c < 4; c++) cv[tid][nn][c] = v[c];
            nn++;
        };
        if (gg < (int64_t)S.nF * S.K) {
            const int fi = (int)(gg / S.K), t = (int)(gg % S.K);
            const int i = S.free_nodes[fi];
            const int k = (int)S.nbr[(int64_t)i * S.K + t];
            const int fk = S.free_id[k];
            const float s = S.edge_scale[(int64_t)i * S.K + t];
            const float d[3] = {S.g[3 * k] - S.g[3 * i], S.g[3 * k + 1] - S.g[3 * i + 1], S.g[3 * k + 2] - S.g[3 * i + 2]};
            push(fi, S.sw_reg * s * d[0], S.sw_reg * s * d[1], S.sw_reg * s * d[2], S.sw_reg);
            if (fk >= 0) push(fk, 0.0f, 0.0f, 0.0f, -S.sw_reg * s);
        } else {
            gg -= (int64_t)S.nF * S.K;
            if (gg < (int64_t)S.nS * S.K) {
                const int si = (int)(gg / S.K), t = (int)(gg % S.K);
                const int i = S.static_nodes[si];
                const int k = (int)S.nbr[(int64_t)i * S.K + t];
                const int fk = S.free_id[k];
                if (fk < 0) continue;
                push(fk, 0.0f, 0.0f, 0.0f, -S.sw_reg);
            } else {
                gg -= (int64_t)S.nS * S.K;
                // groups of several members: coarse_multi_kernel
                if (S.con_ptr[gg + 1] - S.con_ptr[gg] != 1) continue;
                const int c = S.con_mem[S.con_ptr[gg]];
                const float v[3] = {S.g[3 * c], S.g[3 * c + 1], S.g[3 * c + 2]};
                for (int j = 0; j < S.K; j++) {
                    const int nj = (int)S.src_idx[(int64_t)c * S.K + j];
                    const int fj = S.free_id[nj];
                    if (fj < 0) continue;
                    const float w = S.src_w[(int64_t)c * S.K + j];
                    const float dd[3] = {v[0] - S.g[3 * nj], v[1] - S.g[3 * nj + 1], v[2] - S.g[3 * nj + 2]};
                    push(fj, S.sw_con * w * dd[0], S.sw_con * w * dd[1], S.sw_con * w * dd[2], S.sw_con * w);
                }
            }
        }
        for (int q = 0; q < nn; q++) {
            const int a1 = ca[tid][q];
            for (int q2 = 0; q2 < nn; q2++) {
                const int a2 = ca[tid][q2];
                for (int p1 = 0; p1 < 4; p1++) {
                    const float c1 = cv[tid][q][p1];
                    for (int p2 = 0; p2 < 4; p2++) {
                        const float val = c1 * cv[tid][q2][p2];
                        if (val == 0.0f) continue;
                        for (int r = 0; r < 3; r++) {
                            const int u = p1 < 3 ? 3 * p1 + r : 9 + r;
                            const int w = p2 < 3 ? 3 * p2 + r : 9 + r;
                            coarse_add(Ac, (int64_t)(12 * a1 + u) * nc + 12 * a2 + w, val);
                        }
                    }
                }
            }
        }
    }
}

// Blocked in-place Gauss-Jordan inverse of the SPD coarse matrix (no pivoting;
// n a multiple of 64, padding rows/cols hold the identity). Per 64-wide pivot
// block K with P = A_KK^-1: A_Kj <- P A_Kj, A_ij <- A_ij - A_iK (P A_Kj),
// A_iK <- -A_iK P, A_KK <- P. One launch per K: every tile reads the old
// column and row panels from copies made by the previous launch (so tiles are
// updated in place without races) and writes the copies of the next panels;
// the tile (K+1, K+1), dispatched first, inverts its new value into the next
// pivot buffer after its update.
constexpr int GJ_T = 64;

// Inverse of the 64x64 pivot tile by in-place Gauss-Jordan, four waves: lane j
// of wave w keeps rows 16w..16w+15 of column j in registers. Step k: the wave
// holding row k publishes it in LDS (one barrier; the row buffer alternates, so
// step k+2 overwrites a buffer only after every wave has passed step k+1's
// barrier), every wave reads its rows of the pivot column from lane k with
// v_readlane (the loops are unrolled, so every register index is static). Per
// entry the same operations in the same order as the textbook in-place form:
// row k scaled by 1/d, P_ij -= P_ik (P_kj / d), column k -> -P_ik / d.
constexpr int GJ_R = GJ_T / 4;  // rows per wave

template <int K>
__device__ __forceinline__ void gj_step4(float (&c)[GJ_R], int lane, int wave, float (*prow)[GJ_T]) {
    float* pr = prow[K & 1];
    if (wave == K / GJ_R) pr[lane] = c[K % GJ_R];
    __syncthreads();
    const float d = pr[K];
    const float iv = (fabsf(d) > 1e-30f) ? 1.0f / d : 0.0f;
    const float rk = pr[lane] * iv;
#pragma unroll
    for (int r = 0; r < GJ_R; r++) {
        const float colv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c[r]), K));
        const bool row_k = wave * GJ_R + r == K;
        if (lane == K) c[r] = row_k ? iv : colv * -iv;
        else c[r] = row_k ? rk : c[r] - colv * rk;
    }
}

template <int K>
__device__ __forceinline__ void gj_steps4(float (&c)[GJ_R], int lane, int wave, float (*prow)[GJ_T]) {
    if constexpr (K < GJ_T) {
        gj_step4<K>(c, lane, wave, prow);
        gj_steps4<K + 1>(c, lane, wave, prow);
    }
}

// The inverse of the 64x64 tile T (LDS) into dst (64x64 row-major); 256 threads.
__device__ __forceinline__ void gj_invert_tile(const float (*T)[GJ_T + 1], float* __restrict__ dst,
                                               float (*prow)[GJ_T]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float c[GJ_R];
#pragma unroll
    for (int r = 0; r < GJ_R; r++) c[r] = T[wave * GJ_R + r][lane];
    gj_steps4<0>(c, lane, wave, prow);
#pragma unroll
    for (int r = 0; r < GJ_R; r++) dst[(wave * GJ_R + r) * GJ_T + lane] = c[r];
}

// 64x64 tile (row stride ld) into LDS.
__device__ __forceinline__ void load_tile(float (*T)[GJ_T + 1], const float* __restrict__ src, int64_t ld) {
    for (int e = threadIdx.x; e < GJ_T * GJ_T / 4; e += blockDim.x) {
        const int r = e / (GJ_T / 4), c4 = e % (GJ_T / 4);
        const float4 v = *reinterpret_cast<const float4*>(src + (int64_t)r * ld + 4 * c4);
        T[r][4 * c4] = v.x; T[r][4 * c4 + 1] = v.y; T[r][4 * c4 + 2] = v.z; T[r][4 * c4 + 3] = v.w;
    }
}

// 64x64 = (64x64) (64x64) products from LDS tiles: thread t accumulates the 4x4
// block at rows 4 (t / 16), columns 4 (t % 16), with fused multiply-adds.
__device__ __forceinline__ void tile_mm(const float (*X)[GJ_T + 1], const float (*Y)[GJ_T + 1], float (&acc)[4][4]) {
    const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0f;
#pragma unroll 4
    for (int t = 0; t < GJ_T; t++) {
        float a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = X[4 * tr + i][t];
#pragma unroll
        for (int j = 0; j < 4; j++) b[j] = Y[t][4 * tc + j];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_fmaf(a[i], b[j], acc[i][j]);
    }
}

// Panel copies of the Gauss-Jordan sweep: gj_col holds four arrays of nt tiles
// (64x64, contiguous): column panels [0], [1] and row panels [2], [3], the
// pair K & 1 read by the launch of pivot block K. gj_piv holds two pivot
// inverses (K & 1).
__device__ __forceinline__ float* gj_panel(const SolverEntry& E, int which, int tile) {
    return E.gj_col + ((int64_t)which * E.S.nc_pad + (int64_t)tile * GJ_T) * GJ_T;
}

// Before pivot block 0: the panels of block 0 and the inverse of A_00.
// blockIdx.x = tile b: column-panel tile (b, 0) and row-panel tile (0, b).
__global__ void __launch_bounds__(256) gj_start_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    const int b = blockIdx.x;
    if (b >= E.nt) return;
    const float* __restrict__ A = E.acinv;
    const int n = E.S.nc_pad;
    float* __restrict__ col = gj_panel(E, 0, b);
    float* __restrict__ row = gj_panel(E, 2, b);
    for (int e = threadIdx.x; e < GJ_T * GJ_T; e += blockDim.x) {
        const int r = e / GJ_T, c = e % GJ_T;
        col[e] = A[(int64_t)(b * GJ_T + r) * n + c];
        row[e] = A[(int64_t)r * n + b * GJ_T + c];
    }
    if (b == 0) {
        __shared__ float T[GJ_T][GJ_T + 1];
        __shared__ float prow[2][GJ_T];
        load_tile(T, A, n);
        __syncthreads();
        gj_invert_tile(T, E.gj_piv, prow);
    }
}

// Pivot block k: one workgroup per tile (i, j), linear order with the tile
// (k+1, k+1) first. Products in the order of the three-phase form
// (P A_Kj first, then A_iK times that), so the inverse does not depend on how
// the phases are launched.
__global__ void __launch_bounds__(256) gj_step_kernel(const SolverEntry* __restrict__ B, int k) {
    const SolverEntry E = load_entry(B);
    const int nt = E.nt;
    if (k >= nt || (int)blockIdx.x >= nt * nt) return;
    const int kn = k + 1;
    int lin = blockIdx.x;
    if (kn < nt) {
        const int dn = kn * nt + kn;
        lin = lin == 0 ? dn : (lin == dn ? 0 : lin);
    }
    const int i = lin / nt, j = lin % nt;
    float* __restrict__ A = E.acinv;
    const int n = E.S.nc_pad;
    const float* __restrict__ piv = E.gj_piv + (k & 1) * GJ_T * GJ_T;
    __shared__ float Ps[GJ_T][GJ_T + 1];
    __shared__ float Rs[GJ_T][GJ_T + 1];
    __shared__ float Cs[GJ_T][GJ_T + 1];
    __shared__ float prow[2][GJ_T];
    const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
    float out[4][4];
    if (i == k && j == k) {
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) out[a][b] = piv[(4 * tr + a) * GJ_T + 4 * tc + b];
    } else if (i == k) {
        load_tile(Ps, piv, GJ_T);
        load_tile(Rs, gj_panel(E, 2 + (k & 1), j), GJ_T);
        __syncthreads();
        tile_mm(Ps, Rs, out);
    } else if (j == k) {
        load_tile(Cs, gj_panel(E, k & 1, i), GJ_T);
        load_tile(Ps, piv, GJ_T);
        __syncthreads();
        tile_mm(Cs, Ps, out);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) out[a][b] = -out[a][b];
    } else {
        load_tile(Ps, piv, GJ_T);
        load_tile(Rs, gj_panel(E, 2 + (k & 1), j), GJ_T);
        load_tile(Cs, gj_panel(E, k & 1, i), GJ_T);
        float old[4][4];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) old[a][b] = A[(int64_t)(i * GJ_T + 4 * tr + a) * n + j * GJ_T + 4 * tc + b];
        __syncthreads();
        float t[4][4];
        tile_mm(Ps, Rs, t);  // the new row-panel tile P A_Kj
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) Rs[4 * tr + a][4 * tc + b] = t[a][b];
        __syncthreads();
        tile_mm(Cs, Rs, out);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) out[a][b] = old[a][b] - out[a][b];
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) A[(int64_t)(i * GJ_T + 4 * tr + a) * n + j * GJ_T + 4 * tc + b] = out[a][b];
    if (kn >= nt || (i != kn && j != kn)) return;
    // panels of pivot block k+1
    if (j == kn) {
        float* __restrict__ col = gj_panel(E, kn & 1, i);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) col[(4 * tr + a) * GJ_T + 4 * tc + b] = out[a][b];
    }
    if (i == kn) {
        float* __restrict__ row = gj_panel(E, 2 + (kn & 1), j);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) row[(4 * tr + a) * GJ_T + 4 * tc + b] = out[a][b];
    }
    if (i == kn && j == kn) {
        __syncthreads();  // Cs is free again
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) Cs[4 * tr + a][4 * tc + b] = out[a][b];
        __syncthreads();
        gj_invert_tile(Cs, E.gj_piv + (kn & 1) * GJ_T * GJ_T, prow);
    }
}

__global__ void pad_identity_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    float* __restrict__ A = E.acinv;
    const int n = E.S.nc_pad, nc = 12 * E.S.n_agg;
    for (int i = nc + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) A[(int64_t)i * n + i] = 1.0f;
}

// ---------------------------------------------------------------- GN step
// Choose the first candidate with E <= E_old, update x, record energies and the
// stop rule |h_final| < (|x_old| + 1e-6) 1e-6 (Deform.cpp:142-166). Two
// kernels: every block finds the chosen candidate, accumulates its share of
// |x_old|^2 and |h|^2 in FP64 and moves its share of x; one block then sums
// the partials in block order and decides.
__device__ __forceinline__ int chosen_candidate(const double* cand_e, int ncand, const GnScalars* gs) {
    for (int c = 0; c < ncand; c++)
        if (!(cand_e[c] > gs->e_old)) return c;
    return -1;
}

__global__ void __launch_bounds__(256) accept_partial_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    if ((int)blockIdx.x >= E.acc_blocks || *E.gn_done) return;
    const int n = E.n, ncand = E.n_cand, nblocks = E.acc_blocks;
    float* __restrict__ x = E.x;
    const float* __restrict__ hsol = E.hsol;
    const double* __restrict__ cand_e = E.cand_e;
    const GnScalars* __restrict__ gs = E.gs;
    double* __restrict__ partial = E.acc_partial;
    __shared__ int chosen;
    __shared__ double sh[8];
    if (threadIdx.x == 0) chosen = chosen_candidate(cand_e, ncand, gs);
    __syncthreads();
    const float a = chosen >= 0 ? ldexpf(1.0f, -chosen) : 0.0f;
    double nx = 0.0, nh = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += nblocks * blockDim.x) {
        nx += (double)x[i] * x[i];
        nh += (double)hsol[i] * hsol[i];
        if (chosen >= 0) x[i] += a * hsol[i];
    }
    const double tx = block_reduce_d(nx, sh);
    const double th = block_reduce_d(nh, sh);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = tx;
        partial[2 * blockIdx.x + 1] = th;
    }
}

__global__ void __launch_bounds__(256) accept_final_kernel(const SolverEntry* __restrict__ B, int iter, int stop_rule) {
    const SolverEntry E = load_entry(B);
    int* __restrict__ gn_done = E.gn_done;
    if (*gn_done) return;
    const int nblocks = E.acc_blocks, ncand = E.n_cand;
    const double* __restrict__ partial = E.acc_partial;
    const double* __restrict__ cand_e = E.cand_e;
    GnScalars* __restrict__ gs = E.gs;
    const CgScalars* __restrict__ cs = E.cg;
    double* __restrict__ energies = E.energies;
    __shared__ double sh[8];
    double nx = 0.0, nh = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        nx += partial[2 * b];
        nh += partial[2 * b + 1];
    }
    const double tx = block_reduce_d(nx, sh);
    const double th = block_reduce_d(nh, sh);
    if (threadIdx.x == 0) {
        const int chosen = chosen_candidate(cand_e, ncand, gs);
        energies[iter] = gs->e_old;
        const double scale = chosen >= 0 ? ldexp(1.0, -chosen) : ldexp(1.0, -ncand);
        const double normv = sqrt(tx);
        const double normh = sqrt(th) * scale;
        if (chosen >= 0) gs->e_old = cand_e[chosen];
        gs->iters_done = iter + 1;
        gs->cg_total += cs->iters;
        energies[iter + 1] = gs->e_old;
        if (stop_rule && normh < (normv + 1e-6) * 1e-6) *gn_done = 1;
    }
}

__global__ void init_x_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    const int n = E.n;
    float* __restrict__ x = E.x;
    GnScalars* __restrict__ gs = E.gs;
    int* __restrict__ gn_done = E.gn_done;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c = i % 12;
        x[i] = (c == 0 || c == 4 || c == 8) ? 1.0f : 0.0f;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *gn_done = 0;
        gs->iters_done = 0;
        gs->cg_total = 0;
        // a barrier failure of an earlier persistent solve must not mark this one
        GridBarrier* gb = const_cast<GridBarrier*>(E.gbar);
        gb->count = 0u;
        gb->gen = 0u;
        gb->err = 0;
    }
}

__global__ void set_eold_kernel(const SolverEntry* __restrict__ B) {
    const SolverEntry E = load_entry(B);
    E.gs->e_old = E.cand_e[0];
    E.energies[0] = E.cand_e[0];
}

// Entries read this frame's node positions and aims.
// Node positions of the frame copied for a side-stream coarse rebuild (the
// caller's buffer changes under it: the node blend writes it in place).
__global__ void snapshot_g_kernel(const SolverEntry* __restrict__ B, const SolverEntry* __restrict__ Bb) {
    const SolverEntry E = load_entry(B);
    const SolverEntry Eb = load_entry(Bb);
    float* __restrict__ dst = const_cast<float*>(Eb.S.g);
    const int64_t n = 3 * (int64_t)E.S.M;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = E.S.g[i];
}

struct FramePtrs {
    const float* g[SOLVER_BATCH_MAX];
    const float* aim[SOLVER_BATCH_MAX];
};

__global__ void set_frame_kernel(SolverEntry* __restrict__ B, int n, FramePtrs f) {
    const int s = threadIdx.x;
    if (s < n) {
        B[s].S.g = f.g[s];
        B[s].S.aim = f.aim[s];
    }
}

// flag = 1 when every scene's cg_done (which 0) or gn_done (which 1) is set.
// flag is pinned host memory: the host reads it once the stream reaches this
// point, without a device-to-host copy.
// iters (when not null, which 0): the largest CG iteration count over the scenes.
// gn (when not null): gn[0] = 1 when every scene's gn_done is set, gn[1] =
// the largest Gauss-Newton iteration count so far. The flag itself is stored
// last, with release semantics at system scope, so a host that sees it (in
// pinned memory, without waiting for the stream) also sees iters and gn.
__global__ void batch_flag_kernel(const SolverEntry* __restrict__ B, int n, int which, int* __restrict__ flag,
                                  int* __restrict__ iters, int* __restrict__ gn) {
    if (threadIdx.x != 0) return;
    int all = 1, most = 0, gn_all = 1, gn_most = 0;
    for (int s = 0; s < n; s++) {
        all &= (which ? *B[s].gn_done : *B[s].cg_done) != 0 ? 1 : 0;
        if (iters) most = max(most, B[s].cg->iters);
        gn_all &= *B[s].gn_done != 0 ? 1 : 0;
        gn_most = max(gn_most, B[s].gs->iters_done);
    }
    if (iters) *iters = most;
    if (gn) {
        gn[0] = gn_all;
        gn[1] = gn_most;
    }
    __hip_atomic_store(flag, all, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Per-scene outputs: x, the energies of the iterations run, the iteration count
// (-1 when the persistent PCG kernel's grid barrier gave up: the solve is then
// invalid) and the node table of putFreeInputs (Deform.hpp:140-151).
__global__ void finish_kernel(const SolverEntry* __restrict__ B, SolverOutPtrs out) {
    const int s = blockIdx.z;
    const SolverEntry E = load_entry(B);
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int done = E.gs->iters_done;
    if (float* xo = out.x[s])
        for (int64_t i = i0; i < E.n; i += stride) xo[i] = E.x[i];
    if (double* eo = out.energies[s])
        for (int64_t i = i0; i <= done; i += stride) eo[i] = E.energies[i];
    if (int* d = out.iters_done[s])
        if (i0 == 0) *d = E.gbar->err ? -1 : done;
    if (NodeXf* table = reinterpret_cast<NodeXf*>(out.table[s])) {
        const SolverDev& S = E.S;
        for (int64_t m = i0; m < S.M; m += stride) {
            NodeXf r;
            r.g[0] = S.g[3 * m]; r.g[1] = S.g[3 * m + 1]; r.g[2] = S.g[3 * m + 2];
            const int f = S.free_id[m];
            if (f >= 0) {
                const float* xf = E.x + 12 * (int64_t)f;
#pragma unroll
                for (int rr = 0; rr < 3; rr++)
#pragma unroll
                    for (int c = 0; c < 3; c++) r.R[rr * 3 + c] = xf[3 * c + rr];  // x is column-major
                r.t[0] = xf[9]; r.t[1] = xf[10]; r.t[2] = xf[11];
            } else {
#pragma unroll
                for (int i = 0; i < 9; i++) r.R[i] = (i % 4 == 0) ? 1.0f : 0.0f;
                r.t[0] = r.t[1] = r.t[2] = 0.0f;
            }
            r.pad = 0.0f;
            table[m] = r;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------- host driver
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

static SolverEntry entry_of(const SolverWork& W) {
    SolverEntry E;
    E.S = W.dev;
    E.n = W.n;
    E.nb_owner = (int)nblk(W.dev.nF, 256 / LANES);
    E.nb_apply = (int)nblk(W.dev.nF, W.cg_block / LANES);
    E.e_blocks = W.e_blocks;
    E.acc_blocks = W.acc_blocks;
    E.n_cand = W.n_cand;
    E.nt = W.dev.nc_pad / GJ_T;
    E.cg_max_iters = W.cg_max_iters;
    E.tol2 = W.cg_tol * W.cg_tol;
    E.x = W.x; E.hsol = W.hsol; E.r = W.r; E.z = W.z; E.p = W.p; E.q = W.q; E.binv = W.binv; E.rc = W.rc;
    E.dmat = W.dmat; E.pair_m = W.pair_m; E.cg_asm = (W.cg_asm && W.dev.n_pairs > 0) ? 1 : 0;
    E.acinv = W.acinv; E.gj_piv = W.gj_piv; E.gj_col = W.gj_col; E.hwarm = W.hwarm; E.ac64 = W.ac64;
    E.partial = W.partial; E.partial2 = W.partial2; E.rzp = W.rzp; E.e_partial = W.e_partial; E.cand_e = W.cand_e;
    E.acc_partial = W.acc_partial; E.energies = W.energies;
    E.cg = W.cg; E.gs = W.gs; E.gn_done = W.gn_done; E.cg_done = W.cg_done;
    E.gbar = reinterpret_cast<const GridBarrier*>(W.gbar);
    return E;
}

// The entry of a side-stream coarse rebuild: node positions from the snapshot,
// x0 = identity, results into the spare centre / inverse buffers.
static SolverEntry build_entry_of(const SolverWork& W) {
    SolverEntry E = entry_of(W);
    E.S.g = W.g_snap;
    E.S.agg_center = W.agg_center_alt;
    E.acinv = W.acinv_alt;
    E.x = W.x_ident;
    return E;
}

void solver_batch_drop_graphs(SolverBatch& B) {
    for (auto& kv : B.graph_cache)
        if (kv.second) (void)hipGraphExecDestroy(kv.second);
    B.graph_cache.clear();
}

// Launches the graph cached under key, capturing body() on st first.
template <typename Body>
static int graph_launch(SolverBatch& B, const std::array<int64_t, 6>& key, hipStream_t st, Body&& body) {
    auto it = B.graph_cache.find(key);
    if (it == B.graph_cache.end()) {
        HIP_RET(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = body();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : -(int)e - 1000;
        }
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_RET(ei);
        it = B.graph_cache.emplace(key, x).first;
    }
    HIP_RET(hipGraphLaunch(it->second, st));
    return 0;
}

void solver_batch_free(SolverBatch& B) {
    solver_batch_drop_graphs(B);
    if (B.entries) (void)hipFree(B.entries);
    if (B.entries_build) (void)hipFree(B.entries_build);
    if (B.fork_ev) (void)hipEventDestroy(B.fork_ev);
    if (B.side) (void)hipStreamDestroy(B.side);
    B.entries = nullptr;
    B.entries_build = nullptr;
    B.fork_ev = nullptr;
    B.side = nullptr;
    B.w.clear();
}

int solver_batch_init(SolverBatch& B, const std::vector<SolverWork*>& w, std::string& err) {
    // settings survive a rebuild of the entries
    const bool async = B.async, graphs = B.graphs;
    const int lag = B.lag;
    hipStream_t side_ext = B.side_ext;
    solver_batch_free(B);
    if (w.empty() || (int)w.size() > SOLVER_BATCH_MAX) {
        err = "solver batch: need 1.." + std::to_string(SOLVER_BATCH_MAX) + " solvers";
        return -1;
    }
    B = SolverBatch();
    B.async = async;
    B.graphs = graphs;
    B.lag = lag;
    B.side_ext = side_ext;
    B.w = w;
    B.kt = 8;
    std::vector<SolverEntry> host;
    for (SolverWork* W : w) {
        if (!W) { err = "solver batch: null solver"; return -1; }
        host.push_back(entry_of(*W));
        const SolverEntry& E = host.back();
        if (W->dev.K != 8) B.kt = 0;
        B.max_nb_owner = std::max(B.max_nb_owner, E.nb_owner);
        B.max_wt_blocks = std::max(B.max_wt_blocks,
                                   (int)nblk((int64_t)E.S.nF * E.S.K + E.S.n_in + E.S.n_cm + E.S.n_mem, 256));
        B.max_nb_apply = std::max(B.max_nb_apply, E.nb_apply);
        if (W->cg_block != w[0]->cg_block) {
            err = "solver batch: every solver of a batch needs the same CG block size";
            return -1;
        }
        B.max_agg = std::max(B.max_agg, E.S.n_agg);
        B.max_e_blocks = std::max(B.max_e_blocks, E.e_blocks);
        B.max_acc_blocks = std::max(B.max_acc_blocks, E.acc_blocks);
        B.max_n = std::max(B.max_n, std::max(E.n, E.S.M));
        B.max_groups128 = std::max(B.max_groups128, (int)nblk(E.S.n_groups, 128));
        B.max_multi = std::max(B.max_multi, E.S.n_multi);
        B.max_nt = std::max(B.max_nt, E.nt);
        B.max_cand = std::max(B.max_cand, E.n_cand);
        B.g.push_back(nullptr);  // the first set_frame always writes the entries' pointers
        B.aim.push_back(nullptr);
    }
    std::vector<SolverEntry> host_build;
    for (SolverWork* W : w) host_build.push_back(build_entry_of(*W));
    const size_t bytes = sizeof(SolverEntry) * host.size();
    if (hipMalloc(&B.entries, bytes) != hipSuccess ||
        hipMemcpy(B.entries, host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&B.entries_build, bytes) != hipSuccess ||
        hipMemcpy(B.entries_build, host_build.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipStreamCreateWithFlags(&B.side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&B.fork_ev, hipEventDisableTiming) != hipSuccess) {
        err = "solver batch: device allocation failed";
        solver_batch_free(B);
        return -2;
    }
    return 0;
}

int solver_batch_set_frame(SolverBatch& B, const float* const* g, const float* const* aim, hipStream_t st) {
    const int n = (int)B.w.size();
    bool changed = false;
    FramePtrs f;
    for (int s = 0; s < n; s++) {
        changed |= g[s] != B.g[s] || aim[s] != B.aim[s];
        f.g[s] = g[s];
        f.aim[s] = aim[s];
        // the persistent PCG kernel takes the scene's SolverDev by value
        B.w[s]->dev.g = g[s];
        B.w[s]->dev.aim = aim[s];
    }
    if (!changed) return 0;
    hipLaunchKernelGGL(set_frame_kernel, dim3(1), dim3(SOLVER_BATCH_MAX), 0, st, B.entries, n, f);
    HIP_RET(hipGetLastError());
    for (int s = 0; s < n; s++) {
        B.g[s] = g[s];
        B.aim[s] = aim[s];
    }
    return 0;
}

int solver_batch_flag(SolverBatch& B, int which, hipStream_t st, int* flag_h, int k_done, int* iters_h, int* gn_h) {
    if (which != 0) iters_h = nullptr;
    if (which == 0 && B.fused)
        hipLaunchKernelGGL(fused_flag_kernel, dim3(1), dim3(64), 0, st, (const SolverEntry*)B.entries, (int)B.w.size(),
                           k_done, flag_h, iters_h, gn_h);
    else
        hipLaunchKernelGGL(batch_flag_kernel, dim3(1), dim3(64), 0, st, (const SolverEntry*)B.entries, (int)B.w.size(),
                           which, flag_h, iters_h, gn_h);
    HIP_RET(hipGetLastError());
    return 0;
}

int solver_cg_chunk(SolverBatch& B, int k0, int L, hipStream_t st, int* flag_h, int* iters_h) {
    const std::array<int64_t, 6> key = {0, k0, L, (int64_t)B.fused, (int64_t)reinterpret_cast<uintptr_t>(flag_h),
                                        (int64_t)reinterpret_cast<uintptr_t>(iters_h)};
    return graph_launch(B, key, st, [&]() -> int {
        int rc = 0;
        for (int k = k0; k < k0 + L && !rc; k++) rc = solver_cg_iteration(B, k, st);
        if (!rc) rc = solver_batch_flag(B, 0, st, flag_h, k0 + L, iters_h);
        return rc;
    });
}

int solver_gn_begin_graph(SolverBatch& B, int iter, hipStream_t st) {
    const std::array<int64_t, 6> key = {1, iter, (int64_t)B.w[0]->warm, (int64_t)B.w[0]->warm_slots, 0, 0};
    return graph_launch(B, key, st, [&]() -> int { return solver_gn_begin(B, iter, st); });
}

int solver_gn_end_graph(SolverBatch& B, int iter, int stop_rule, hipStream_t st, int* flag_h) {
    const std::array<int64_t, 6> key = {2, iter, stop_rule, (int64_t)B.w[0]->warm, (int64_t)B.w[0]->warm_slots,
                                        (int64_t)reinterpret_cast<uintptr_t>(flag_h)};
    return graph_launch(B, key, st, [&]() -> int {
        int rc = solver_gn_end(B, iter, stop_rule, st);
        if (!rc) rc = solver_batch_flag(B, 1, st, flag_h);
        return rc;
    });
}

int solver_batch_finish(SolverBatch& B, const SolverOutPtrs& out, hipStream_t st) {
    const unsigned nz = (unsigned)B.w.size();
    hipLaunchKernelGGL(finish_kernel, dim3(std::max(1u, std::min(64u, nblk(B.max_n, 256))), 1, nz), dim3(256), 0, st,
                       (const SolverEntry*)B.entries, out);
    HIP_RET(hipGetLastError());
    return 0;
}

// Galerkin coarse operator at x0 = identity and its inverse for the entries E
// (it is only a preconditioner, so callers may keep it across frames).
static int coarse_build(const SolverBatch& B, const SolverEntry* E, hipStream_t st) {
    const unsigned nz = (unsigned)B.w.size();
    hipLaunchKernelGGL(agg_center_kernel, dim3(B.max_agg, 1, nz), dim3(256), 0, st, E);
    hipLaunchKernelGGL(zero_coarse_kernel, dim3(256, 1, nz), dim3(256), 0, st, E);
    hipLaunchKernelGGL(coarse_assemble_kernel, dim3(B.max_groups128, 1, nz), dim3(128), 0, st, E);
    if (B.max_multi > 0)
        hipLaunchKernelGGL(coarse_multi_kernel, dim3(B.max_multi, 1, nz), dim3(256), (size_t)24 * B.max_agg, st, E);
    hipLaunchKernelGGL(coarse_to_float_kernel, dim3(256, 1, nz), dim3(256), 0, st, E);
    hipLaunchKernelGGL(pad_identity_kernel, dim3(1, 1, nz), dim3(64), 0, st, E);
    const unsigned nt = (unsigned)B.max_nt;
    hipLaunchKernelGGL(gj_start_kernel, dim3(nt, 1, nz), dim3(256), 0, st, E);
    for (int kb = 0; kb < (int)nt; kb++)
        hipLaunchKernelGGL(gj_step_kernel, dim3(nt * nt, 1, nz), dim3(256), 0, st, E, kb);
    HIP_RET(hipGetLastError());
    return 0;
}

int solver_frame_setup(SolverBatch& B, bool refresh_coarse, hipStream_t st) {
    const SolverEntry* E = B.entries;
    const unsigned nz = (unsigned)B.w.size();
    // adopt side-stream rebuilds that the lag-th earlier solve started
    bool any_pending = false;
    for (SolverWork* W : B.w) {
        if (!W->pending) continue;
        if (++W->pending_age < B.lag) {
            any_pending = true;
            continue;
        }
        HIP_RET(hipStreamWaitEvent(st, W->built_ev, 0));
        const size_t nc = (size_t)W->dev.nc_pad;
        HIP_RET(hipMemcpyAsync(W->acinv, W->acinv_alt, sizeof(float) * nc * nc, hipMemcpyDeviceToDevice, st));
        HIP_RET(hipMemcpyAsync(W->dev.agg_center, W->agg_center_alt, sizeof(float) * 3 * (size_t)W->dev.n_agg,
                               hipMemcpyDeviceToDevice, st));
        W->pending = false;
    }
    hipLaunchKernelGGL(init_x_kernel, dim3(nblk(B.max_n, 256), 1, nz), dim3(256), 0, st, E);
    hipLaunchKernelGGL(walk_table_kernel, dim3(B.max_wt_blocks, 1, nz), dim3(256), 0, st, E);
    if (B.w[0]->cg_asm)
        hipLaunchKernelGGL(pair_coef_kernel, dim3(nblk(B.max_n, 256), 1, nz), dim3(256), 0, st, E);
    if (refresh_coarse && !any_pending) {
        bool all_have = true;
        for (const SolverWork* W : B.w) all_have = all_have && W->have_coarse;
        if (!B.async || !all_have) {
            int rc = coarse_build(B, E, st);
            if (rc) return rc;
            for (SolverWork* W : B.w) W->have_coarse = true;
        } else {
            // rebuild from this frame's node positions beside the solve; the
            // next frame adopts it (this one keeps the current operator)
            hipLaunchKernelGGL(snapshot_g_kernel, dim3(64, 1, nz), dim3(256), 0, st, E,
                               (const SolverEntry*)B.entries_build);
            HIP_RET(hipGetLastError());
            HIP_RET(hipEventRecord(B.fork_ev, st));
            hipStream_t side = B.side_ext ? B.side_ext : B.side;
            HIP_RET(hipStreamWaitEvent(side, B.fork_ev, 0));
            int rc = coarse_build(B, B.entries_build, side);
            if (rc) return rc;
            for (SolverWork* W : B.w) {
                HIP_RET(hipEventRecord(W->built_ev, side));
                W->pending = true;
                W->pending_age = 0;
            }
        }
    }
    // E(x0)
    hipLaunchKernelGGL(energy_kernel, dim3(B.max_e_blocks, 1, nz), dim3(256), 0, st, E, 0, 0, 0, 0,
                       (const float*)nullptr);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1, 1, nz), dim3(256), 0, st, E, 0, 1, 0, 0, (double*)nullptr);
    hipLaunchKernelGGL(set_eold_kernel, dim3(1, 1, nz), dim3(1), 0, st, E);
    HIP_RET(hipGetLastError());
    return 0;
}

// The apply kernel of the batch's K (unrolled for 8) and CG block size.
static int launch_apply(const SolverBatch& B, unsigned nz, hipStream_t st, int k, int first, int warm_iter,
                        const float* v_ext) {
    const SolverEntry* E = B.entries;
    const dim3 grid(B.max_nb_apply, 1, nz);
    const int bs = B.w[0]->cg_block;
    if (B.w[0]->cg_asm && B.w[0]->dev.n_pairs > 0) {
#define ARAPGS_APPLY_ASM(BSV) \
    hipLaunchKernelGGL((apply_asm_kernel<BSV>), grid, dim3(BSV), 0, st, E, k, first, warm_iter, v_ext)
        if (bs == 64) ARAPGS_APPLY_ASM(64);
        else if (bs == 128) ARAPGS_APPLY_ASM(128);
        else if (bs == 256) ARAPGS_APPLY_ASM(256);
        else ARAPGS_APPLY_ASM(512);
#undef ARAPGS_APPLY_ASM
        HIP_RET(hipGetLastError());
        return 0;
    }
#define ARAPGS_APPLY(KTV, BSV) \
    hipLaunchKernelGGL((apply_kernel<KTV, BSV>), grid, dim3(BSV), 0, st, E, k, first, warm_iter, v_ext)
    if (B.kt == 8) {
        if (bs == 64) ARAPGS_APPLY(8, 64);
        else if (bs == 128) ARAPGS_APPLY(8, 128);
        else if (bs == 256) ARAPGS_APPLY(8, 256);
        else ARAPGS_APPLY(8, 512);
    } else {
        if (bs == 64) ARAPGS_APPLY(0, 64);
        else if (bs == 128) ARAPGS_APPLY(0, 128);
        else if (bs == 256) ARAPGS_APPLY(0, 256);
        else ARAPGS_APPLY(0, 512);
    }
#undef ARAPGS_APPLY
    HIP_RET(hipGetLastError());
    return 0;
}

int solver_gn_begin(SolverBatch& B, int iter, hipStream_t st) {
    const SolverEntry* E = B.entries;
    const unsigned nz = (unsigned)B.w.size();
    hipLaunchKernelGGL(rhs_diag_kernel, dim3(B.max_nb_owner, 1, nz), dim3(256), 0, st, E);
    hipLaunchKernelGGL(cg_init_kernel, dim3(1, 1, nz), dim3(64), 0, st, E);
    if (B.w[0]->warm && iter < B.w[0]->warm_slots) {
        // q = A h0 (the first-iteration form of the apply), then h = h0, r = b - A h0
        int rc = launch_apply(B, nz, st, 0, 1, iter, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(warm_init_kernel, dim3(nblk(B.max_n, 256), 1, nz), dim3(256), 0, st, E, iter);
    } else {
        hipLaunchKernelGGL(zero_step_kernel, dim3(256, 1, nz), dim3(256), 0, st, E);
    }
    // z0 = M^-1 r0, rz_0 -> rz[0]
    hipLaunchKernelGGL(update_restrict_kernel, dim3(B.max_agg, 1, nz), dim3(256), 0, st, E, 0, 0);
    hipLaunchKernelGGL(coarse_prolong_kernel, dim3(B.max_agg, 1, nz), dim3(256), 0, st, E, 0, 0);
    HIP_RET(hipGetLastError());
    return 0;
}

int solver_cg_iteration(SolverBatch& B, int k, hipStream_t st) {
    const SolverEntry* E = B.entries;
    const unsigned nz = (unsigned)B.w.size();
    if (B.fused) {
        const dim3 grid(B.max_agg, 1, nz);
        const int bs = B.w[0]->cg_block;
#define ARAPGS_FUSED(BSV)                                                                                   \
    do {                                                                                                    \
        if (B.kt == 8) hipLaunchKernelGGL((cg_apply_agg_kernel<8, BSV>), grid, dim3(BSV), 0, st, E, k);     \
        else hipLaunchKernelGGL((cg_apply_agg_kernel<0, BSV>), grid, dim3(BSV), 0, st, E, k);               \
        hipLaunchKernelGGL(cg_update_coarse_kernel<BSV>, grid, dim3(BSV), 0, st, E, k);                     \
    } while (0)
        if (bs == 64) ARAPGS_FUSED(64);
        else if (bs == 128) ARAPGS_FUSED(128);
        else if (bs == 256) ARAPGS_FUSED(256);
        else ARAPGS_FUSED(512);
#undef ARAPGS_FUSED
        HIP_RET(hipGetLastError());
        return 0;
    }
    int rc = launch_apply(B, nz, st, k, k == 0 ? 1 : 0, -1, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(update_restrict_kernel, dim3(B.max_agg, 1, nz), dim3(256), 0, st, E, 1, k);
    hipLaunchKernelGGL(coarse_prolong_kernel, dim3(B.max_agg, 1, nz), dim3(256), 0, st, E, k, 1);
    HIP_RET(hipGetLastError());
    return 0;
}

int solver_cg_persistent(SolverWork& W, hipStream_t st) {
    const SolverDev& S = W.dev;
    if (W.coop_grid <= 0) {
        int per_cu = 0, dev = 0, ncu = 0;
        HIP_RET(hipGetDevice(&dev));
        HIP_RET(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HIP_RET(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(pcg_persistent_kernel<0>),
                                                             256, 0));
        const int want = std::max((int)nblk(S.nF, 256 / LANES), S.n_agg);
        W.coop_grid = std::max(1, std::min(want, per_cu * ncu));
    }
    GridBarrier* gb = reinterpret_cast<GridBarrier*>(W.gbar);
    hipLaunchKernelGGL(grid_barrier_reset_kernel, dim3(1), dim3(1), 0, st, gb);
    const float* x = W.x;
    const float* binv = W.binv;
    const float* acinv = W.acinv;
    SolverDev Sv = S;
    void* args[] = {&Sv,  (void*)&x,   &W.hsol,    &W.r,     &W.z,   &W.p, &W.q,       (void*)&binv,
                    &W.rc, (void*)&acinv, &W.partial, &W.partial2, &W.rzp, &W.cg, &gb, &W.cg_done};
    const void* kern = S.K == 8 ? reinterpret_cast<const void*>(pcg_persistent_kernel<8>)
                                : reinterpret_cast<const void*>(pcg_persistent_kernel<0>);
    HIP_RET(hipLaunchCooperativeKernel(kern, dim3(W.coop_grid),
                                       dim3(256), args, 0, st));
    return 0;
}

int solver_barrier_error(SolverWork& W, int* err) {
    GridBarrier gb;
    HIP_RET(hipMemcpy(&gb, W.gbar, sizeof gb, hipMemcpyDeviceToHost));
    *err = gb.err;
    return 0;
}

int solver_gn_end(SolverBatch& B, int iter, int stop_rule, hipStream_t st) {
    const SolverEntry* E = B.entries;
    const unsigned nz = (unsigned)B.w.size();
    if (B.w[0]->warm && iter < B.w[0]->warm_slots)
        hipLaunchKernelGGL(save_warm_kernel, dim3(nblk(B.max_n, 256), 1, nz), dim3(256), 0, st, E, iter);
    // the full step first; the 49 halvings only if it was rejected
    hipLaunchKernelGGL(energy_kernel, dim3(B.max_e_blocks, 1, nz), dim3(256), 0, st, E, 1, 1, 0, 0,
                       (const float*)nullptr);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1, 1, nz), dim3(256), 0, st, E, 0, 1, 1, 0, (double*)nullptr);
    hipLaunchKernelGGL(energy_kernel, dim3(B.max_e_blocks, B.max_cand - 1, nz), dim3(256), 0, st, E, 1, 1, 1, 1,
                       (const float*)nullptr);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1, 1, nz), dim3(256), 0, st, E, 1, B.max_cand - 1, 1, 1,
                       (double*)nullptr);
    hipLaunchKernelGGL(accept_partial_kernel, dim3(B.max_acc_blocks, 1, nz), dim3(256), 0, st, E);
    hipLaunchKernelGGL(accept_final_kernel, dim3(1, 1, nz), dim3(256), 0, st, E, iter, stop_rule);
    HIP_RET(hipGetLastError());
    return 0;
}

int solver_apply_test(SolverBatch& B, const float* x, const float* v, float* q_out, float* b_out, hipStream_t st) {
    SolverWork& W = *B.w[0];
    const SolverEntry* E = B.entries;
    const unsigned nb_owner = nblk(W.dev.nF, 256 / LANES);
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(W.n, 256)), dim3(256), 0, st, W.x, x, (int64_t)W.n, (const int*)nullptr);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<float*>(W.cg_done), (int64_t)1,
                       (const int*)nullptr);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<float*>(W.gn_done), (int64_t)1,
                       (const int*)nullptr);
    hipLaunchKernelGGL(walk_table_kernel,
                       dim3(nblk((int64_t)W.dev.nF * W.dev.K + W.dev.n_in + W.dev.n_cm + W.dev.n_mem, 256)), dim3(256),
                       0, st, E);
    hipLaunchKernelGGL(rhs_diag_kernel, dim3(nb_owner), dim3(256), 0, st, E);
    if (W.cg_asm) hipLaunchKernelGGL(pair_coef_kernel, dim3(nblk(W.dev.nF, 256)), dim3(256), 0, st, E);
    int rc = launch_apply(B, 1, st, 0, 1, -1, v);
    if (rc) return rc;
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(W.n, 256)), dim3(256), 0, st, q_out, (const float*)W.q, (int64_t)W.n,
                       (const int*)nullptr);
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(W.n, 256)), dim3(256), 0, st, b_out, (const float*)W.r, (int64_t)W.n,
                       (const int*)nullptr);
    HIP_RET(hipGetLastError());
    return 0;
}

int solver_energy(SolverBatch& B, const float* x, double* out, hipStream_t st) {
    const SolverEntry* E = B.entries;
    hipLaunchKernelGGL(energy_kernel, dim3(B.max_e_blocks, 1, 1), dim3(256), 0, st, E, 0, 0, 0, 0, x);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1, 1, 1), dim3(256), 0, st, E, 0, 1, 0, 0, out);
    HIP_RET(hipGetLastError());
    return 0;
}

}  // namespace arapgs
