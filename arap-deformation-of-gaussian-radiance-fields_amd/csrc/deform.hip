// Per-frame deformation of Gaussians, end points, nodes and samples (gfx950).
//
//   * end points (GaussianView.cpp:4643-4668)
//   * node transform table from the solved unknowns (DeformGraph::putFreeInputs,
//     Deform.hpp:140-151)
//   * K-node blend p' = sum_j w_j (R_j (p - g_j) + t_j + g_j) (predict_mesh /
//     predict_samples, Deform.hpp:230-268) for samples and nodes
//   * the fused per-Gaussian update: blend of the 6 end points, FP64 polar of
//     M = [(e0+ - e0-)/2 | (e1+ - e1-)/2 | (e2+ - e2-)/2], quaternion, scales,
//     centre and SH rotation (UpdatePosition + UpdateAsSixPointsWithdrawBad,
//     GaussianView.cpp:2986-3166)
//   * node quaternions and the per-sample slerp-blend SH rotation
//     (FastUpdateSamplesSH + RotateSHs, GaussianView.cpp:3169-3186,
//     cudakdtree.cu:201-222)
//
// Node tables are 64-byte records gathered through L2 (a frame's table is at
// most a few hundred KB); all per-point streams are read once and written once.
#include "common.hpp"

namespace arapgs {

namespace {

// One blended point with the arithmetic of Deform.hpp:238-246: p - g_j in
// float, the term w_j (R_j(p - g_j) + t_j + g_j) evaluated in double left to
// right, and a float accumulator updated once per neighbour.
__device__ __forceinline__ void blend_point(const NodeXf* __restrict__ table, const uint32_t* __restrict__ idx,
                                            const float* __restrict__ w, int k, const float p[3], float o[3]) {
#pragma clang fp contract(off)
    o[0] = o[1] = o[2] = 0.0f;
    for (int j = 0; j < k; j++) {
        const NodeXf& n = table[idx[j]];
        const double wj = (double)w[j];
        const float tx = p[0] - n.g[0], ty = p[1] - n.g[1], tz = p[2] - n.g[2];
        o[0] = (float)((double)o[0] + wj * ((double)n.R[0] * tx + (double)n.R[1] * ty + (double)n.R[2] * tz + (double)n.t[0] + (double)n.g[0]));
        o[1] = (float)((double)o[1] + wj * ((double)n.R[3] * tx + (double)n.R[4] * ty + (double)n.R[5] * tz + (double)n.t[1] + (double)n.g[1]));
        o[2] = (float)((double)o[2] + wj * ((double)n.R[6] * tx + (double)n.R[7] * ty + (double)n.R[8] * tz + (double)n.t[2] + (double)n.g[2]));
    }
}

template <int K>
__device__ __forceinline__ void blend_point_regs(const NodeXf* __restrict__ table, const uint32_t (&id)[K],
                                                 const float (&wt)[K], const float p[3], float o[3]) {
#pragma clang fp contract(off)
    o[0] = o[1] = o[2] = 0.0f;
#pragma unroll
    for (int j = 0; j < K; j++) {
        const float4* r = reinterpret_cast<const float4*>(table + id[j]);
        const float4 a = r[0], b = r[1], c = r[2], d = r[3];
        // a = g0 g1 g2 R0 | b = R1..R4 | c = R5..R8 | d = t0 t1 t2 pad
        const float tx = p[0] - a.x, ty = p[1] - a.y, tz = p[2] - a.z;
        const double wj = (double)wt[j];
        o[0] = (float)((double)o[0] + wj * ((double)a.w * tx + (double)b.x * ty + (double)b.y * tz + (double)d.x + (double)a.x));
        o[1] = (float)((double)o[1] + wj * ((double)b.z * tx + (double)b.w * ty + (double)c.x * tz + (double)d.y + (double)a.y));
        o[2] = (float)((double)o[2] + wj * ((double)c.y * tx + (double)c.z * ty + (double)c.w * tz + (double)d.z + (double)a.z));
    }
}

// Node-row cache of a workgroup (CACHE): the block's 256 Gaussians lie close
// together in the cell order, so their 12,288 end-point neighbour ids name a
// few dozen nodes. A direct-mapped table of NODE_CACHE_SLOTS 64-byte rows in
// LDS (slot = id mod slots) is filled by the first thread that needs a node
// (compare-and-swap on the slot's tag; tags only ever go from empty to one id,
// so a filled row never changes), and every gather whose slot holds its id
// reads LDS instead of going through the texture path; ids whose slot another
// node took read the global table. The rows are the same floats, so the blend
// is bit-identical.
constexpr int NODE_CACHE_SLOTS = 256;
// A cached row takes 4 16-byte units. Padding rows to 5 units (80 bytes) spread
// neighbouring slots over more bank groups but made both kernels slower (six
// point 357 vs 303 us, sample blend 222 vs 169 us at 1M Gaussians / 5.6M
// samples): the extra 4 KB per workgroup costs occupancy, which matters more
// than the bank conflicts of the unpadded rows.
constexpr int NODE_CACHE_ROW4 = 4;

// Claims the empty slots of a point's neighbours and fills the claimed rows;
// the caller synchronises the block before any lane reads the cache.
template <int K>
__device__ __forceinline__ void node_cache_claim(const NodeXf* __restrict__ table, const uint32_t (&id)[K], int* s_tag,
                                                 float4* s_rows) {
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int slot = (int)(id[j] & (NODE_CACHE_SLOTS - 1));
        if (s_tag[slot] == -1 && atomicCAS(&s_tag[slot], -1, (int)id[j]) == -1) {
            const float4* src = reinterpret_cast<const float4*>(table + id[j]);
#pragma unroll
            for (int q = 0; q < 4; q++) s_rows[NODE_CACHE_ROW4 * slot + q] = src[q];
        }
    }
}

template <int K>
__device__ __forceinline__ void blend_point_cached(const NodeXf* __restrict__ table, const int* s_tag,
                                                   const float4* s_rows, const uint32_t (&id)[K],
                                                   const float (&wt)[K], const float p[3], float o[3]) {
#pragma clang fp contract(off)
    o[0] = o[1] = o[2] = 0.0f;
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int slot = (int)(id[j] & (NODE_CACHE_SLOTS - 1));
        const float4* r = s_tag[slot] == (int)id[j] ? s_rows + NODE_CACHE_ROW4 * slot
                                                    : reinterpret_cast<const float4*>(table + id[j]);
        const float4 a = r[0], b = r[1], c = r[2], d = r[3];
        const float tx = p[0] - a.x, ty = p[1] - a.y, tz = p[2] - a.z;
        const double wj = (double)wt[j];
        o[0] = (float)((double)o[0] + wj * ((double)a.w * tx + (double)b.x * ty + (double)b.y * tz + (double)d.x + (double)a.x));
        o[1] = (float)((double)o[1] + wj * ((double)b.z * tx + (double)b.w * ty + (double)c.x * tz + (double)d.y + (double)a.y));
        o[2] = (float)((double)o[2] + wj * ((double)c.y * tx + (double)c.z * ty + (double)c.w * tz + (double)d.z + (double)a.z));
    }
}

// K neighbour ids and weights of one point; 16-byte vector loads when the
// rows are 16-byte aligned (K % 4 == 0 and aligned base pointers).
template <int K, bool VEC>
__device__ __forceinline__ void load_nbrs(const uint32_t* __restrict__ idx, const float* __restrict__ w,
                                          uint32_t (&id)[K], float (&wt)[K]) {
    if constexpr (VEC && K % 4 == 0) {
#pragma unroll
        for (int j = 0; j < K; j += 4) {
            const uint4 a = *reinterpret_cast<const uint4*>(idx + j);
            const float4 b = *reinterpret_cast<const float4*>(w + j);
            id[j] = a.x; id[j + 1] = a.y; id[j + 2] = a.z; id[j + 3] = a.w;
            wt[j] = b.x; wt[j + 1] = b.y; wt[j + 2] = b.z; wt[j + 3] = b.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < K; j++) { id[j] = idx[j]; wt[j] = w[j]; }
    }
}

__global__ void end_points_kernel(const float* __restrict__ pos, const float* __restrict__ rot,
                                  const float* __restrict__ scale, int64_t N, float pad, float coeff,
                                  float* __restrict__ ends) {
#pragma clang fp contract(off)
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    float q[4] = {rot[4 * g], rot[4 * g + 1], rot[4 * g + 2], rot[4 * g + 3]};
    quat_normalize(q);
    float R[3][3];
    quat_to_mat(q[0], q[1], q[2], q[3], R);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float len = (scale[3 * g + i] + pad) * coeff;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float off = R[c][i] * len;
            ends[g * 18 + i * 6 + c] = pos[3 * g + c] + off;
            ends[g * 18 + i * 6 + 3 + c] = pos[3 * g + c] - off;
        }
    }
}

__global__ void node_table_kernel(int M, const float* __restrict__ node_pos, const float* __restrict__ x,
                                  const int* __restrict__ free_id, NodeXf* __restrict__ table) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= M) return;
    NodeXf r;
    r.g[0] = node_pos[3 * n]; r.g[1] = node_pos[3 * n + 1]; r.g[2] = node_pos[3 * n + 2];
    const int f = free_id ? free_id[n] : -1;
    if (f >= 0) {
        const float* xf = x + 12 * (int64_t)f;
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
    table[n] = r;
}

// VEC: the K indices and weights of a point are read as 16-byte vectors (rows
// 16-byte aligned): two loads per array instead of K, each covering half of
// every lane's 32-byte row.
// CACHE: the node rows go through the workgroup's LDS node-row cache (every
// thread reaches the block barrier; results identical).
template <int K, bool VEC, bool CACHE>
__global__ void __launch_bounds__(256) blend_points_kernel(const float* pts, int64_t P,
                                                           const uint32_t* __restrict__ idx,
                                                           const float* __restrict__ w,
                                                           const NodeXf* __restrict__ table,
                                                           const uint8_t* __restrict__ skip,
                                                           float* out) {
    __shared__ int s_tag[CACHE ? NODE_CACHE_SLOTS : 1];
    __shared__ float4 s_rows[CACHE ? NODE_CACHE_ROW4 * NODE_CACHE_SLOTS : 1];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool here = i < P;
    const bool moves = here && !(skip && skip[i]);
    uint32_t id[K];
    float wt[K];
    if (moves) load_nbrs<K, VEC>(idx + i * K, w + i * K, id, wt);
    if constexpr (CACHE) {
        for (int t = threadIdx.x; t < NODE_CACHE_SLOTS; t += blockDim.x) s_tag[t] = -1;
        __syncthreads();
        if (moves) node_cache_claim<K>(table, id, s_tag, s_rows);
        __syncthreads();
    }
    if (!here) return;
    const float p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    if (!moves) {
        out[3 * i] = p[0]; out[3 * i + 1] = p[1]; out[3 * i + 2] = p[2];
        return;
    }
    float o[3];
    if constexpr (CACHE)
        blend_point_cached<K>(table, s_tag, s_rows, id, wt, p, o);
    else
        blend_point_regs<K>(table, id, wt, p, o);
    out[3 * i] = o[0]; out[3 * i + 1] = o[1]; out[3 * i + 2] = o[2];
}

__global__ void __launch_bounds__(256) blend_points_generic_kernel(const float* pts, int64_t P,
                                                                   const uint32_t* __restrict__ idx,
                                                                   const float* __restrict__ w, int k,
                                                                   const NodeXf* __restrict__ table,
                                                                   const uint8_t* __restrict__ skip,
                                                                   float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    if (skip && skip[i]) {
        out[3 * i] = p[0]; out[3 * i + 1] = p[1]; out[3 * i + 2] = p[2];
        return;
    }
    float o[3];
    blend_point(table, idx + i * k, w + i * k, k, p, o);
    out[3 * i] = o[0]; out[3 * i + 1] = o[1]; out[3 * i + 2] = o[2];
}

// 48-float rows (SH coefficients / sample features) of a block's 256
// consecutive items moved between global memory and the owning threads'
// registers through LDS, in passes of as many rows as `lds` holds. The global
// side is whole contiguous float4 runs (a wave's access covers 1 KB), so every
// cache line is fetched once. LDS is accessed in 16-byte units with a row
// stride of 13 units (52 floats): 13 is odd, so the 8 lanes that share a bank
// cycle touch 8 different 16-byte bank groups both when a wave writes 64
// consecutive units of the global side and when each lane reads its own row
// (a 49-float stride with 4-byte accesses lost 26-28% of the LDS cycles to
// bank conflicts). Every thread of the block must call these (they
// synchronise); rows [0, n_rows) exist.
constexpr int ROW48_STRIDE4 = 13;

__device__ __forceinline__ void rows48_load(const float* __restrict__ src, int n_rows, float* lds, int lds_floats,
                                            float (&c)[48]) {
    const int t = threadIdx.x;
    const int per = lds_floats / (4 * ROW48_STRIDE4);
    float4* l4 = reinterpret_cast<float4*>(lds);
    for (int r0 = 0; r0 < n_rows; r0 += per) {
        const int nr = min(per, n_rows - r0);
        const float4* s4 = reinterpret_cast<const float4*>(src + (int64_t)r0 * 48);
        for (int q = t; q < nr * 12; q += blockDim.x) l4[(q / 12) * ROW48_STRIDE4 + q % 12] = s4[q];
        __syncthreads();
        if (t >= r0 && t < r0 + nr) {
            const float4* d = l4 + (t - r0) * ROW48_STRIDE4;
#pragma unroll
            for (int k = 0; k < 12; k++) {
                const float4 v = d[k];
                c[4 * k] = v.x; c[4 * k + 1] = v.y; c[4 * k + 2] = v.z; c[4 * k + 3] = v.w;
            }
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void rows48_store(float* __restrict__ dst, int n_rows, float* lds, int lds_floats,
                                             const float (&c)[48]) {
    const int t = threadIdx.x;
    const int per = lds_floats / (4 * ROW48_STRIDE4);
    float4* l4 = reinterpret_cast<float4*>(lds);
    for (int r0 = 0; r0 < n_rows; r0 += per) {
        const int nr = min(per, n_rows - r0);
        if (t >= r0 && t < r0 + nr) {
            float4* d = l4 + (t - r0) * ROW48_STRIDE4;
#pragma unroll
            for (int k = 0; k < 12; k++) d[k] = make_float4(c[4 * k], c[4 * k + 1], c[4 * k + 2], c[4 * k + 3]);
        }
        __syncthreads();
        float4* d4 = reinterpret_cast<float4*>(dst + (int64_t)r0 * 48);
        for (int q = t; q < nr * 12; q += blockDim.x) d4[q] = l4[(q / 12) * ROW48_STRIDE4 + q % 12];
        __syncthreads();
    }
}

// The fused per-Gaussian update, one workgroup per 256 Gaussians in two
// phases. Phase 1 is point-parallel: each thread blends 6 of the block's 1536
// end points, so consecutive lanes read consecutive neighbour rows (32 B each
// at K = 8) and the rows are consumed right after they arrive; blended points
// go to LDS and back to `ends`. Phase 2 is Gaussian-parallel: polar factor,
// quaternion, scale, centre and the SH rotation of one Gaussian per thread;
// the SH rows then move through the same LDS in coalesced passes (loaded after
// the polar so the 48 coefficients are not live across it: 113 VGPRs, four
// waves per SIMD; 327 instead of 356 us at 1M Gaussians).
// Static Gaussians (every end-point neighbour a static node) still blend their
// end points, as the reference does, but keep their attributes (their SH rows
// are written back unchanged).
template <int K, bool VEC, bool CACHE>
__global__ void __launch_bounds__(256) six_point_kernel(int64_t N, const NodeXf* __restrict__ table,
                                                        const uint32_t* __restrict__ ends_idx,
                                                        const float* __restrict__ ends_w, float* __restrict__ ends,
                                                        const float* __restrict__ scale0,
                                                        const uint8_t* __restrict__ is_static, float pad, float coeff,
                                                        float* pos, float* rot, float* scale, float* sh,
                                                        const ShRotTables* __restrict__ T, SixPointIn in) {
    __shared__ __align__(16) float s_e[256 * 18];
    __shared__ int s_tag[CACHE ? NODE_CACHE_SLOTS : 1];
    __shared__ float4 s_rows[CACHE ? NODE_CACHE_ROW4 * NODE_CACHE_SLOTS : 1];
    const int64_t g0 = (int64_t)blockIdx.x * 256;
    const int64_t npts = 6 * min((int64_t)256, N - g0);
    if constexpr (CACHE) {
        for (int i = threadIdx.x; i < NODE_CACHE_SLOTS; i += 256) s_tag[i] = -1;
        __syncthreads();
    }
#pragma unroll 1
    for (int r = 0; r < 6; r++) {
        const int k = r * 256 + (int)threadIdx.x;
        const bool here = k < npts;
        const int64_t pt = g0 * 6 + k;
        uint32_t id[K];
        float wt[K];
        if (here) load_nbrs<K, VEC>(ends_idx + pt * K, ends_w + pt * K, id, wt);
        if constexpr (CACHE) {
            // claim empty slots, fill the claimed rows, then every lane reads
            if (here) node_cache_claim<K>(table, id, s_tag, s_rows);
            __syncthreads();
        }
        if (here) {
            const float in[3] = {ends[pt * 3], ends[pt * 3 + 1], ends[pt * 3 + 2]};
            float o[3];
            if constexpr (CACHE)
                blend_point_cached<K>(table, s_tag, s_rows, id, wt, in, o);
            else
                blend_point_regs<K>(table, id, wt, in, o);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ends[pt * 3 + c] = o[c];
                s_e[k * 3 + c] = o[c];
            }
        }
    }
    __syncthreads();
    const int64_t g = g0 + threadIdx.x;
    const int n_here = (int)min((int64_t)256, N - g0);
    const bool active = g < N && !(is_static && is_static[g]);
    float e[6][3];
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
        for (int c = 0; c < 3; c++) e[p][c] = s_e[threadIdx.x * 18 + p * 3 + c];
    __syncthreads();  // s_e is reused for the SH rows
    float Rsh[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    if (active) {

    float center[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float sum = 0.0f;
#pragma unroll
        for (int p = 0; p < 6; p++) sum += e[p][c];
        center[c] = sum / 6.0f;
    }
    double M[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) M[c][i] = 0.5 * ((double)e[2 * i][c] - (double)e[2 * i + 1][c]);
    double Rd[3][3], Sd[3];
    polar3_fp64(M, Rd, Sd);
    float Rf[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Rf[i][j] = (float)Rd[i][j];
    float q[4];
    mat_to_quat(Rf, q);
    quat_normalize(q);
    const float4 qo = *reinterpret_cast<const float4*>(in.rot + 4 * g);
    *reinterpret_cast<float4*>(rot + 4 * g) = make_float4(q[0], q[1], q[2], q[3]);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float s0 = scale0[3 * g + i];
        scale[3 * g + i] = (float)Sd[i] / ((s0 + pad) * coeff) * s0;
        pos[3 * g + i] = center[i];
    }
    // relative rotation q * q_old^-1, normalised
    const float n2 = qo.x * qo.x + qo.y * qo.y + qo.z * qo.z + qo.w * qo.w;
    const float qi[4] = {qo.x / n2, -qo.y / n2, -qo.z / n2, -qo.w / n2};
    float qr[4];
    quat_mul(q, qi, qr);
    quat_normalize(qr);
    float Rr[3][3];
    quat_to_mat(qr[0], qr[1], qr[2], qr[3], Rr);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Rsh[i][j] = Rr[i][j];
    }
    if (!active && g < N && in.pos != pos) {
        // a static Gaussian keeps its attributes: carried into the written set
#pragma unroll
        for (int i = 0; i < 3; i++) {
            pos[3 * g + i] = in.pos[3 * g + i];
            scale[3 * g + i] = in.scale[3 * g + i];
        }
        *reinterpret_cast<float4*>(rot + 4 * g) = *reinterpret_cast<const float4*>(in.rot + 4 * g);
    }
    float shc[48];
    rows48_load(in.sh + g0 * 48, n_here, s_e, 256 * 18, shc);
    if (active) sh_rotate_regs(*T, Rsh, shc);
    rows48_store(sh + g0 * 48, n_here, s_e, 256 * 18, shc);
}

// Node rotation -> orthogonal polar factor -> unit quaternion (w,x,y,z).
__global__ void node_quats_kernel(int M, const NodeXf* __restrict__ table, float* __restrict__ q) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= M) return;
    double A[3][3];
#pragma unroll
    for (int i = 0; i < 9; i++) A[i / 3][i % 3] = (double)table[n].R[i];
    double Rd[3][3], Sd[3];
    polar3_fp64(A, Rd, Sd);
    float Rf[3][3];
#pragma unroll
    for (int i = 0; i < 9; i++) Rf[i / 3][i % 3] = (float)Rd[i / 3][i % 3];
    float qq[4];
    mat_to_quat(Rf, qq);
    quat_normalize(qq);
    *reinterpret_cast<float4*>(q + 4 * n) = make_float4(qq[0], qq[1], qq[2], qq[3]);
}

// Running slerp from identity towards each neighbour node's quaternion with
// t = w_j / (w_j + sum of previous weights) (Q_SlerpCUDA, cudakdtree.cu:113-148),
// then SH rotation of the sample's 48-float feature. The features move through
// LDS (rows48_load / rows48_store): with one row per lane, a wave's float4 load
// touches 64 rows 192 B apart and the lines are fetched again for each of the
// row's 12 loads once the L1 has evicted them (bypassing the caches made the
// kernel 2.6x slower); the coalesced form takes 613 instead of 682 us per
// 5.64M samples.
template <int K>
__device__ __forceinline__ void sample_rotation(int64_t s, const float* __restrict__ w, const uint32_t* __restrict__ idx,
                                                const float* __restrict__ node_q, float (&R)[3][3]);

// Five waves per SIMD: the register budget (at most 96) still holds the 48
// features and the band matrices without spills, and the extra wave hides more
// of the gathers' latency (683 vs 712 us per 5.64M samples at four waves).
template <int K>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) sample_sh_kernel(int64_t S, const float* __restrict__ w,
                                                        const uint32_t* __restrict__ idx,
                                                        const float* __restrict__ node_q,
                                                        const uint8_t* __restrict__ is_static,
                                                        float* __restrict__ feat, const ShRotTables* __restrict__ T) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ __align__(16) float lds[256 * 18];
    const int64_t s0 = (int64_t)blockIdx.x * blockDim.x;
    const int n_here = (int)min((int64_t)blockDim.x, S - s0);
    float c[48];
    rows48_load(feat + s0 * 48, n_here, lds, 256 * 18, c);
    if (s < S && !(is_static && is_static[s])) {
        float R[3][3];
        sample_rotation<K>(s, w, idx, node_q, R);
        sh_rotate_regs(*T, R, c);
    }
    rows48_store(feat + s0 * 48, n_here, lds, 256 * 18, c);
}

// Unit quaternion by the hardware reciprocal square root and one Newton step
// (within an ulp or two of q / |q|; the slerp runs it after every neighbour).
__device__ __forceinline__ void quat_normalize_rsq(float q[4]) {
    const float n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (!(n2 > 0.0f)) return;
    float r = __builtin_amdgcn_rsqf(n2);
    r = r * (1.5f - 0.5f * n2 * r * r);
#pragma unroll
    for (int c = 0; c < 4; c++) q[c] *= r;
}

// theta = acos(x) for x in [0, 1]: Abramowitz & Stegun 4.4.46 (Hastings),
// sqrt(1 - x) times a degree-7 polynomial, |error| <= 2e-8 in exact
// arithmetic (about 2.5e-7 evaluated in FP32).
__device__ __forceinline__ float acos_unit(float x) {
    float p = -0.0012624911f;
    p = __builtin_fmaf(p, x, 0.0066700901f);
    p = __builtin_fmaf(p, x, -0.0170881256f);
    p = __builtin_fmaf(p, x, 0.0308918810f);
    p = __builtin_fmaf(p, x, -0.0501743046f);
    p = __builtin_fmaf(p, x, 0.0889789874f);
    p = __builtin_fmaf(p, x, -0.2145988016f);
    p = __builtin_fmaf(p, x, 1.5707963050f);
    return __builtin_sqrtf(1.0f - x) * p;
}

// sin and cos of a in [0, pi/2] by their Taylor series through a^11 / a^12
// (truncation below 6e-8 and 7e-9 at pi/2; about 1.6e-7 evaluated in FP32).
__device__ __forceinline__ void sincos_quadrant(float a, float& sn, float& cs) {
    const float a2 = a * a;
    float ps = -2.5052108e-8f;               // -1/11!
    ps = __builtin_fmaf(ps, a2, 2.7557319e-6f);   // 1/9!
    ps = __builtin_fmaf(ps, a2, -1.9841270e-4f);  // -1/7!
    ps = __builtin_fmaf(ps, a2, 8.3333333e-3f);   // 1/5!
    ps = __builtin_fmaf(ps, a2, -1.6666667e-1f);  // -1/3!
    sn = __builtin_fmaf(a * a2, ps, a);
    float pc = 2.0876757e-9f;                // 1/12!
    pc = __builtin_fmaf(pc, a2, -2.7557319e-7f);  // -1/10!
    pc = __builtin_fmaf(pc, a2, 2.4801587e-5f);   // 1/8!
    pc = __builtin_fmaf(pc, a2, -1.3888889e-3f);  // -1/6!
    pc = __builtin_fmaf(pc, a2, 4.1666667e-2f);   // 1/4!
    pc = __builtin_fmaf(pc, a2, -0.5f);
    cs = __builtin_fmaf(a2, pc, 1.0f);
}

// One step of the running slerp of Q_SlerpCUDA (cudakdtree.cu:113-148): acc
// moves towards e (flipped into acc's hemisphere) by t = cw / (cw + last).
// The first neighbour (last = 0, t = 1) lands on e itself. Otherwise, with
// cos(theta) = cs and s = sin(t theta) / sin(theta), the slerp weights are
// (cos(t theta) - cs s, s): one reciprocal instead of two divisions
// (angle-addition formula). theta = acos(cs) lies in [0, pi/2] after the flip
// and t theta too, so theta and sin/cos(t theta) come from the short
// polynomial forms above instead of the general atan2 and sincos (the weights
// stay within a few 1e-7 of the reference's FP64 ones).
__device__ __forceinline__ void slerp_step(float (&acc)[4], float4 e4, float cw, float last, bool first) {
    float e[4] = {e4.x, e4.y, e4.z, e4.w};
    float cs = acc[1] * e[1] + acc[2] * e[2] + acc[3] * e[3] + acc[0] * e[0];
    if (cs < 0.0f) {
        e[0] = -e[0]; e[1] = -e[1]; e[2] = -e[2]; e[3] = -e[3];
        cs = -cs;
    }
    if (first) {
#pragma unroll
        for (int c = 0; c < 4; c++) acc[c] = e[c];
    } else {
        const float t = cw / (cw + last);
        float ra, rb;
        if (cs > 0.99995f) {
            ra = 1.0f - t;
            rb = t;
        } else {
            const float sn = __builtin_sqrtf(1.0f - cs * cs);
            const float ang = acos_unit(cs);
            float st, ct;
            sincos_quadrant(t * ang, st, ct);
            rb = st * (1.0f / sn);
            ra = ct - cs * rb;
        }
#pragma unroll
        for (int c = 0; c < 4; c++) acc[c] = ra * acc[c] + rb * e[c];
    }
    quat_normalize_rsq(acc);
}

template <int K>
__device__ __forceinline__ void sample_rotation(int64_t s, const float* __restrict__ w, const uint32_t* __restrict__ idx,
                                                const float* __restrict__ node_q, float (&R)[3][3]) {
    float acc[4] = {1.0f, 0.0f, 0.0f, 0.0f};
    float last = 0.0f;
#pragma unroll
    for (int j = 0; j < K; j++) {
        const float cw = w[s * K + j];
        const float4 e4 = *reinterpret_cast<const float4*>(node_q + 4 * (int64_t)idx[s * K + j]);
        slerp_step(acc, e4, cw, last, j == 0);
        last += cw;
    }
    quat_to_mat(acc[0], acc[1], acc[2], acc[3], R);
}

// sample_rotation with the sample's ids and weights already in registers (the
// same arithmetic, operation for operation).
template <int K>
__device__ __forceinline__ void sample_rotation_regs(const uint32_t (&id)[K], const float (&wt)[K],
                                                     const float* __restrict__ node_q, float (&R)[3][3]) {
    float acc[4] = {1.0f, 0.0f, 0.0f, 0.0f};
    float last = 0.0f;
#pragma unroll
    for (int j = 0; j < K; j++) {
        const float cw = wt[j];
        const float4 e4 = *reinterpret_cast<const float4*>(node_q + 4 * (int64_t)id[j]);
        slerp_step(acc, e4, cw, last, j == 0);
        last += cw;
    }
    quat_to_mat(acc[0], acc[1], acc[2], acc[3], R);
}

// Several consecutive drag frames of the sample update in one pass. During a
// drag nothing reads the samples (they feed the density-preserving step after
// the release), so frame f's sample blend (UpdatePositionforSamples) and SH
// rotation (FastUpdateSamplesSH) can wait for later frames: for f = 0..nf-1 in
// order, the position is blended with frame f's node table and the feature
// rotated by frame f's slerped node rotation, each with the per-frame kernels'
// arithmetic, so the result equals nf launches of blend_points_kernel<8, true>
// and sample_sh_kernel<8> bit for bit while every sample's 536 bytes (position,
// ids, weights, feature) move through HBM once instead of nf times.
// compose != 0: the feature is rotated once by R_{nf-1} ... R_1 R_0 instead of
// nf times (the SH rotation is a representation: D(R_b) D(R_a) = D(R_b R_a));
// the positions stay bit-identical, the features agree within float rounding.
template <int K>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
samples_frames_kernel(int64_t S, SampleFrames F, const uint32_t* __restrict__ idx, const float* __restrict__ w,
                      const uint8_t* __restrict__ is_static, float* __restrict__ pts, float* __restrict__ feat,
                      const ShRotTables* __restrict__ T, int compose) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ __align__(16) float lds[256 * 18];
    const int64_t s0 = (int64_t)blockIdx.x * blockDim.x;
    const int n_here = (int)min((int64_t)blockDim.x, S - s0);
    const bool live = s < S && !(is_static && is_static[s]);
    uint32_t id[K];
    float wt[K];
    if (live) {
        load_nbrs<K, true>(idx + s * K, w + s * K, id, wt);
        // positions first, all frames, before the features occupy registers
        float p[3] = {pts[3 * s], pts[3 * s + 1], pts[3 * s + 2]};
#pragma unroll 1
        for (int f = 0; f < F.n; f++) {
            float o[3];
            blend_point_regs<K>(F.table[f], id, wt, p, o);
            p[0] = o[0]; p[1] = o[1]; p[2] = o[2];
        }
        pts[3 * s] = p[0]; pts[3 * s + 1] = p[1]; pts[3 * s + 2] = p[2];
    }
    if (compose) {
        // the product of the frames' rotations before the features occupy registers
        float Rt[3][3];
        if (live) {
            sample_rotation_regs<K>(id, wt, F.q[0], Rt);
#pragma unroll 1
            for (int f = 1; f < F.n; f++) {
                float R[3][3], P[3][3];
                sample_rotation_regs<K>(id, wt, F.q[f], R);
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) P[i][j] = R[i][0] * Rt[0][j] + R[i][1] * Rt[1][j] + R[i][2] * Rt[2][j];
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) Rt[i][j] = P[i][j];
            }
        }
        float c[48];
        rows48_load(feat + s0 * 48, n_here, lds, 256 * 18, c);
        if (live) sh_rotate_regs(*T, Rt, c);
        rows48_store(feat + s0 * 48, n_here, lds, 256 * 18, c);
        return;
    }
    float c[48];
    rows48_load(feat + s0 * 48, n_here, lds, 256 * 18, c);
    if (live) {
#pragma unroll 1
        for (int f = 0; f < F.n; f++) {
            float R[3][3];
            sample_rotation_regs<K>(id, wt, F.q[f], R);
            sh_rotate_regs(*T, R, c);
        }
    }
    rows48_store(feat + s0 * 48, n_here, lds, 256 * 18, c);
}

__global__ void sh_rotate_kernel(int64_t n, const float* __restrict__ R9, float* __restrict__ sh,
                                 const ShRotTables* __restrict__ T) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float R[3][3];
#pragma unroll
    for (int k = 0; k < 9; k++) R[k / 3][k % 3] = R9[9 * i + k];
    sh_rotate_f4(*T, R, sh + 48 * i);
}

__global__ void polar_kernel(int64_t n, const double* __restrict__ M9, double* __restrict__ R9,
                             double* __restrict__ Sd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double M[3][3], R[3][3], S[3];
#pragma unroll
    for (int k = 0; k < 9; k++) M[k / 3][k % 3] = M9[9 * i + k];
    polar3_fp64(M, R, S);
#pragma unroll
    for (int k = 0; k < 9; k++) R9[9 * i + k] = R[k / 3][k % 3];
    Sd[3 * i] = S[0]; Sd[3 * i + 1] = S[1]; Sd[3 * i + 2] = S[2];
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

int launch_end_points(const float* pos, const float* rot, const float* scale, int64_t N, float pad, float coeff,
                      float* ends, hipStream_t st) {
    if (N == 0) return 0;
    hipLaunchKernelGGL(end_points_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, pos, rot, scale, N, pad, coeff, ends);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_build_node_table(int M, const float* node_pos, const float* x_free, const int* free_id, NodeXf* table,
                            hipStream_t st) {
    hipLaunchKernelGGL(node_table_kernel, dim3(nblk(M, 256)), dim3(256), 0, st, M, node_pos, x_free, free_id, table);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_blend_points(const float* pts, int64_t P, const uint32_t* idx, const float* w, int k, const NodeXf* table,
                        int M, const uint8_t* skip, float* out, hipStream_t st, bool node_cache) {
    (void)M;
    if (P == 0) return 0;
    const bool vec = ((uintptr_t)idx % 16 == 0) && ((uintptr_t)w % 16 == 0);
    if (k == 8 && vec && node_cache)
        hipLaunchKernelGGL((blend_points_kernel<8, true, true>), dim3(nblk(P, 256)), dim3(256), 0, st, pts, P, idx, w,
                           table, skip, out);
    else if (k == 8 && vec)
        hipLaunchKernelGGL((blend_points_kernel<8, true, false>), dim3(nblk(P, 256)), dim3(256), 0, st, pts, P, idx, w,
                           table, skip, out);
    else if (k == 8)
        hipLaunchKernelGGL((blend_points_kernel<8, false, false>), dim3(nblk(P, 256)), dim3(256), 0, st, pts, P, idx,
                           w, table, skip, out);
    else
        hipLaunchKernelGGL(blend_points_generic_kernel, dim3(nblk(P, 256)), dim3(256), 0, st, pts, P, idx, w, k,
                           table, skip, out);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_six_point(int64_t N, int k, const NodeXf* table, int M, const uint32_t* ends_idx, const float* ends_w,
                     float* ends, const float* scale0, const uint8_t* is_static, float pad, float coeff, float* pos,
                     float* rot, float* scale, float* sh, const ShRotTables* T, hipStream_t st, bool node_cache,
                     const SixPointIn* in_p) {
    (void)M;
    if (N == 0) return 0;
    const bool vec = ((uintptr_t)ends_idx % 16 == 0) && ((uintptr_t)ends_w % 16 == 0);
    const SixPointIn in = in_p ? *in_p : SixPointIn{pos, rot, scale, sh};
    if (node_cache && k == 8 && vec) {
        hipLaunchKernelGGL((six_point_kernel<8, true, true>), dim3(nblk(N, 256)), dim3(256), 0, st, N, table, ends_idx,
                           ends_w, ends, scale0, is_static, pad, coeff, pos, rot, scale, sh, T, in);
        HIP_RET(hipGetLastError());
        return 0;
    }
    switch (k) {
#define ARAPGS_SP_CASE(KV)                                                                                             \
    case KV:                                                                                                           \
        if (vec)                                                                                                       \
            hipLaunchKernelGGL((six_point_kernel<KV, true, false>), dim3(nblk(N, 256)), dim3(256), 0, st, N, table,    \
                               ends_idx, ends_w, ends, scale0, is_static, pad, coeff, pos, rot, scale, sh, T, in);     \
        else                                                                                                           \
            hipLaunchKernelGGL((six_point_kernel<KV, false, false>), dim3(nblk(N, 256)), dim3(256), 0, st, N, table,   \
                               ends_idx, ends_w, ends, scale0, is_static, pad, coeff, pos, rot, scale, sh, T, in);     \
        break;
        ARAPGS_SP_CASE(1) ARAPGS_SP_CASE(2) ARAPGS_SP_CASE(3) ARAPGS_SP_CASE(4) ARAPGS_SP_CASE(5) ARAPGS_SP_CASE(6)
        ARAPGS_SP_CASE(7) ARAPGS_SP_CASE(8) ARAPGS_SP_CASE(9) ARAPGS_SP_CASE(10) ARAPGS_SP_CASE(11) ARAPGS_SP_CASE(12)
#undef ARAPGS_SP_CASE
        default:
            return -1;
    }
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_node_quats(int M, const NodeXf* table, float* q, hipStream_t st) {
    hipLaunchKernelGGL(node_quats_kernel, dim3(nblk(M, 128)), dim3(128), 0, st, M, table, q);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_sample_sh(int64_t S, int k, const float* w, const uint32_t* idx, const float* node_q,
                     const uint8_t* is_static, float* feat, const ShRotTables* T, hipStream_t st) {
    if (S == 0) return 0;
    switch (k) {
#define ARAPGS_SS_CASE(KV)                                                                                        \
    case KV:                                                                                                      \
        hipLaunchKernelGGL(sample_sh_kernel<KV>, dim3(nblk(S, 256)), dim3(256), 0, st, S, w, idx, node_q, is_static, \
                           feat, T);                                                                              \
        break;
        ARAPGS_SS_CASE(1) ARAPGS_SS_CASE(2) ARAPGS_SS_CASE(3) ARAPGS_SS_CASE(4) ARAPGS_SS_CASE(5) ARAPGS_SS_CASE(6)
        ARAPGS_SS_CASE(7) ARAPGS_SS_CASE(8) ARAPGS_SS_CASE(9) ARAPGS_SS_CASE(10) ARAPGS_SS_CASE(11) ARAPGS_SS_CASE(12)
#undef ARAPGS_SS_CASE
        default:
            return -1;
    }
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_samples_frames(int64_t S, int k, const SampleFrames& F, const uint32_t* idx, const float* w,
                          const uint8_t* is_static, float* pts, float* feat, const ShRotTables* T, hipStream_t st,
                          bool compose) {
    if (S == 0 || F.n == 0) return 0;
    if (k != 8 || F.n < 0 || F.n > SAMPLE_FRAMES_MAX || ((uintptr_t)idx % 16) || ((uintptr_t)w % 16)) return -1;
    hipLaunchKernelGGL(samples_frames_kernel<8>, dim3(nblk(S, 256)), dim3(256), 0, st, S, F, idx, w, is_static, pts,
                       feat, T, compose ? 1 : 0);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_sh_rotate(int64_t n, const float* R9, float* sh, const ShRotTables* T, hipStream_t st) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(sh_rotate_kernel, dim3(nblk(n, 128)), dim3(128), 0, st, n, R9, sh, T);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_polar(int64_t n, const double* M9, double* R9, double* Sd, hipStream_t st) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(polar_kernel, dim3(nblk(n, 128)), dim3(128), 0, st, n, M9, R9, Sd);
    HIP_RET(hipGetLastError());
    return 0;
}

}  // namespace arapgs
