// Gauss-Newton solver state (device side) for solver.hip / solver_host.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <array>
#include <map>
#include <vector>

namespace arapgs {

// Everything the residual / Jacobian kernels read. Plain device pointers,
// passed to kernels by value.
struct SolverDev {
    int M, K;
    int nF, nS, nC;        // free nodes, static nodes, constraint groups
    int64_t n_groups;      // nF + nF*K + nS*K + nC
    int n_agg, nc_pad;     // aggregates, padded coarse dimension (multiple of 64)
    float sw_rot, sw_reg, sw_con;
    const int* free_nodes;     // nF   node id of each free rank (ascending)
    const int* static_nodes;   // nS
    const int* free_id;        // M    rank among free nodes or -1
    const int* static_id;      // M    rank among static nodes or -1
    const uint32_t* nbr;       // M*K  graph neighbours
    const float* edge_scale;   // M*K  weighting factor of (i, t) when i and k share a non-excluded block
    const uint32_t* src_idx;   // M*K  KNN of each node's source vertex
    const float* src_w;        // M*K
    // Constraint group c: rows sw_con * sum over members m of (blend(m) - centre_c),
    // centre_c = (float) sum of the aims of its aim nodes / their count. Plain
    // mode: one member, aim node = the member. Centre mode: the ascending key
    // nodes of a block and its first min(20, |block|) nodes.
    const int* con_ptr;        // nC+1 members of each group
    const int* con_mem;
    const int* caim_ptr;       // nC+1 aim nodes of each group
    const int* caim_node;
    const int* in_ptr;         // nF+1 in-edges (src*K + t) of each free node
    const int* in_edge;
    const int* cm_ptr;         // nF+1 constraint memberships of each free node, one per group
    const int* cm_con;         // group of each membership
    const int* cm_sbeg;        // (memberships+1) range of the membership in cm_src
    const int* cm_src;         // source-weight entries (member*K + slot) that reach the free node
    const int* multi_con;      // n_multi groups with more than one member
    int n_multi;
    // Walk tables of the CG operator: the index chains of the owner walk
    // resolved once (static: neighbour free ids, member ranges) and the
    // frame's geometry per record (rebuilt by walk_table_kernel every solve).
    int n_mem;                 // constraint members (con_mem entries)
    int n_in, n_cm;            // in-edges of free nodes, memberships
    const int2* wt_out_ik;     // nF*K  (node i, neighbour k) of out-edge (free rank, t)
    const int2* wt_in_is;      // in-edges: (owner node i, source node)
    const int* wt_cm_i;        // memberships: owner node i
    const int* wt_out_nb;      // nF*K  free id of out-edge t's neighbour (-1: static)
    const int* wt_in_nb;       // in-edges: free id of the source (-1: static)
    const int2* wt_cm_rng;     // memberships: the group's member range [con_ptr[c], con_ptr[c+1])
    const int* wt_con_nb;      // n_mem*K  free id of each member's source slot (-1: static)
    float4* wt_out;            // nF*K  (g_k - g_i, edge scale) of out-edge (i, t)
    float4* wt_in;             // in-edges: (g_i - g_src, edge scale)
    float4* wt_cm;             // memberships: the group's I3 block (a, b) wrt the node
    float4* wt_con;            // n_mem*K  (g_member - g_src, weight) of each source slot
    // Assembled CG operator (plain constraint mode): J^T J = blockdiag(D) +
    // sum over free pairs (i, j), i != j, of (M_ij (x) I_3), M_ij a 4x4 matrix
    // over (R column 0, 1, 2, t). Pair p of owner i: j = pair_j[p], the terms
    // [term_ptr[p], term_ptr[p+1]) each add u_i u_j^T for one residual group
    // shared by i and j (int4: type, node, slot, 0; type 0: edge (node, slot)
    // = (i, t) with nbr = j; type 1: edge (j, t) with nbr = i; type 2:
    // constraint group `node`). n_pairs = 0 when not built (centre mode).
    const int* pair_ptr;       // nF+1
    const int* pair_j;         // n_pairs
    const int* term_ptr;       // n_pairs+1
    const int4* terms;
    int n_pairs;
    const int* agg_ptr;        // n_agg+1 free ranks grouped by aggregate
    const int* agg_member;
    const int* agg_of;         // nF
    float* agg_center;         // n_agg*3 (per frame)
    const float* g;            // M*3 current node positions (per frame)
    const float* aim;          // M*3 aims (per frame)
};

struct CgScalars {
    double rz, alpha, beta, rr, bb, tol2;
    double rr_stop;   // max(tol^2 |b|^2, tol_floor^2 |b_0|^2), b_0 = the solve's first gradient
    int iters, max_iters;
};

struct GnScalars {
    double e_old;
    double bb0;       // |b|^2 of the solve's first Gauss-Newton iteration
    int iters_done;
    int cg_total;     // CG iterations of the solve so far
};

// Counter, generation and error flag of the persistent PCG kernel's grid barrier.
struct GridBarrier {
    unsigned count;
    unsigned gen;
    int err;
    int pad;
};

// Largest Gauss-Newton iteration count of one solve (energies are kept for
// iterations 0..GN_MAX_ITERS on the device).
constexpr int GN_MAX_ITERS = 64;

// One scene of a solver launch: everything the solver kernels read or write,
// in device memory. Every solver kernel takes an array of entries and serves
// entry blockIdx.z, so a batch of independent scenes runs each phase of the
// solve in one launch; a single solve is a batch of one. Grid sizes are per
// entry (blocks past an entry's count return at once), so each scene's blocks
// do exactly the work of a launch for that scene alone.
struct SolverEntry {
    SolverDev S;
    int n, nb_owner, e_blocks, acc_blocks, n_cand, nt;  // nt = nc_pad / 64 coarse tiles
    int nb_apply;      // workgroups of the CG apply (cg_block / 16 owners each)
    int cg_max_iters;
    double tol2;
    float *x, *hsol, *r, *z, *p, *q, *binv, *rc, *acinv, *gj_piv, *gj_col, *hwarm;
    float* dmat;       // nF*144: the diagonal blocks D of J^T J (assembled operator)
    float* pair_m;     // n_pairs*16: M_ij of the assembled operator (per frame)
    int cg_asm;        // the CG operator is applied assembled
    unsigned long long* ac64;
    double *partial, *partial2, *rzp, *e_partial, *cand_e, *acc_partial;
    double* energies;  // GN_MAX_ITERS + 1: E(x) before and after every Gauss-Newton iteration
    CgScalars* cg;
    GnScalars* gs;
    int* gn_done;
    int* cg_done;
    const GridBarrier* gbar;
};

constexpr int SOLVER_BATCH_MAX = 64;

// Per-scene outputs of a batched solve (kernel argument, by value).
struct SolverOutPtrs {
    float* x[SOLVER_BATCH_MAX];
    double* energies[SOLVER_BATCH_MAX];
    int* iters_done[SOLVER_BATCH_MAX];
    void* table[SOLVER_BATCH_MAX];  // NodeXf rows of the scene's nodes (putFreeInputs), or null
};

struct SolverWork {
    SolverDev dev;
    int n;  // 12 * nF
    float *x, *hsol, *r, *z, *p, *q, *binv, *rc, *acinv, *gj_piv, *gj_col;
    float* dmat = nullptr;     // nF*144 diagonal blocks of J^T J (assembled operator)
    float* pair_m = nullptr;   // n_pairs*16 off-diagonal 4x4 coefficients (per frame)
    bool cg_asm = false;       // apply the CG operator assembled (needs dev.n_pairs > 0)
    unsigned long long* ac64;  // fixed-point coarse assembly (order independent)
    double *partial, *partial2, *rzp, *e_partial, *cand_e;  // rzp: 2 x n_agg r.z partials
    int e_blocks, n_cand;
    double* acc_partial;  // 2 x acc_blocks: |x|^2, |h|^2 partials of the step acceptance
    int acc_blocks;
    // Warm start: the CG of GN iteration k starts from the step of GN iteration
    // k of the previous solve (hwarm[k], zero before the first), when warm != 0.
    float* hwarm;
    int warm;
    int warm_slots;
    CgScalars* cg;
    GnScalars* gs;
    int* gn_done;
    int* cg_done;
    int cg_max_iters;
    double cg_tol;
    int cg_block = 256;  // workgroup size of the CG apply: 64, 128 or 256 (16 lanes per owner node)
    void* gbar;        // grid barrier of the persistent PCG kernel
    int coop_grid;     // its grid size (0: not yet chosen)
    double* energies;  // GN_MAX_ITERS + 1
    // Asynchronous coarse refresh: a rebuild of the coarse space (aggregate
    // centres and the inverse Galerkin matrix) runs on a side stream from a
    // snapshot of the node positions into the *_alt buffers; the next frame's
    // setup waits for built_ev and copies them over the current ones.
    float* acinv_alt = nullptr;
    float* agg_center_alt = nullptr;
    float* g_snap = nullptr;   // M*3
    float* x_ident = nullptr;  // n: identity rotations, zero translations (x0 of the Galerkin operator)
    hipEvent_t built_ev = nullptr;
    bool pending = false;      // a side-stream rebuild waits to be adopted
    int pending_age = 0;       // solves set up since that rebuild started
    bool have_coarse = false;  // a coarse inverse exists (the first one is built in line)
    std::vector<void*> allocs;  // owned device buffers
};

// A launch group of solvers (one entry per scene, in device memory). The
// per-frame node positions and aims live in the entries and are rewritten by
// a small kernel only when a caller passes different pointers.
struct SolverBatch {
    std::vector<SolverWork*> w;
    SolverEntry* entries = nullptr;   // device, w.size()
    SolverEntry* entries_build = nullptr;  // device: the entries of a side-stream coarse rebuild
    hipStream_t side = nullptr;       // side stream of the asynchronous coarse refresh
    hipStream_t side_ext = nullptr;   // a caller's stream used instead of it (not owned)
    hipEvent_t fork_ev = nullptr;
    bool async = true;                // coarse refreshes on the side stream (the first one in line)
    int lag = 1;                      // a side-stream rebuild is adopted by the lag-th solve after the one
                                      // that started it; no new rebuild starts while one is pending
    std::vector<const float*> g, aim; // what the device entries hold
    // Launch sequences replayed as captured HIP graphs: the CG chunks (CG
    // iterations [k0, k0 + L) and the flag kernel) and the Gauss-Newton
    // iteration's begin and end phases, each cached under the arguments it was
    // captured with (a graph replays its kernels' arguments as captured).
    bool graphs = false;
    // CG iterations as two kernels (restriction by recurrence) instead of three
    bool fused = false;
    std::map<std::array<int64_t, 6>, hipGraphExec_t> graph_cache;
    int kt = 0;                       // 8 when every scene has K == 8 (unrolled walks), else 0
    int max_wt_blocks = 0;  // walk_table_kernel workgroups (largest entry)
    int max_nb_owner = 0, max_nb_apply = 0, max_agg = 0, max_e_blocks = 0, max_acc_blocks = 0, max_n = 0, max_groups128 = 0,
        max_multi = 0, max_nt = 0, max_cand = 0;
};

int solver_batch_init(SolverBatch& B, const std::vector<SolverWork*>& w, std::string& err);
void solver_batch_free(SolverBatch& B);
// Points the entries at this frame's node positions and aims (one launch when any changed).
int solver_batch_set_frame(SolverBatch& B, const float* const* g, const float* const* aim, hipStream_t st);
// *flag_h (pinned host memory) = AND over the scenes of cg_done (which 0) or
// gn_done (which 1), written by a kernel queued on st. In the two-kernel CG
// form (B.fused) the kernel first runs the convergence test after k_done CG
// iterations. iters_h (which 0, when not null): the largest CG iteration count
// over the scenes so far.
// flag_h = 1 when every scene's CG (which 0) or Gauss-Newton loop (which 1) is
// done; iters_h: the largest CG count (which 0); gn_h (when not null):
// gn_h[0] = every Gauss-Newton loop done, gn_h[1] = the largest GN count.
int solver_batch_flag(SolverBatch& B, int which, hipStream_t st, int* flag_h, int k_done = 0, int* iters_h = nullptr,
                      int* gn_h = nullptr);
// Queues CG iterations [k0, k0 + L) followed by the cg_done flag kernel
// writing *flag_h (and the CG count into *iters_h) as one graph launch,
// capturing it on st the first time (st must not be the null stream).
int solver_cg_chunk(SolverBatch& B, int k0, int L, hipStream_t st, int* flag_h, int* iters_h);
// solver_gn_begin / solver_gn_end + the gn_done flag kernel writing *flag_h,
// as graph launches (captured on first use; st must not be the null stream).
int solver_gn_begin_graph(SolverBatch& B, int iter, hipStream_t st);
int solver_gn_end_graph(SolverBatch& B, int iter, int stop_rule, hipStream_t st, int* flag_h);
void solver_batch_drop_graphs(SolverBatch& B);
int solver_batch_finish(SolverBatch& B, const SolverOutPtrs& out, hipStream_t st);
// The warm-start branch is taken from the first scene; a batch needs them all alike.
inline bool solver_batch_warm_uniform(const SolverBatch& B) {
    for (const SolverWork* W : B.w)
        if ((W->warm != 0) != (B.w[0]->warm != 0)) return false;
    return true;
}

struct SolverDesc {
    int M, K;
    const uint32_t* nbr;
    const int* block_ptr;
    const int* block_nodes;
    const int* block_type;
    int nblocks;
    const uint32_t* src_idx;
    const float* src_w;
    const float* node_pos;
    float w_rot, w_reg, w_con, wf;
    int center;
    int n_agg;
};

// Host-side setup; returns 0 or a negative error with a message in err.
int solver_build(const SolverDesc& d, SolverWork& W, std::string& err);
void solver_free(SolverWork& W);

int solver_frame_setup(SolverBatch& B, bool refresh_coarse, hipStream_t st);
int solver_gn_begin(SolverBatch& B, int iter, hipStream_t st);
// CG iteration k: three kernels, or two when B.fused.
int solver_cg_iteration(SolverBatch& B, int k, hipStream_t st);
// The whole PCG solve of one GN iteration as one cooperative launch (after
// solver_gn_begin); same arithmetic and stopping rule as solver_cg_iteration.
int solver_cg_persistent(SolverWork& W, hipStream_t st);
// Reads the persistent kernel's barrier error flag (synchronous).
int solver_barrier_error(SolverWork& W, int* err);
int solver_gn_end(SolverBatch& B, int iter, int stop_rule, hipStream_t st);
// Single-scene helpers (B holds one solver).
int solver_energy(SolverBatch& B, const float* x, double* out, hipStream_t st);
// Test hook: q = J^T J v and b = -J^T f at x (the CG operator and right-hand side).
int solver_apply_test(SolverBatch& B, const float* x, const float* v, float* q_out, float* b_out, hipStream_t st);

}  // namespace arapgs
