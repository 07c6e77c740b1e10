/*
 * arapgs — MI355X (gfx950) per-frame "ARAP solve -> deform Gaussians -> render"
 * path behind a plain C ABI.
 *
 * Conventions
 *   - One context per GPU; all work is queued on the context's stream
 *     (arapgs_set_stream) and is asynchronous unless stated otherwise.
 *   - Pointer arguments are DEVICE pointers unless the name ends in `_h`.
 *   - Every call returns 0 on success, a negative code on error; the message is
 *     available from arapgs_last_error(ctx). Handles are not thread-safe.
 *   - Gaussian attributes use the reference's activated SoA layout
 *     (GaussianView.cpp:43-153 after activation): pos f32[N*3], rot f32[N*4]
 *     as (w,x,y,z), scale f32[N*3], opacity f32[N], sh f32[N*48] with
 *     coefficient-major interleaving (coef*3 + channel).
 *   - K (nearest nodes per point) is 1..12 (KNN_MAX, helper.hpp:48).
 *   - Steady-state per-frame calls do not allocate device memory once a call
 *     of the same size has run (workspaces grow on first use).
 *   - Solvers and grids keep their own device buffers and may be destroyed
 *     before or after the context that created them; calls that take both need
 *     both alive.
 */
#ifndef ARAPGS_H
#define ARAPGS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct arapgs_ctx arapgs_ctx;
typedef struct arapgs_solver arapgs_solver;

/* Per-node deformation record (64 bytes): current node position g, rotation R
 * row-major, translation t. Replaces DeformGraph::rot / trans / nodes[].Position
 * (Deform.hpp:16-38) in device memory. */
typedef struct arapgs_node_xf {
    float g[3];
    float R[9];
    float t[3];
    float pad;
} arapgs_node_xf;

/* ------------------------------------------------------------ context */
int arapgs_create(int device, arapgs_ctx** out);
void arapgs_destroy(arapgs_ctx* ctx);
const char* arapgs_last_error(const arapgs_ctx* ctx);
/* hip_stream is a hipStream_t (NULL = the null stream). */
int arapgs_set_stream(arapgs_ctx* ctx, void* hip_stream);
int arapgs_synchronize(arapgs_ctx* ctx);
/* A HIP stream on `device` whose kernels run only on the compute units
 * [cu_begin, cu_begin + cu_count) (hipExtStreamCreateWithCUMask); cu_count = 0:
 * an ordinary non-blocking stream. Lets a latency-bound stage (the solve) keep
 * CUs of its own while a bandwidth-bound stage fills the rest. No counterpart
 * in the reference. Destroy with arapgs_stream_destroy. */
int arapgs_stream_create(int device, int cu_begin, int cu_count, void** stream_out);
/* The same with an explicit CU mask: bit i of mask_h (words 32-bit words)
 * enables compute unit i (the order of hipExtStreamCreateWithCUMask). */
int arapgs_stream_create_mask(int device, const uint32_t* mask_h, int words, void** stream_out);
int arapgs_stream_destroy(void* stream);
/* Library version string (static storage). */
const char* arapgs_version(void);

/* ------------------------------------------------------------ graph build */
/* farthest_control_points_sampling + FetchFirstNodeIdx (helper.cpp:139-195).
 * out_idx: device int[m+1]; *count_h receives the number written (min(m,N),
 * or 2 in the reference's m==1 corner case). Synchronous. */
int arapgs_fps(arapgs_ctx* ctx, const float* pos, int64_t n, int m, int* out_idx, int* count_h);

/* DeformGraph::computeWeights / findNearestNodes (Deform.hpp:153-208) for Q
 * query points against M rest node positions, with the reference's exact
 * selection-sort tie semantics. out_idx u32[Q*(k+1)] (slot k is the normaliser
 * node), out_w f32[Q*k] (may be NULL): (1 - d_j/d_k)^2 normalised, computed in
 * FP64 and rounded. Requires M >= k+1. */
int arapgs_knn(arapgs_ctx* ctx, const float* nodes, int m, const float* queries, int64_t q, int k,
               uint32_t* out_idx, float* out_w);

/* No counterpart in the reference: arapgs_knn for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). Same arguments and bit-identical results; it calls no hipMalloc,
 * hipFree or allocating routine. Its scratch (the fallback count and list) is
 * a static device array of 32 MB per GPU, held until the process ends, so q is
 * at most 2^23 (8,388,608) queries; larger q is rejected with an error. Every
 * call on one GPU shares that GPU's scratch (the count is cleared by a kernel
 * at the start of each call): calls that overlap on one GPU, on different
 * streams or threads, race. */
int arapgs_knn_noalloc(arapgs_ctx* ctx, const float* nodes, int m, const float* queries, int64_t q, int k,
                       uint32_t* out_idx, float* out_w);

/* DeformGraph::setupEdges (Deform.hpp:84-95): out_nbr u32[M*k]. */
int arapgs_graph_edges(arapgs_ctx* ctx, const float* nodes, int m, int k, uint32_t* out_nbr);

/* GaussianView::GetEndPoints (GaussianView.cpp:4643-4668): ends f32[N*18] as
 * [e0+, e0-, e1+, e1-, e2+, e2-] per Gaussian. */
int arapgs_end_points(arapgs_ctx* ctx, const float* pos, const float* rot, const float* scale, int64_t n,
                      float axis_padding, float end_coeff, float* ends);

/* ------------------------------------------------------------ ARAP solve */
/* Replaces the Deform constructor (Deform.hpp:414-455) for one block layout:
 * host arrays describing the graph and the edit blocks. block_type: -1
 * excluded/static, 0 pinned, 1 active (GaussianView.cpp:1995-2021). src_idx /
 * src_w (M*k) are the KNN of each node's source vertex
 * (cand_vertices[nodes[i].Vertex_index]). node_pos_h seeds the coarse
 * aggregation of the preconditioner. */
typedef struct arapgs_solver_desc {
    int m, k;
    const uint32_t* nbr_h;
    const int* block_ptr_h;      /* nblocks+1 */
    const int* block_nodes_h;
    const int* block_type_h;     /* nblocks */
    int nblocks;
    const uint32_t* src_idx_h;
    const float* src_w_h;
    const float* node_pos_h;
    float w_rot, w_reg, w_con;   /* DeformGraph weights (Deform.hpp:54-59): 1, 10, 100 */
    float weighting_factor;      /* GaussianView::_weighting_factor (GaussianView.hpp:459) */
    int constraints_on_center;   /* Deform::_constraints_on_center: one 3-row constraint per control
                                    block, on the centre of its first min(20, |block|) nodes
                                    (Deform.cpp:34-75, 483-529) */
    int n_aggregates;            /* coarse aggregates, 0 = 128 */
    int cg_max_iters;            /* 0 = 100 */
    float cg_tol;                /* relative residual, 0 = 1e-4 */
} arapgs_solver_desc;

int arapgs_solver_create(arapgs_ctx* ctx, const arapgs_solver_desc* desc_h, arapgs_solver** out);
void arapgs_solver_destroy(arapgs_solver* s);
/* n_free_h: free nodes; n_unknowns_h = 12 * n_free. */
int arapgs_solver_dims(const arapgs_solver* s, int* n_free_h, int* n_unknowns_h);

/* Deform::real_time_deform + optimize (Deform.cpp:77-169): GN from identity,
 * `iters` iterations at most, stop rule |h| < (|x|+1e-6)1e-6 when stop_rule.
 * node_pos f32[M*3] current node positions, aims f32[M*3].
 * x_out f32[12*n_free] (optional), energies f64[iters+1] (energy at the start
 * of each iteration; entry iters_done is the final energy), iters_done int[1].
 * All outputs are device buffers; nothing is read back. iters_done is -1 when
 * the persistent PCG kernel's grid barrier gave up waiting (solve invalid). */
int arapgs_gn_solve(arapgs_ctx* ctx, arapgs_solver* s, const float* node_pos, const float* aims, int iters,
                    int stop_rule, float* x_out, double* energies, int* iters_done);

/* Convergence checks on the host every `every` CG iterations and after every
 * GN iteration (synchronising; the loop then stops early). 0 = none: the
 * whole solve is queued with device-side early exit (graph-capturable). */
int arapgs_solver_set_host_checks(arapgs_solver* s, int every);

/* The coarse level of the preconditioner (Galerkin operator of the node
 * aggregates and its inverse) is rebuilt on the first solve and then every
 * `frames` solves (default 8); in between it is reused. It only affects the
 * CG iteration count, never the solution. */
int arapgs_solver_set_coarse_refresh(arapgs_solver* s, int frames);
/* on = 1 (default): every rebuild after the first runs on a side stream of the
 * solver from a snapshot of that solve's node positions, beside the solve, and
 * the next solve adopts it (waiting for it if needed); the solve that starts
 * it keeps the previous operator. on = 0: rebuilds run in line on the
 * context's stream before the solve that needs them. Either way the sequence
 * of operators is a function of the solve count only, so the results do not
 * depend on the streams or the host. */
int arapgs_solver_set_async_refresh(arapgs_solver* s, int on);
/* With the asynchronous refresh: a side-stream rebuild is adopted by the
 * solves-th solve after the one that started it (1..8, default 1), so a rebuild
 * that takes longer than a frame does not hold up the next solve; no new
 * rebuild starts while one is waiting. The operator sequence stays a function
 * of the solve count (deterministic). No counterpart in the reference. */
int arapgs_solver_set_coarse_lag(arapgs_solver* s, int solves);
/* The stream the asynchronous coarse rebuilds run on: a caller's stream (e.g.
 * one restricted to some compute units; it must outlive its use by the solver)
 * or null for the solver's own. Synchronises the device. No counterpart in
 * the reference. */
int arapgs_solver_set_side_stream(arapgs_solver* s, void* stream);
/* The CG block size (no counterpart in the reference; set by the autotuner):
 * workgroup size of the CG operator kernel, 64, 128, 256 or 512 threads (16
 * lanes per free node, so 4 to 32 nodes per workgroup; default 256); in the
 * two-kernel CG form (cg_mode 2) the workgroup size of both kernels (one
 * workgroup per aggregate, BS / 16 member nodes per pass). It changes the
 * order of the p.q partial sums only: the solve stops at the same tolerance.
 * Synchronises the device. Solvers of one batch need the same value. */
int arapgs_solver_set_cg_block(arapgs_solver* s, int threads);

/* on = 1: the CG operator J^T J is applied assembled instead of matrix-free:
 * per frame, every pair (i, j) of free nodes that share a residual group gets
 * its 4x4 coefficient matrix M_ij (J^T J = blockdiag(D) + sum M_ij (x) I_3,
 * the Jacobian blocks being I_3-structured), and the operator kernel reads
 * M_ij and v_j instead of re-deriving every group's Jacobian from the graph.
 * Same operator up to rounding (the solve stops at the same tolerance). Plain
 * constraint mode only (-2 in centre mode, whose multi-member groups couple
 * every pair of their sources). Synchronises the device. No counterpart in
 * the reference (its J^T J is assembled for a sparse Cholesky). Default 0. */
int arapgs_solver_set_cg_assembled(arapgs_solver* s, int on);

/* on = 1: the CG of GN iteration k starts from the step that GN iteration k of
 * the previous solve on this solver found (zero before the first), instead of
 * from zero. Consecutive drag frames solve nearly the same systems, so the CG
 * starts closer to its tolerance; the tolerance and stopping rule are
 * unchanged. Default 0 (every solve independent of the ones before). Every
 * call clears the stored steps (synchronous). */
int arapgs_solver_set_warm_start(arapgs_solver* s, int on);

/* Returns the solver's frame-to-frame state to that of a new solver: the
 * warm-start steps are zero, the next solve builds the coarse operator in line
 * (also with the asynchronous refresh on; a side-stream rebuild in flight is
 * waited for and dropped) and the coarse-refresh count restarts. The reference
 * builds a new Deform every drag frame (GaussianView.cpp:1481); this solver
 * keeps that state across frames as an optimisation, and a caller resets it
 * where a new drag starts or where two runs must compute the same frames.
 * Synchronises the device. */
int arapgs_solver_reset(arapgs_solver* s);

/* on = 1: with host checks every `every` CG iterations (arapgs_solver_set_host_checks
 * > 0) on a non-null stream, the solve's launch sequences replay as captured HIP
 * graphs (each captured on first use under the arguments it was captured with,
 * and kept): every Gauss-Newton iteration's begin phase (right-hand side,
 * preconditioner, CG start) and end phase (line search, acceptance, gn_done
 * flag), and the CG in chunks (the predicted count of set_cg_predict, or
 * `every`, then `every` iterations) with their convergence flag; the host
 * reads chunk j-1's flag while chunk j is already queued, so the stream does
 * not wait for the host between chunks, and a chunk queued past convergence
 * exits kernel by kernel on the device flag. The kernels, their arguments and
 * their order are those of the launch loop, so the iterates are identical bit
 * for bit. No counterpart in the reference (its solver is a CPU Cholesky).
 * Default 0. */
int arapgs_solver_set_cg_graphs(arapgs_solver* s, int on);

/* GN iterations and total CG iterations of the solver's last solve
 * (synchronises the device). Diagnostics; no counterpart in the reference. */
int arapgs_solver_last_stats(arapgs_solver* s, int* gn_iters_h, int* cg_iters_h);

/* How each GN iteration's PCG solve (the dense Eigen solve of Deform.cpp:118-128
 * in the reference) is launched: 0 (default) = three kernels per CG iteration
 * (combine with set_host_checks to bound the launches queued past convergence);
 * 1 = one persistent cooperative kernel per solve with grid barriers between
 * the CG phases (the same iterates as 0 bit for bit; on MI355X a grid barrier
 * costs more than a kernel boundary); 2 = two kernels per CG iteration: the
 * coarse restriction is carried by the recurrence rc_{k+1} = rc_k - alpha P^T q_k
 * (P^T q formed in the operator kernel, one workgroup per aggregate), so the
 * update and the coarse correction share one launch. Mode 2 is the same
 * preconditioned CG up to rounding (same tolerance and stopping rule); its
 * iterates do not depend on the host-check interval or on CG graphs. */
int arapgs_solver_set_cg_mode(arapgs_solver* s, int mode);

/* on = 1: the host checks of arapgs_gn_solve poll the flag's pinned slot,
 * which the flag kernel publishes with a system-scope release, instead of
 * waiting for the stream (a slot still unset after 50 ms falls back to the
 * stream wait). Iterates are unchanged; only the host's wake-up is faster.
 * No counterpart in the reference (its solve runs on the CPU). */
int arapgs_solver_set_host_spin(arapgs_solver* s, int on);

/* The CG's relative residual tolerance after creation (desc.cg_tol; 0 < tol
 * < 1). The reference solves each Gauss-Newton step exactly (sparse
 * Cholesky); a tolerance is an inexact-Newton choice whose accuracy is judged
 * by the final energies against an exact FP64 solve. Synchronises the device. */
int arapgs_solver_set_cg_tol(arapgs_solver* s, float tol);

/* on = 1: with host checks on (set_host_checks > 0), the first host check of
 * GN iteration k's CG comes after as many CG iterations as GN iteration k of
 * the previous solve needed, then every set_host_checks iterations (no
 * counterpart in the reference; host scheduling only: the device-side
 * convergence test decides, so the results are the same bit for bit).
 * Consecutive drag frames need similar counts, so fewer launches run past
 * convergence. With the CG as separate launches (cg_mode 0 or 2) the
 * Gauss-Newton stop flag is predicted too: before the previous solve's GN count, the next GN
 * iteration is queued without waiting for the flag (its kernels do nothing
 * once the loop has stopped) and the flag is read at that iteration's first
 * CG check. arapgs_solver_reset forgets the counts. */
int arapgs_solver_set_cg_predict(arapgs_solver* s, int on);

/* Energy |f(x)|^2 (FP64 reduction) of a given unknown vector (Deform::CalcEnergyFunc,
 * Deform.cpp:378-581). energy: device f64[1]. */
int arapgs_gn_energy(arapgs_ctx* ctx, arapgs_solver* s, const float* node_pos, const float* aims, const float* x,
                     double* energy);

/* DeformGraph::putFreeInputs (Deform.hpp:140-151): node table from the solved
 * unknowns (identity for non-free nodes). */
int arapgs_put_free_inputs(arapgs_ctx* ctx, arapgs_solver* s, const float* node_pos, const float* x,
                           arapgs_node_xf* table);

/* ------------------------------------------------------------ scene batches
 * Independent scenes of uneven size solved together: every phase of the
 * Gauss-Newton solve (frame set-up, right-hand side, each CG phase, line
 * search, step acceptance) is ONE launch for all scenes of the batch, scene =
 * blockIdx.z, each scene's blocks doing exactly the work of a solve of that
 * scene alone. No counterpart in the reference (it edits one scene); per scene
 * the result is bit-identical to arapgs_gn_solve + arapgs_put_free_inputs on
 * the same solver. A batch holds 1..64 solvers created on ctx's device, each
 * at most once, sharing their state (warm-start steps, coarse operator) with
 * single solves on them; all need the same warm-start setting. The batch
 * refers to the solvers' device buffers: destroy it before any of its
 * solvers. Host convergence checks default to every 8 CG iterations (on the
 * AND of the scenes' flags), the coarse refresh to 8 batch solves. */
typedef struct arapgs_solver_batch arapgs_solver_batch;
int arapgs_solver_batch_create(arapgs_ctx* ctx, arapgs_solver* const* solvers_h, int n, arapgs_solver_batch** out);
void arapgs_solver_batch_destroy(arapgs_solver_batch* b);
int arapgs_solver_batch_set_host_checks(arapgs_solver_batch* b, int every);
int arapgs_solver_batch_set_coarse_refresh(arapgs_solver_batch* b, int frames);
/* arapgs_solver_set_cg_graphs for a batch's solve. */
int arapgs_solver_batch_set_cg_graphs(arapgs_solver_batch* b, int on);
int arapgs_solver_batch_set_async_refresh(arapgs_solver_batch* b, int on);
/* Host arrays of n device pointers, one per scene in batch order: node_pos and
 * aims f32[M_s*3] (required); x_out f32[12*n_free_s], energies f64[iters+1],
 * iters_done int[1] and tables (putFreeInputs node table, M_s rows) are
 * optional (a null array or a null entry skips that output). */
int arapgs_solver_batch_solve(arapgs_ctx* ctx, arapgs_solver_batch* b, const float* const* node_pos_h,
                              const float* const* aims_h, int iters, int stop_rule, float* const* x_out_h,
                              double* const* energies_h, int* const* iters_done_h, arapgs_node_xf* const* tables_h);

/* ------------------------------------------------------------ deformation */
/* DeformGraph::predict_mesh / predict_samples (Deform.hpp:230-268):
 * out[i] = sum_j w_j (R_j (p_i - g_j) + t_j + g_j); skip (u8[P], optional)
 * copies p_i unchanged. Used for samples (UpdatePositionforSamples,
 * GaussianView.cpp:3060-3077) and nodes (GaussianView.cpp:3046-3056). */
int arapgs_blend_points(arapgs_ctx* ctx, const float* pts, int64_t p, const uint32_t* idx, const float* w, int k,
                        const arapgs_node_xf* table, int m, const uint8_t* skip, float* out);

/* UpdatePosition's end-point blend + UpdateAsSixPointsWithdrawBad
 * (GaussianView.cpp:3021-3043, 3081-3166), fused: ends (N*18) are blended in
 * place with ends_idx u32[N*6*k] / ends_w f32[N*6*k]; for non-static Gaussians
 * pos, rot, scale and sh are rewritten (polar in FP64). scale0 = backup scales. */
int arapgs_six_point_update(arapgs_ctx* ctx, int64_t n, int k, const arapgs_node_xf* table, int m,
                            const uint32_t* ends_idx, const float* ends_w, float* ends, const float* scale0,
                            const uint8_t* is_static, float axis_padding, float end_coeff, float* pos, float* rot,
                            float* scale, float* sh);

/* The same update reading the attributes from one set (pos_in, rot_in,
 * scale_in, sh_in) and writing them to another (pos, rot, scale, sh; static
 * Gaussians' attributes are copied over), so that a render may still read the
 * previous frame's set meanwhile (double-buffered Gaussians); ends are blended
 * in place. Results bit-identical to the in-place update. No counterpart in the
 * reference (its update is in place). */
int arapgs_six_point_update_from(arapgs_ctx* ctx, int64_t n, int k, const arapgs_node_xf* table, int m,
                                 const uint32_t* ends_idx, const float* ends_w, float* ends, const float* scale0,
                                 const uint8_t* is_static, float axis_padding, float end_coeff, const float* pos_in,
                                 const float* rot_in, const float* scale_in, const float* sh_in, float* pos,
                                 float* rot, float* scale, float* sh);

/* FastUpdateSamplesSH (GaussianView.cpp:3169-3186): node quaternions
 * q f32[M*4] (w,x,y,z) from the table's rotations (orthogonal polar factor). */
int arapgs_node_quats(arapgs_ctx* ctx, const arapgs_node_xf* table, int m, float* q);

/* RotateSamplesSHsCUDA / RotateSHs (cudakdtree.cuh:25, cudakdtree.cu:201-222):
 * per non-static sample, running slerp of node quaternions, then rotation of
 * the 48-float feature in place. */
int arapgs_sample_sh(arapgs_ctx* ctx, int64_t s, int k, const float* w, const uint32_t* idx, const float* node_q,
                     const uint8_t* is_static, float* feat);

/* Batching form with no single counterpart in the reference: nf (<= 8)
 * consecutive drag frames of the sample update (UpdatePositionforSamples, then
 * FastUpdateSamplesSH, GaussianView.cpp:1500-1522) in one pass. For f = 0..nf-1
 * in order the samples f32[s*3] are blended with tables_h[f] (like
 * arapgs_blend_points) and the features f32[s*48] rotated with node_q_h[f]
 * (like arapgs_sample_sh): bit-identical to those 2*nf calls. tables_h /
 * node_q_h are host arrays of nf device pointers; k must be 8 and idx, w, feat
 * 16-byte aligned. */
int arapgs_samples_frames(arapgs_ctx* ctx, int64_t s, int k, int nf, const arapgs_node_xf* const* tables_h,
                          const float* const* node_q_h, const uint32_t* idx, const float* w, const uint8_t* is_static,
                          float* samples, float* feat);
/* on = 1: arapgs_samples_frames on this context rotates each feature once, by
 * the product R_{nf-1} ... R_1 R_0 of the frames' slerped rotations, instead of
 * nf times (the SH rotation is a group representation, D(R_b) D(R_a) =
 * D(R_b R_a)): positions stay bit-identical, features agree with the per-frame
 * calls within float rounding. No counterpart in the reference. Default 0. */
int arapgs_set_sh_compose(arapgs_ctx* ctx, int on);

/* ------------------------------------------------------------ grid + density */
typedef struct arapgs_grid arapgs_grid;

/* getOverallAABB (GaussianView.cpp:3601-3631) + the grid step of
 * getGridSamples (:3873-3877): out7_h (host) = {min.xyz, max.xyz, step} for a
 * grid_num^3 grid. Synchronous. */
int arapgs_scene_aabb(arapgs_ctx* ctx, const float* pos, int64_t n, int grid_num, float* out7_h);

/* The cell reorder of getGridSamples (GaussianView.cpp:3889-3953): order u32[n]
 * with order[new] = old, a stable sort of the Gaussians by centre cell. */
int arapgs_grid_reorder(arapgs_ctx* ctx, const float* pos, int64_t n, const float* aabb7_h, int grid_num,
                        uint32_t* order);

/* Per-Gaussian alpha-cutoff AABBs (GaussianView.cpp:3960-4019): f32[n*6]. */
int arapgs_gaussian_aabbs(arapgs_ctx* ctx, const float* pos, const float* rot, const float* scale,
                          const float* opacity, int64_t n, float* aabbs);

/* Cell lists + valid cells + 4^3 samples per valid cell (GaussianView.cpp:4021-4133,
 * GetBoxesGsGrid cudakdtree.cu:425-451). The grid owns its device buffers.
 * Synchronous (list and valid-cell counts are data dependent). The lists are
 * built in chunks of whole x-slabs of cells (at most 2^28 cell-Gaussian pairs
 * per chunk, 16 temporary bytes per pair) and addressed with 64-bit offsets,
 * so the pair total is not limited to 2^31 (the reference's int prefix sums
 * are); one slab may hold at most 2^31-1 pairs and n at most 2^31-1. */
int arapgs_grid_build(arapgs_ctx* ctx, const float* aabbs, int64_t n, const float* aabb7_h, int grid_num,
                      int padding, arapgs_grid** out);
void arapgs_grid_destroy(arapgs_grid* grid);
/* UpdateContainingRelationship (GaussianView.cpp:3634-3744): new cell lists
 * (prefix and list) from the Gaussians' current boxes `aabbs` (n must equal
 * the grid's Gaussian count) on the geometry aabb7_h (arapgs_scene_aabb of the
 * current positions); the grid's resolution and padding, its valid cells and
 * their samples are kept. Synchronous; views from arapgs_grid_info before the
 * call are invalid after it. */
int arapgs_grid_rebin(arapgs_ctx* ctx, arapgs_grid* grid, const float* aabbs, int64_t n, const float* aabb7_h);
/* Test hook with no counterpart in the reference: the number of cell-Gaussian
 * pairs per chunk of later arapgs_grid_build calls on this context (0: the
 * default 2^28). The lists do not depend on it. */
int arapgs_set_grid_chunk(arapgs_ctx* ctx, int64_t pairs);
/* Tuning switch with no counterpart in the reference: the K = 8 blends (the
 * six-point update's end points, arapgs_blend_points, the frame's sample and
 * node blends) read node rows through a per-workgroup LDS cache (1, default)
 * or gather every row from the global table (0). Results are identical bit
 * for bit. */
int arapgs_set_node_cache(arapgs_ctx* ctx, int on);
/* Device views of a grid: prefix int64[G^3] (inclusive per-cell counts, the
 * reference's grid_gs_prefix_sum widened to 64 bits), list int[n_pairs]
 * (grided_gs_idx), valid int[n_valid] (valid_grid_idx), samples
 * f32[n_valid*64*3] (sample_positions). */
int arapgs_grid_info(const arapgs_grid* grid, int64_t* n_pairs, int64_t* n_valid, const int64_t** prefix,
                     const int** list, const int** valid, const float** samples);

/* GetAdaLpfRatio (GaussianView.cpp:4670-4749, on by default: GaussianView.hpp:180;
 * called at F6, :1703-1705): the adaptive low-pass filter of every valid cell.
 * M is the linear map that best carries the unit cube's corners onto the
 * cell's 8 corner samples of `samples` (device f32[n_valid*64*3], NULL = the
 * grid's own samples; taken relative to the cell's first sample and divided by
 * the 3 sample intervals), in the least-squares sense of the reference's 24 x 9
 * system, and the cell's filter is lpf_param * M M^T (at rest, lpf_param times
 * the squared sample interval, times I). While set, arapgs_density_sample,
 * arapgs_density_backward, the realignment and the chain on this grid add the
 * cell's filter to each Gaussian's covariance in place of lpf I (their lpf
 * argument is then unused). lpf_param = 0 switches back to lpf I. */
int arapgs_grid_set_ada_lpf(arapgs_ctx* ctx, arapgs_grid* grid, const float* samples, float lpf_param);
/* Device view of the adaptive filters, f32[n_valid*6] (xx, yy, zz, xy, xz, yz
 * per valid cell), or NULL while the grid uses lpf I. */
int arapgs_grid_ada_lpf(const arapgs_grid* grid, const float** mats);

/* Density sample of every grid sample (the forward3d_grid pass,
 * GaussianView.cpp:4159-4186; the reference kernel is not available, so the
 * formula is this library's own):
 * a_g(x) = o_g exp(-1/2 d^T (Sigma_g + L)^-1 d) over the cell's list, L = lpf I
 * or the cell's adaptive filter (arapgs_grid_set_ada_lpf),
 * feat f32[S*48] = sum a_g sh_g, opac f32[S] = sum a_g. */
int arapgs_density_sample(arapgs_ctx* ctx, const arapgs_grid* grid, const float* pos, const float* rot,
                          const float* scale, const float* opacity, const float* sh, float lpf, float* feat,
                          float* opac);

/* The vector-Jacobian product of arapgs_density_sample (the role of the
 * un-vendored backward3d_grid, called at GaussianView.cpp:4415-4447) at sample
 * positions `samples` (device f32[S*3]; NULL = the grid's own samples), with
 * the grid's cell lists held fixed: given dL/dfeat (f32[S*48], may be NULL) and
 * dL/dopac (f32[S], may be NULL), writes dL/d of every Gaussian's pos (N*3),
 * rot (N*4, with respect to the stored, unnormalised quaternion), scale (N*3,
 * activated), opacity (N, activated) and sh (N*48); dsamples (S*3) optional.
 * Float atomics accumulate over cells, so the last bits vary between runs. */
int arapgs_density_backward(arapgs_ctx* ctx, const arapgs_grid* grid, const float* samples, const float* pos,
                            const float* rot, const float* scale, const float* opacity, const float* sh, float lpf,
                            const float* dfeat, const float* dopac, float* dpos, float* drot, float* dscale,
                            float* dopacity, float* dsh, float* dsamples);

/* Density-preserving realignment (the F6 loop of GaussianView.cpp:1695-1733,
 * GPUOptimize at :4363-4486: forward3d_grid, L1loss3d, backward3d_grid,
 * update3d). The Gaussians' raw parameters — pos, rot (stored quaternion), log
 * scales, logit opacities, SH — are moved by Adam so that the density sample
 * at the (deformed) samples matches the aim features / densities in L1:
 * L = sum |feat - aim_feat| + w_density sum |dens - aim_dens|. The step uses
 * the grid's current cell lists (arapgs_grid_rebin replaces them: the
 * reference re-bins every _containing_relation_interval steps,
 * GaussianView.cpp:1718-1721). The reference's kernels are not vendored: the
 * loss, the masks and the convergence rule below are this library's
 * definitions. ExpandOpRange and the adaptive low-pass are not modelled. */
typedef struct arapgs_realign_desc {
    float lr_pos, lr_rot, lr_scale, lr_opacity, lr_sh;  /* Adam step sizes (raw parameters) */
    float beta1, beta2, eps;                             /* 0 = 0.9, 0.999, 1e-15 */
    float w_density;                                     /* weight of the density term */
} arapgs_realign_desc;
typedef struct arapgs_realign arapgs_realign;
int arapgs_realign_create(arapgs_ctx* ctx, const arapgs_grid* grid, int64_t n, const arapgs_realign_desc* desc_h,
                          arapgs_realign** out);
void arapgs_realign_destroy(arapgs_realign* r);
/* One Adam step; loss (device f64[1], optional) receives L before the step.
 * samples f32[S*3], aim_feat f32[S*48], aim_dens f32[S] (S = 64 * valid cells). */
int arapgs_realign_step(arapgs_ctx* ctx, arapgs_realign* r, const float* samples, const float* aim_feat,
                        const float* aim_dens, float lpf, float* pos, float* rot, float* log_scale,
                        float* logit_opacity, float* sh, double* loss);
/* Options of later steps (defaults: all NULL / 0):
 *  moved      device u8[n]: only Gaussians with moved[g] != 0 are updated ("Only
 *             Optimize Deformed Region", the moved_gaussians of CheckMovedGaussians,
 *             GaussianView.cpp:2092-2109, passed to update3d at :4452-4478), and
 *             valid cells whose lists hold no moved Gaussian are static (the
 *             current_static_grids of :3850-3852): they pass no gradient;
 *  max_scale  device f32[n*3]: after each step the activated scales are clamped
 *             to it (the max_scale argument of update3d);
 *  conv_tol   per-cell convergence (grid_is_converged of L1loss3d, :4191-4207):
 *             the first step after creation or new options records every cell's
 *             L1 loss, and a cell whose loss later falls to conv_tol times its
 *             record is converged and passes no gradient; the flag is sticky,
 *             also across re-binnings (the reference clears it only with new
 *             features, :4154) (0: off).
 * The reported loss always covers every cell. */
typedef struct arapgs_realign_options {
    const uint8_t* moved;
    const float* max_scale;
    float conv_tol;
} arapgs_realign_options;
int arapgs_realign_set_options(arapgs_ctx* ctx, arapgs_realign* r, const arapgs_realign_options* opts_h);
/* ExpandOpRange (GaussianView.cpp:4320-4360, run before an F6 pass in high
 * quality mode, :1699-1701): every converged valid cell that is, or is a face
 * neighbour (clamped to the grid) of, a non-converged valid cell is no longer
 * converged, so the next steps optimise a margin around the cells still
 * moving. Uses a G^3-byte buffer allocated on the first call. */
int arapgs_realign_expand(arapgs_ctx* ctx, arapgs_realign* r);
/* Device views of the per-valid-cell state after the last step: skip u8[n_valid]
 * (static or converged) and the cell's L1 loss f32[n_valid]. */
int arapgs_realign_cells(const arapgs_realign* r, int64_t* n_valid, const uint8_t** skip, const float** cell_loss);

/* ------------------------------------------------------------ render */
/* CudaRasterizer::Rasterizer::forward as called at GaussianView.cpp:1099-1128
 * (positional order of that call): P Gaussians, SH degree D (M = 16
 * coefficients per Gaussian), background bg (device f32[3]), W x H image,
 * means3D, shs, opacities, scales (x scale_modifier), rotations (w,x,y,z),
 * viewmatrix / projmatrix (device f32[16], column-major; projmatrix is the full
 * world->clip transform), cam_pos (device f32[3]), tan_fovx/y.
 * out_color: device f32[3*H*W], planar CHW (shaders/copy.frag:38-40).
 * surface (device int[P], optional): set to 1 for every Gaussian that
 * contributed to a pixel. radii (device int[P], optional). rects (device
 * int[2*P], optional): when given, tile rectangles use per-axis 3-sigma extents
 * (the "Fast culling" mode, GaussianView.cpp:1062) and the extents are written.
 * boxmin_h / boxmax_h (host float[3], optional): crop box on the means.
 * colors_precomp, cov3D_precomp must be NULL and prefiltered 0. to_ortho = 1
 * renders through an orthographic camera (GaussianView's setOrthoCam(ortho_scale,
 * ortho_scale), GaussianView.cpp:1018-1019, 1094-1095): projmatrix is that
 * camera's world->clip matrix (w = 1) and the EWA Jacobian is the constant
 * diag(W, H) / (2 ortho_scale); ortho_scale > 0. tile: 16 (the 3DGS tile) or 8. Returns the number of (tile, Gaussian)
 * instances in *num_rendered_h. Synchronises once (instance count). */
typedef struct arapgs_render_desc {
    int P, D, M;
    const float* background;
    int width, height;
    const float* means3D;
    const float* shs;
    const float* colors_precomp;
    const float* opacities;
    const float* scales;
    float scale_modifier;
    const float* rotations;
    const float* cov3D_precomp;
    const float* viewmatrix;
    const float* projmatrix;
    const float* cam_pos;
    float tan_fovx, tan_fovy;
    int prefiltered;
    float* out_color;
    int* surface;
    int* radii;
    int* rects;
    const float* boxmin_h;
    const float* boxmax_h;
    int to_ortho;
    float ortho_scale;
    int tile;
} arapgs_render_desc;

int arapgs_render(arapgs_ctx* ctx, const arapgs_render_desc* desc_h, int* num_rendered_h);

/* The same forward pass for n_views cameras of one Gaussian set in one go
 * (GaussianView renders its views one Rasterizer::forward call at a time):
 * descs_h[v] differ only in viewmatrix, projmatrix, cam_pos, tan_fovx/y,
 * to_ortho, ortho_scale, out_color, surface, radii and rects; every other field
 * must equal view 0's.
 * Each Gaussian's SH is read once for all views, the (tile, depth) keys of all
 * views (combined tile id view * ntiles + tile) are sorted together and one
 * blend launch covers every tile of every view, so the deepest tiles of the
 * views overlap. Images are identical bit for bit to n_views separate
 * arapgs_render calls. n_views <= 64; *num_rendered_h = total instances. */
int arapgs_render_views(arapgs_ctx* ctx, const arapgs_render_desc* descs_h, int n_views, int* num_rendered_h);

/* Independent scenes stored back to back, one view each, in one pass (no
 * counterpart in the reference, which renders one scene): view v draws only
 * Gaussians [seg_h[v], seg_h[v+1]) (host int64[n_views+1], ascending, from 0
 * to P). The descs follow arapgs_render_views' rules; a view's surface, radii
 * and rects are indexed within its segment. One record per Gaussian, keys of
 * all views sorted together, one blend launch for every tile of every view;
 * each image is bit-identical to arapgs_render of that segment alone. No
 * backward. */
int arapgs_render_segments(arapgs_ctx* ctx, const arapgs_render_desc* descs_h, int n_views, const int64_t* seg_h,
                           int* num_rendered_h);

/* Tuning of the 16x16 blend: how many Gaussians (2, 4 or 8; default 4) have
 * their alphas evaluated together before being composited in order. Images do
 * not depend on it. No counterpart in the reference (set by the autotuner). */
int arapgs_set_blend_unroll(arapgs_ctx* ctx, int unroll);
/* Tuning of the 16x16 blend with no counterpart in the reference: a tile whose
 * list holds more than `entries` instances (a multiple of 256; 0 = never, the
 * default) is blended in segments of `entries` by separate workgroups, each
 * from transmittance 1, and the segments' (colour, transmittance) are then
 * composited front to back per pixel. The deepest tiles otherwise set the
 * blend's length. Tiles of one segment are unchanged bit for bit; in a split
 * tile the segment where a pixel's transmittance falls below 1e-4 is blended
 * again from the true incoming transmittance, so the contributors are those of
 * the tile-by-tile blend and only the association of the sums and products
 * differs (float rounding). Renders that ask for `surface` and the adjoint
 * chain always blend tile by tile. */
int arapgs_set_blend_split(arapgs_ctx* ctx, int entries);

/* on = 1: a render's read of its instance count (the one host read of a
 * render) polls the pinned slot, which the count kernel publishes with a
 * system-scope release, instead of waiting for the stream (bounded at 50 ms,
 * then the stream wait). Images unchanged. No counterpart in the reference. */
int arapgs_set_host_spin(arapgs_ctx* ctx, int on);

/* on = 1: the context's renders emit, for a Gaussian whose tile rectangle spans
 * at most 64 tiles, only the tiles whose box of pixel centres passes the
 * conservative alpha test the blend applies to every 8x8 quadrant (the
 * Gaussian cannot reach alpha 1/255 there): the sort, range and blend work on
 * shorter lists and the images, contributor marks and gradients are identical
 * bit for bit; the instance keys and values are the subsequence of the full
 * binning's that passes the test. No counterpart in the reference (its
 * binning emits the whole rectangle). Default 0. */
int arapgs_set_tile_cull(arapgs_ctx* ctx, int on);

/* Debug view of the last arapgs_render on this context: sorted keys
 * (u64, tile << 32 | depth bits) and values (Gaussian ids) copied to host
 * buffers of capacity cap; tile ranges (u32 pairs [start,end)) to ranges_h
 * (capacity in tiles). Synchronous. */
int arapgs_render_debug(arapgs_ctx* ctx, uint64_t* keys_h, uint32_t* vals_h, int64_t cap, uint32_t* ranges_h,
                        int64_t ranges_cap);

/* ------------------------------------------------------------ one frame */
/* Device state of a deforming scene (all device pointers). */
typedef struct arapgs_frame_buffers {
    int64_t n;               /* Gaussians */
    int k, m;                /* neighbours per point, nodes */
    float* pos;              /* N*3  (pos_vector) */
    float* rot;              /* N*4  (rot_vector, w,x,y,z) */
    float* scale;            /* N*3  (scale_vector) */
    float* sh;               /* N*48 (shs_vector) */
    const float* scale0;     /* N*3  (scale_backup_vector) */
    float* ends;             /* N*18 (ends_vector) */
    const uint32_t* ends_idx;/* N*6*k (ends_vertices[].Neighbor_Nodes) */
    const float* ends_w;     /* N*6*k */
    const uint8_t* gs_static;/* N, optional (static_gaussians) */
    float* node_pos;         /* M*3  (deform_graph.nodes[].Position) */
    const uint32_t* node_src_idx; /* M*k (cand_vertices[Vertex_index]) */
    const float* node_src_w;
    int64_t s;               /* samples (0 = none) */
    float* samples;          /* S*3  (sample_positions) */
    const uint32_t* sample_idx;   /* S*k */
    const float* sample_w;        /* S*k */
    const uint8_t* sample_static; /* S, optional (static_samples) */
    float* sample_feat;      /* S*48 (aim_feature_cuda), optional */
    float axis_padding, end_coeff;   /* 1e-3, 2 (GaussianView.hpp:116-117) */
} arapgs_frame_buffers;

/* One drag frame of GaussianView::onUpdate (GaussianView.cpp:1481-1522) after
 * the aims are set: Deform + real_time_deform, putFreeInputs,
 * UpdatePositionforSamples, UpdatePosition (end points + nodes),
 * UpdateAsSixPointsWithdrawBad with FastUpdateSamplesSH. */
int arapgs_deform_frame(arapgs_ctx* ctx, arapgs_solver* s, const arapgs_frame_buffers* fb_h, const float* aims,
                        int gn_iters, int stop_rule, double* energies, int* iters_done);

/* UpdateAimPosition (GaussianView.cpp:2920-2934) after ReloadAimPositions:
 * aims = node_pos, then aims[active] += delta_h (host float[3]). */
int arapgs_aims_translate(arapgs_ctx* ctx, const float* node_pos, int m, const int* active, int n_active,
                          const float* delta_h, float* aims);
/* UpdateAimPositionTwist (GaussianView.cpp:2936-2962): rotate the active aims
 * by `radian` about axis_h (host float[3]) through the active nodes' centroid. */
int arapgs_aims_twist(arapgs_ctx* ctx, const float* node_pos, int m, const int* active, int n_active,
                      const float* axis_h, float radian, float* aims);
/* UpdateAimPositionScale (GaussianView.cpp:2961-2983): aims = node_pos, then
 * the active aims scaled by s about the active nodes' centroid. */
int arapgs_aims_scale(arapgs_ctx* ctx, const float* node_pos, int m, const int* active, int n_active, float s,
                      float* aims);
/* Aims of a DeformScript step (GaussianView.cpp:1921-1935): aims = node_pos,
 * then aims[idx[j]] = vals[j*3 .. j*3+3) in order (device idx int[n], vals
 * f32[n*3]; ids outside [0, m) are skipped). */
int arapgs_aims_set(arapgs_ctx* ctx, const float* node_pos, int m, const int* idx, int n, const float* vals,
                    float* aims);
/* CheckStaticSamples (GaussianView.cpp:2023-2090): out[r] = 1 when every one of
 * the len node ids idx[r*len .. r*len+len) is static (node_static u8[m]).
 * Samples: len = k (static_samples); Gaussians: the 6k ids of their end points
 * (static_gaussians). All device pointers. */
int arapgs_static_rows(arapgs_ctx* ctx, const uint8_t* node_static, const uint32_t* idx, int64_t n, int len,
                       uint8_t* out);

/* Measurement hook (no counterpart in the reference): when enabled,
 * arapgs_deform_frame brackets its fused six-point kernel (which = 0), the
 * sample blend (1) and the sample SH rotation (2) with HIP events on the
 * context stream; arapgs_profile_read_kernel synchronises, returns that
 * kernel's per-launch durations (ms) and resets its list; arapgs_profile_read
 * is which = 0. Return the count written (>= 0) or a negative error. */
int arapgs_profile_enable(arapgs_ctx* ctx, int on);
int arapgs_profile_read(arapgs_ctx* ctx, float* ms_h, int cap);
int arapgs_profile_read_kernel(arapgs_ctx* ctx, int which, float* ms_h, int cap);

/* ------------------------------------------------------------ adjoint chain */
/* The differentiable deform -> density-sample -> render chain from rest
 * state: end points of every Gaussian (GetEndPoints, GaussianView.cpp:
 * 4643-4668) from its rest rotation q0 and scale s0, the K-node blend of end
 * points and samples (predict_mesh / predict_samples, Deform.hpp:230-268), the
 * six-point update (UpdateAsSixPointsWithdrawBad, GaussianView.cpp:3081-3166,
 * with q_old = q0 and the backup scales = s0), the renders of n_views cameras
 * (Rasterizer::forward as in arapgs_render, 16-pixel tiles, no fast culling)
 * and the density sample of the deformed Gaussians at the blended samples of a
 * grid (arapgs_density_sample's formula over the grid's fixed cell lists).
 * Inputs differentiated: the node transforms (R, t of an arapgs_node_xf table;
 * g is a constant), q0 and s0. arapgs_chain_backward takes the gradients of a
 * scalar loss with respect to the outputs and returns those of the inputs; it
 * recomputes the forward (including every tile's blend state) instead of
 * keeping it. Replaces the Jacobians the reference obtains by finite
 * differences of its CUDA forward for the density-preserving step
 * (L1loss3d / backward3d_grid / update3d, GaussianView.cpp:1695-1733). */
typedef struct arapgs_chain arapgs_chain;
typedef struct arapgs_chain_desc {
    int64_t n;                    /* Gaussians */
    int k, m;                     /* neighbours per point, nodes */
    const float* pos0;            /* N*3 rest means */
    const float* sh0;             /* N*48 rest SH */
    const float* opacity;         /* N activated opacities */
    const uint32_t* ends_idx;     /* N*6*k blend ids of the end points */
    const float* ends_w;          /* N*6*k */
    const arapgs_grid* grid;      /* density grid (NULL: no density term); S = 64 * valid cells */
    const uint32_t* sample_idx;   /* S*k blend ids of the grid's samples */
    const float* sample_w;        /* S*k */
    float lpf;                    /* density low-pass variance */
    int n_views, width, height, sh_degree;
    const float* background;      /* device f32[3] */
    const float* const* viewmatrix_h;  /* host arrays of n_views device pointers */
    const float* const* projmatrix_h;
    const float* const* cam_pos_h;
    const float* tan_fovx_h;      /* host arrays of n_views */
    const float* tan_fovy_h;
    float axis_padding, end_coeff;  /* 1e-3, 2 */
} arapgs_chain_desc;

/* Validates the description and builds the node -> (end point | sample, slot)
 * inverse index used by the backward (a stable sort by node id). The device
 * arrays of the description must outlive the chain. */
int arapgs_chain_create(arapgs_ctx* ctx, const arapgs_chain_desc* desc_h, arapgs_chain** out);
void arapgs_chain_destroy(arapgs_chain* ch);
/* images_h: host array of n_views device pointers (f32[3*H*W] each), feat
 * f32[S*48] and dens f32[S] (both NULL when the chain has no grid). */
int arapgs_chain_forward(arapgs_ctx* ctx, arapgs_chain* ch, const arapgs_node_xf* table, const float* rot0,
                         const float* scale0, float* const* images_h, float* feat, float* dens);
/* dimages_h: host array of n_views device pointers (dL/d image) or NULL;
 * dfeat / ddens may be NULL. Outputs (all written): dtable (R and t
 * gradients in the table layout, g and pad 0), drot0 f32[N*4], dscale0
 * f32[N*3]. Float atomics accumulate over tiles and cells, so the last bits
 * vary from run to run. */
int arapgs_chain_backward(arapgs_ctx* ctx, arapgs_chain* ch, const arapgs_node_xf* table, const float* rot0,
                          const float* scale0, const float* const* dimages_h, const float* dfeat,
                          const float* ddens, arapgs_node_xf* dtable, float* drot0, float* dscale0);
/* Device views of the deformed state of the last forward or backward call:
 * pos f32[N*3], rot f32[N*4], scale f32[N*3], sh f32[N*48], samples f32[S*3]. */
int arapgs_chain_state(const arapgs_chain* ch, const float** pos, const float** rot, const float** scale,
                       const float** sh, const float** samples);

/* ------------------------------------------------------------ PLY codec */
/* The 3DGS PLY of loadPly / savePly (GaussianView.cpp:43-339): binary little-
 * endian, 62 float properties per vertex (x y z nx ny nz f_dc_0..2
 * f_rest_0..44 opacity scale_0..2 rot_0..3), or 63 with a trailing `index`
 * (the RichPoint_Soup variant). Header check only, no GPU: */
int arapgs_ply_info(const char* path, int64_t* count_h, int* has_index_h);
/* Reads the file into device SoA with the reference's activations (unit
 * quaternion, exp(scale), sigmoid(opacity), SH interleaved coefficient * 3 +
 * channel). morton != 0 reorders by the 3D Morton code of the positions over
 * their bounding box (21 bits per axis, as the reference; ties keep file order,
 * where the reference's std::sort leaves them unspecified) and returns the box
 * in minmax6_h (host, optional). index (device int[count], optional): the
 * `index` property, or each vertex's position in the file. Buffers sized for
 * arapgs_ply_info's count. Synchronous. */
int arapgs_ply_load(arapgs_ctx* ctx, const char* path, int morton, float* pos, float* rot, float* scale,
                    float* opacity, float* sh, int* index, float* minmax6_h);
/* Writes the Gaussians inside the crop box (host float[3] each, optional) and
 * not flagged in too_small (device u8[n], optional), in their order, with the
 * inverse activations (log(scale), logit(opacity)) and normals 0.
 * *written_h receives the vertex count. Synchronous. */
int arapgs_ply_save(arapgs_ctx* ctx, const char* path, int64_t n, const float* pos, const float* rot,
                    const float* scale, const float* opacity, const float* sh, const float* boxmin_h,
                    const float* boxmax_h, const uint8_t* too_small, int64_t* written_h);

/* ------------------------------------------------------------ test hooks */
/* The solver's CG operator and right-hand side at unknowns x (f32[n]):
 * q = J^T J v and b = -J^T f (device f32[n] each), J and f those of
 * Deform::FastCalcJacobiMat / CalcEnergyFunc (Deform.cpp:180-581). */
int arapgs_solver_apply(arapgs_ctx* ctx, arapgs_solver* s, const float* node_pos, const float* aims, const float* x,
                        const float* v, float* q, float* b);
/* SH rotation f'(d) = f(R^T d) of n 48-float sets (helper.cpp:938-1075 with the
 * GaussianView.cpp:3134-3156 sign wrapper); R9 f32[n*9] row-major. */
int arapgs_sh_rotate(arapgs_ctx* ctx, int64_t n, const float* r9, float* sh);
/* Polar decomposition M = R S in FP64 (helper.cpp:429-440): m9, r9 f64[n*9]
 * row-major, sdiag f64[n*3] = diag(S). */
int arapgs_polar(arapgs_ctx* ctx, int64_t n, const double* m9, double* r9, double* sdiag);
/* Host-side evaluation (no GPU, no context) of the blend's conservative
 * Gaussian-vs-tile test: keep[i] = 0 only if Gaussian i (xy f32[n*2], conic
 * f32[n*4] = a, b, c, opacity) has alpha < 1/255 at every pixel centre of
 * box f32[n*4] = x0, y0, x1, y1. Skipping such pairs leaves the composite of
 * Rasterizer::forward (un-vendored XinhaoT/CudaRasterizer@96ea96c, called at
 * GaussianView.cpp:1099-1128) unchanged. */
int arapgs_tile_test_host(int64_t n, const float* xy, const float* conic, const float* box, uint8_t* keep);

#ifdef __cplusplus
}
#endif

#endif /* ARAPGS_H */
