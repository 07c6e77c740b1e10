// C-ABI of the arapgs library (include/arapgs.h): context, workspaces, argument
// checks and dispatch to the gfx950 kernels.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/arapgs.h"
#include "adjoint.hpp"
#include "common.hpp"
#include "grid.hpp"
#include "ply.hpp"
#include "raster.hpp"
#include "solver.hpp"
#include "sort.hpp"

using namespace arapgs;

// byte offset (per instance) of the sorted record ids in arapgs_ctx::r_bin
constexpr size_t BIN_VALS_AT = 8;

struct arapgs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    ShRotTables shrot;
    ShRotTables* shrot_dev = nullptr;  // device copy read by the SH kernels
    // growable workspaces
    void* knn_list = nullptr;
    size_t knn_list_cap = 0;
    uint32_t* knn_count = nullptr;
    void* fps_dist = nullptr;
    size_t fps_dist_cap = 0;
    unsigned long long* fps_partial = nullptr;
    unsigned int* fps_counter = nullptr;
    unsigned long long* fps_best = nullptr;
    uint32_t* edge_tmp = nullptr;
    size_t edge_tmp_cap = 0;
    // rasteriser workspaces
    void* r_geo = nullptr;
    size_t r_geo_cap = 0;
    void* r_scan = nullptr;
    size_t r_scan_cap = 0;
    void* r_bin = nullptr;          // keys, vals, keys_alt, vals_alt
    size_t r_bin_cap = 0;
    void* r_sort = nullptr;
    size_t r_sort_cap = 0;
    void* r_ranges = nullptr;
    size_t r_ranges_cap = 0;
    void* r_rec = nullptr;          // depth sort of the records
    size_t r_rec_cap = 0;
    void* r_split = nullptr;        // deep-tile split: segment lists and per-segment pixel partials
    size_t r_split_cap = 0;
    void* db_ws = nullptr;          // density backward accumulators
    size_t db_cap = 0;
    uint64_t grid_chunk_pairs = 0;  // cell-list build chunk (0: GRID_CHUNK_PAIRS)
    uint32_t* r_total = nullptr;
    uint32_t last_rendered = 0;
    int last_tiles = 0;
    uint32_t last_dmin = 0;   // depth-bit offset and width of the last render's keys
    int last_dbits = 32;
    RasterView* r_views = nullptr;    // device copy of the views of the current render
    RasterView* r_views_h = nullptr;  // pinned host staging for it
    int* r_seg = nullptr;             // device: Gaussian segments of a segment-mode render (RASTER_MAX_VIEWS + 1)
    int* r_seg_h = nullptr;           // pinned host staging for it
    int* flag_h = nullptr;            // pinned, coherent: convergence flags and render counts written by kernels
                                      // (16 ints; [0]: a host check's flag; [3]: the CG iteration count at a
                                      // host check; [4], [5]: the flags of consecutive CG chunk graphs; [6]:
                                      // the unread flag of the two-kernel CG's final test; [7]: gn_done after
                                      // a GN-end graph; [8], [9]: every GN loop stopped and the GN count, at a
                                      // host check; [10..12]: a render's instance counts, [13]: their ready word)
    hipEvent_t chk_ev[2] = {nullptr, nullptr};  // recorded after each CG chunk graph
    RasterParams last_rp{};           // parameters of the last render (its binning stays in r_geo / r_bin)
    int blend_unroll = 4;             // Gaussians per compositing step of the 16x16 blend
    int blend_split = 0;              // instances per blend segment of a deep 16x16 tile (0: no split)
    bool node_cache = true;
    bool host_spin = false;           // the render's instance-count read polls its pinned slot
    bool sh_compose = false;          // samples_frames: one SH rotation by the product of the frames' rotations
    bool tile_cull = false;           // renders emit only the tiles that pass the conservative tile test           // six-point end-point blend through the LDS node-row cache
    // optional HIP-event timing of the fused six-point kernel inside deform_frame
    bool prof = false;
    // HIP events around the six-point (0), sample blend (1) and sample SH (2)
    // launches of arapgs_deform_frame while profiling is on
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_events[3];
    size_t prof_used[3] = {0, 0, 0};
};

struct arapgs_solver {
    int device = 0;  // the creating context's device; a solver never touches its context again
    SolverWork w;
    int host_check_every = 0;   // 0: device-side convergence only
    int cg_mode = 0;            // 0: one launch per CG phase, 1: persistent cooperative PCG, 2: two launches per iteration
    bool cg_predict = false;    // first host check of each GN iteration's CG at the previous solve's count
    int cg_pred[GN_MAX_ITERS] = {};  // those counts (0: none yet)
    int gn_pred = 0;            // the previous solve's GN count (0: none yet): the stop flag is read only after it
    bool host_spin = false;     // host checks poll the pinned flag instead of waiting for the stream
    int coarse_refresh = 8;     // frames between coarse-operator rebuilds
    int64_t solves = 0;
    SolverBatch one;            // this solver as a launch group of one
    // per-frame scratch (M nodes)
    NodeXf* table = nullptr;
    float* node_q = nullptr;
    float* node_tmp = nullptr;
};

namespace {

constexpr int FPS_BLOCKS = 1024;

int fail(arapgs_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code < 0 ? code : -1;
}

int hip_fail(arapgs_ctx* c, hipError_t e, const char* what) {
    return fail(c, -1000 - (int)e, "%s: %s", what, hipGetErrorString(e));
}

int check_launch(arapgs_ctx* c, int rc, const char* what) {
    if (rc == 0) return 0;
    if (rc <= -1000) return fail(c, rc, "%s: HIP error %d (%s)", what, -rc - 1000, hipGetErrorString((hipError_t)(-rc - 1000)));
    return fail(c, rc, "%s: unsupported arguments", what);
}

// Workspaces grow with 25 % headroom: per-frame sizes (tile instances) drift
// from frame to frame, and a re-allocation synchronises the device.
int grow(arapgs_ctx* c, void** p, size_t* cap, size_t bytes) {
    if (*cap >= bytes) return 0;
    if (*p) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "workspace");
        (void)hipFree(*p);
        *p = nullptr;
    }
    const size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(p, want);
    if (e != hipSuccess) { *cap = 0; return hip_fail(c, e, "workspace allocation"); }
    *cap = want;
    return 0;
}

#define CHECK_CTX(c) \
    if (!(c)) return -1;

}  // namespace

extern "C" {

const char* arapgs_version(void) { return "arapgs 0.1 (gfx950)"; }

int arapgs_create(int device, arapgs_ctx** out) {
    if (!out) return -1;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return -2;
    if (device < 0 || device >= n) return -3;
    if (hipSetDevice(device) != hipSuccess) return -4;
    arapgs_ctx* c = new arapgs_ctx();
    c->device = device;
    build_shrot_tables(&c->shrot);
    hipError_t e;
    if ((e = hipMalloc(&c->knn_count, sizeof(uint32_t))) != hipSuccess ||
        (e = hipMalloc(&c->fps_partial, sizeof(unsigned long long) * FPS_BLOCKS)) != hipSuccess ||
        (e = hipMalloc(&c->fps_counter, sizeof(unsigned int))) != hipSuccess ||
        (e = hipMalloc(&c->fps_best, sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMalloc(&c->r_total, 4 * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMalloc(&c->shrot_dev, sizeof(ShRotTables))) != hipSuccess ||
        (e = hipMalloc(&c->r_views, sizeof(RasterView) * RASTER_MAX_VIEWS)) != hipSuccess ||
        (e = hipHostMalloc(&c->r_views_h, sizeof(RasterView) * RASTER_MAX_VIEWS, hipHostMallocDefault)) != hipSuccess ||
        (e = hipMalloc(&c->r_seg, sizeof(int) * (RASTER_MAX_VIEWS + 1))) != hipSuccess ||
        (e = hipHostMalloc(&c->r_seg_h, sizeof(int) * (RASTER_MAX_VIEWS + 1), hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc(&c->flag_h, sizeof(int) * 16, hipHostMallocCoherent)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->chk_ev[0], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->chk_ev[1], hipEventDisableTiming)) != hipSuccess ||
        (e = hipMemcpy(c->shrot_dev, &c->shrot, sizeof(ShRotTables), hipMemcpyHostToDevice)) != hipSuccess) {
        arapgs_destroy(c);
        return -5;
    }
    (void)hipMemset(c->fps_counter, 0, sizeof(unsigned int));
    // the static KNN scratch's address on this device (arapgs_knn_noalloc);
    // a failure here is reported by that entry point, which retries
    uint32_t* knn_static = nullptr;
    (void)knn_noalloc_scratch(device, &knn_static);
    *out = c;
    return 0;
}

void arapgs_destroy(arapgs_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    for (auto& list : c->prof_events)
        for (auto& pr : list) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (void* p : {c->knn_list, (void*)c->knn_count, c->fps_dist, (void*)c->fps_partial, (void*)c->fps_counter,
                    (void*)c->fps_best, (void*)c->edge_tmp, c->r_geo, c->r_scan, c->r_bin, c->r_sort, c->r_ranges, c->r_rec, c->r_split, c->db_ws,
                    (void*)c->r_total, (void*)c->shrot_dev, (void*)c->r_views, (void*)c->r_seg})
        if (p) (void)hipFree(p);
    if (c->r_views_h) (void)hipHostFree(c->r_views_h);
    if (c->r_seg_h) (void)hipHostFree(c->r_seg_h);
    if (c->flag_h) (void)hipHostFree(c->flag_h);
    for (hipEvent_t ev : c->chk_ev)
        if (ev) (void)hipEventDestroy(ev);
    delete c;
}

const char* arapgs_last_error(const arapgs_ctx* c) { return c ? c->err.c_str() : "null context"; }

int arapgs_set_stream(arapgs_ctx* c, void* s) {
    CHECK_CTX(c);
    c->stream = (hipStream_t)s;
    return 0;
}

int arapgs_stream_create(int device, int cu_begin, int cu_count, void** stream_out) {
    if (!stream_out || cu_begin < 0 || cu_count < 0) return -1;
    *stream_out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return -1;
    hipStream_t st = nullptr;
    hipError_t e;
    if (cu_count == 0) {
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    } else {
        int ncu = 0;
        e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
        if (e != hipSuccess) return -1000 - (int)e;
        if (cu_begin + cu_count > ncu) return -1;
        std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
        for (int cu = cu_begin; cu < cu_begin + cu_count; cu++) mask[cu / 32] |= 1u << (cu % 32);
        e = hipExtStreamCreateWithCUMask(&st, (uint32_t)mask.size(), mask.data());
    }
    if (e != hipSuccess) return -1000 - (int)e;
    *stream_out = st;
    return 0;
}

int arapgs_stream_create_mask(int device, const uint32_t* mask_h, int words, void** stream_out) {
    if (!stream_out || !mask_h || words < 1) return -1;
    *stream_out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return -1;
    hipStream_t st = nullptr;
    const hipError_t e = hipExtStreamCreateWithCUMask(&st, (uint32_t)words, mask_h);
    if (e != hipSuccess) return -1000 - (int)e;
    *stream_out = st;
    return 0;
}

int arapgs_stream_destroy(void* stream) {
    if (!stream) return -1;
    return hipStreamDestroy(static_cast<hipStream_t>(stream)) == hipSuccess ? 0 : -1;
}

int arapgs_set_node_cache(arapgs_ctx* c, int on) {
    CHECK_CTX(c);
    if (on < 0 || on > 1) return fail(c, -1, "set_node_cache: on must be 0 or 1");
    c->node_cache = on != 0;
    return 0;
}

int arapgs_set_grid_chunk(arapgs_ctx* c, int64_t pairs) {
    CHECK_CTX(c);
    if (pairs < 0) return fail(c, -1, "set_grid_chunk: negative chunk");
    c->grid_chunk_pairs = (uint64_t)pairs;
    return 0;
}

int arapgs_synchronize(arapgs_ctx* c) {
    CHECK_CTX(c);
    hipError_t e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? 0 : hip_fail(c, e, "synchronize");
}

// ------------------------------------------------------------ graph build
int arapgs_fps(arapgs_ctx* c, const float* pos, int64_t n, int m, int* out_idx, int* count_h) {
    CHECK_CTX(c);
    if (!pos || !out_idx || !count_h || n <= 0 || m <= 0) return fail(c, -1, "fps: bad arguments");
    (void)hipSetDevice(c->device);
    const int mm = (int)std::min<int64_t>(m, n);
    int rc = grow(c, &c->fps_dist, &c->fps_dist_cap, sizeof(float) * (size_t)n);
    if (rc) return rc;
    // Number of argmax steps of the reference loop (helper.cpp:157-174): m-1,
    // except m == 1 with a first node != 0, which takes one step.
    int first_h = 0;
    int steps = mm - 1;
    if (mm == 1) {
        rc = check_launch(c, launch_fps(pos, n, 0, out_idx, (float*)c->fps_dist, c->fps_partial, FPS_BLOCKS,
                                        c->fps_counter, c->fps_best, c->stream), "fps");
        if (rc) return rc;
        hipError_t e = hipMemcpyAsync(&first_h, out_idx, sizeof(int), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "fps");
        steps = first_h != 0 ? 1 : 0;
        if (steps == 0) { *count_h = 1; return 0; }
    }
    rc = check_launch(c, launch_fps(pos, n, steps, out_idx, (float*)c->fps_dist, c->fps_partial, FPS_BLOCKS,
                                    c->fps_counter, c->fps_best, c->stream), "fps");
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "fps");
    *count_h = steps + 1;
    return 0;
}

int arapgs_knn(arapgs_ctx* c, const float* nodes, int m, const float* queries, int64_t q, int k, uint32_t* out_idx,
               float* out_w) {
    CHECK_CTX(c);
    if (!nodes || !queries || !out_idx || q < 0) return fail(c, -1, "knn: bad arguments");
    if (k < 1 || k > KNN_MAX_T) return fail(c, -1, "knn: k must be in 1..12");
    if (m < k + 1) return fail(c, -1, "knn: need at least k+1 nodes (m=%d, k=%d)", m, k);
    if ((size_t)m * 3 * sizeof(float) > 64 * 1024) return fail(c, -1, "knn: m=%d nodes exceed the LDS stage", m);
    if (q > 0xffffffffLL) return fail(c, -1, "knn: too many queries");
    (void)hipSetDevice(c->device);
    int rc = grow(c, &c->knn_list, &c->knn_list_cap, sizeof(uint32_t) * (size_t)std::max<int64_t>(q, 1));
    if (rc) return rc;
    return check_launch(c, launch_knn(nodes, m, queries, q, k, out_idx, out_w, (uint32_t*)c->knn_list, c->knn_count,
                                      c->stream), "knn");
}

// Scratch-free variant of arapgs_knn for callers that cannot allocate.
int arapgs_knn_noalloc(arapgs_ctx* c, const float* nodes, int m, const float* queries, int64_t q, int k,
                       uint32_t* out_idx, float* out_w) {
    CHECK_CTX(c);
    if (!nodes || !queries || !out_idx || q < 0) return fail(c, -1, "knn_noalloc: bad arguments");
    if (k < 1 || k > KNN_MAX_T) return fail(c, -1, "knn_noalloc: k must be in 1..12");
    if (m < k + 1) return fail(c, -1, "knn_noalloc: need at least k+1 nodes (m=%d, k=%d)", m, k);
    if ((size_t)m * 3 * sizeof(float) > 64 * 1024) return fail(c, -1, "knn_noalloc: m=%d nodes exceed the LDS stage", m);
    if (q > KNN_NOALLOC_MAX_QUERIES)
        return fail(c, -1, "knn_noalloc: %lld queries exceed the static scratch (at most %lld)", (long long)q,
                    (long long)KNN_NOALLOC_MAX_QUERIES);
    (void)hipSetDevice(c->device);
    uint32_t* s = nullptr;
    int rc = knn_noalloc_scratch(c->device, &s);
    if (rc) return check_launch(c, rc, "knn_noalloc (static scratch address)");
    return check_launch(c, launch_knn(nodes, m, queries, q, k, out_idx, out_w, s + 64, s, c->stream), "knn_noalloc");
}

}  // extern "C"

namespace {
__global__ void drop_first_kernel(const uint32_t* __restrict__ idx, int m, int k, uint32_t* __restrict__ nbr) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)m * k) return;
    const int64_t i = e / k, j = e % k;
    nbr[e] = idx[i * (k + 1) + j + 1];
}
}  // namespace

extern "C" {

int arapgs_graph_edges(arapgs_ctx* c, const float* nodes, int m, int k, uint32_t* out_nbr) {
    CHECK_CTX(c);
    if (!out_nbr) return fail(c, -1, "graph_edges: bad arguments");
    int rc = grow(c, (void**)&c->edge_tmp, &c->edge_tmp_cap, sizeof(uint32_t) * (size_t)m * (k + 1));
    if (rc) return rc;
    rc = arapgs_knn(c, nodes, m, nodes, m, k, c->edge_tmp, nullptr);
    if (rc) return rc;
    const int64_t tot = (int64_t)m * k;
    hipLaunchKernelGGL(drop_first_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, c->edge_tmp,
                       m, k, out_nbr);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(c, e, "graph_edges");
}

int arapgs_end_points(arapgs_ctx* c, const float* pos, const float* rot, const float* scale, int64_t n, float pad,
                      float coeff, float* ends) {
    CHECK_CTX(c);
    if (!pos || !rot || !scale || !ends || n < 0) return fail(c, -1, "end_points: bad arguments");
    return check_launch(c, launch_end_points(pos, rot, scale, n, pad, coeff, ends, c->stream), "end_points");
}

// ------------------------------------------------------------ solver
int arapgs_solver_create(arapgs_ctx* c, const arapgs_solver_desc* d, arapgs_solver** out) {
    CHECK_CTX(c);
    if (!d || !out || !d->nbr_h || !d->src_idx_h || !d->src_w_h || !d->node_pos_h || (d->nblocks > 0 && (!d->block_ptr_h || !d->block_nodes_h || !d->block_type_h)))
        return fail(c, -1, "solver_create: bad arguments");
    *out = nullptr;
    (void)hipSetDevice(c->device);
    arapgs_solver* s = new arapgs_solver();
    s->device = c->device;
    SolverDesc sd;
    sd.M = d->m; sd.K = d->k; sd.nbr = d->nbr_h; sd.block_ptr = d->block_ptr_h; sd.block_nodes = d->block_nodes_h;
    sd.block_type = d->block_type_h; sd.nblocks = d->nblocks; sd.src_idx = d->src_idx_h; sd.src_w = d->src_w_h;
    sd.node_pos = d->node_pos_h; sd.w_rot = d->w_rot; sd.w_reg = d->w_reg; sd.w_con = d->w_con;
    sd.wf = d->weighting_factor; sd.center = d->constraints_on_center; sd.n_agg = d->n_aggregates;
    std::string err;
    int rc = solver_build(sd, s->w, err);
    if (rc) {
        delete s;
        return fail(c, rc, "%s", err.c_str());
    }
    if (d->cg_max_iters > 0) s->w.cg_max_iters = d->cg_max_iters;
    if (d->cg_tol > 0.0f) s->w.cg_tol = d->cg_tol;
    rc = solver_batch_init(s->one, {&s->w}, err);
    if (rc) {
        solver_free(s->w);
        delete s;
        return fail(c, rc, "%s", err.c_str());
    }
    hipError_t e = hipMalloc(&s->table, sizeof(NodeXf) * (size_t)d->m);
    if (e == hipSuccess) e = hipMalloc(&s->node_q, sizeof(float) * 4 * (size_t)d->m);
    if (e == hipSuccess) e = hipMalloc(&s->node_tmp, sizeof(float) * 3 * (size_t)d->m);
    if (e != hipSuccess) {
        arapgs_solver_destroy(s);
        return hip_fail(c, e, "solver_create");
    }
    *out = s;
    return 0;
}

void arapgs_solver_destroy(arapgs_solver* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    solver_batch_free(s->one);
    solver_free(s->w);
    for (void* p : {(void*)s->table, (void*)s->node_q, (void*)s->node_tmp})
        if (p) (void)hipFree(p);
    delete s;
}

int arapgs_solver_set_host_checks(arapgs_solver* s, int every) {
    if (!s || every < 0) return -1;
    s->host_check_every = every;
    return 0;
}

int arapgs_solver_set_cg_tol(arapgs_solver* s, float tol) {
    if (!s || !(tol > 0.0f) || !(tol < 1.0f)) return -1;
    if ((double)tol == s->w.cg_tol) return 0;
    (void)hipSetDevice(s->device);
    // the batch entries carry the tolerance; nothing may still read the old ones
    if (hipDeviceSynchronize() != hipSuccess) return -1000;
    s->w.cg_tol = tol;
    std::string err;
    return solver_batch_init(s->one, {&s->w}, err);
}

int arapgs_solver_set_cg_mode(arapgs_solver* s, int mode) {
    if (!s || mode < 0 || mode > 2) return -1;
    s->cg_mode = mode;
    return 0;
}

int arapgs_solver_apply(arapgs_ctx* c, arapgs_solver* s, const float* node_pos, const float* aims, const float* x,
                        const float* v, float* q, float* b) {
    CHECK_CTX(c);
    if (!s || !node_pos || !aims || !x || !v || !q || !b) return fail(c, -1, "solver_apply: bad arguments");
    int rc = check_launch(c, solver_batch_set_frame(s->one, &node_pos, &aims, c->stream), "solver_apply");
    if (rc) return rc;
    return check_launch(c, solver_apply_test(s->one, x, v, q, b, c->stream), "solver_apply");
}

int arapgs_solver_last_stats(arapgs_solver* s, int* gn_iters_h, int* cg_iters_h) {
    if (!s) return -1;
    (void)hipSetDevice(s->device);
    GnScalars gs;
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&gs, s->w.gs, sizeof gs, hipMemcpyDeviceToHost) != hipSuccess)
        return -1000;
    if (gn_iters_h) *gn_iters_h = gs.iters_done;
    if (cg_iters_h) *cg_iters_h = gs.cg_total;
    return 0;
}

int arapgs_solver_set_warm_start(arapgs_solver* s, int on) {
    if (!s || on < 0 || on > 1) return -1;
    (void)hipSetDevice(s->device);
    // every (re)enable starts from zero steps, so a timed run does not depend on earlier ones
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemset(s->w.hwarm, 0, sizeof(float) * (size_t)s->w.warm_slots * s->w.n) != hipSuccess)
        return -1000;
    s->w.warm = on;
    return 0;
}

int arapgs_solver_reset(arapgs_solver* s) {
    if (!s) return -1;
    (void)hipSetDevice(s->device);
    // a side-stream rebuild in flight finishes before its buffers are forgotten
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemset(s->w.hwarm, 0, sizeof(float) * (size_t)s->w.warm_slots * s->w.n) != hipSuccess)
        return -1000;
    s->w.pending = false;
    s->w.pending_age = 0;
    s->w.have_coarse = false;
    s->solves = 0;
    for (int& v : s->cg_pred) v = 0;
    s->gn_pred = 0;
    return 0;
}

int arapgs_solver_set_cg_predict(arapgs_solver* s, int on) {
    if (!s || on < 0 || on > 1) return -1;
    s->cg_predict = on != 0;
    for (int& v : s->cg_pred) v = 0;
    s->gn_pred = 0;
    return 0;
}

int arapgs_solver_set_host_spin(arapgs_solver* s, int on) {
    if (!s || on < 0 || on > 1) return -1;
    s->host_spin = on != 0;
    return 0;
}

int arapgs_solver_set_cg_block(arapgs_solver* s, int threads) {
    if (!s || (threads != 64 && threads != 128 && threads != 256 && threads != 512)) return -1;
    if (threads == s->w.cg_block) return 0;
    (void)hipSetDevice(s->device);
    // the batch entries are rebuilt; nothing may still read the old ones
    if (hipDeviceSynchronize() != hipSuccess) return -1000;
    s->w.cg_block = threads;
    std::string err;
    return solver_batch_init(s->one, {&s->w}, err);
}

int arapgs_solver_set_cg_assembled(arapgs_solver* s, int on) {
    if (!s || on < 0 || on > 1) return -1;
    if (on && s->w.dev.n_pairs == 0) return -2;  // centre mode: not built
    if ((on != 0) == s->w.cg_asm) return 0;
    (void)hipSetDevice(s->device);
    // the batch entries are rebuilt; nothing may still read the old ones
    if (hipDeviceSynchronize() != hipSuccess) return -1000;
    s->w.cg_asm = on != 0;
    std::string err;
    return solver_batch_init(s->one, {&s->w}, err);
}

int arapgs_solver_set_cg_graphs(arapgs_solver* s, int on) {
    if (!s || on < 0 || on > 1) return -1;
    s->one.graphs = on != 0;
    return 0;
}

int arapgs_solver_set_side_stream(arapgs_solver* s, void* stream) {
    if (!s) return -1;
    (void)hipSetDevice(s->device);
    // a rebuild in flight on the old side stream finishes before the switch
    if (hipDeviceSynchronize() != hipSuccess) return -1000;
    s->one.side_ext = (hipStream_t)stream;
    return 0;
}

int arapgs_solver_set_coarse_lag(arapgs_solver* s, int solves) {
    if (!s || solves < 1 || solves > 8) return -1;
    s->one.lag = solves;
    return 0;
}

int arapgs_solver_set_async_refresh(arapgs_solver* s, int on) {
    if (!s || on < 0 || on > 1) return -1;
    s->one.async = on != 0;
    return 0;
}

int arapgs_solver_set_coarse_refresh(arapgs_solver* s, int frames) {
    if (!s || frames < 1) return -1;
    s->coarse_refresh = frames;
    s->solves = 0;
    return 0;
}

int arapgs_solver_dims(const arapgs_solver* s, int* n_free_h, int* n_unknowns_h) {
    if (!s) return -1;
    if (n_free_h) *n_free_h = s->w.dev.nF;
    if (n_unknowns_h) *n_unknowns_h = s->w.n;
    return 0;
}

}  // extern "C"

namespace {
// The Gauss-Newton loop of Deform::optimize (Deform.cpp:95-169) over a launch
// group of solvers: every phase is one launch for all scenes; the host reads
// the groups' convergence flags every `every` CG iterations and after every
// GN iteration (every == 0: device-side flags only, all launches queued).
// pred (when not null): per GN iteration, the CG iteration count of the
// previous solve; the first host check of that GN iteration's CG comes after
// that many iterations (then every `every`), and the count found is stored back.
// gn_pred (when not null, with host checks and a CG of separate launches): the GN
// count of the previous solve; the stop flag of GN iterations before that
// count is not waited for: the next iteration is queued at once (its kernels
// do nothing once the loop has stopped) and the flag is read with that
// iteration's first CG check. The count found is stored back. The iterates do
// not depend on it.
// spin: host checks poll the flag's pinned slot (published with a
// system-scope release by the flag kernel) instead of waiting for the stream;
// a slot still unset after HOST_SPIN_US falls back to the stream wait, which
// also reports a failed launch.
constexpr int HOST_SPIN_US = 50000;

int run_gn(arapgs_ctx* c, SolverBatch& B, int every, int cg_mode, bool refresh, int iters, int stop_rule,
           int* pred = nullptr, bool spin = false, int* gn_pred = nullptr) {
    int rc = check_launch(c, solver_frame_setup(B, refresh, c->stream), "gn_solve setup");
    if (rc) return rc;
    // the flag kernel writes into pinned host memory: no device-to-host copy
    B.fused = cg_mode == 2;
    int* const gn_slot = c->flag_h + 8;
    auto host_flag = [&](int which, int* out, int k_done, int* gn = nullptr) -> int {
        __atomic_store_n(c->flag_h, -1, __ATOMIC_SEQ_CST);
        int r = check_launch(c, solver_batch_flag(B, which, c->stream, c->flag_h, k_done, c->flag_h + 3, gn),
                             "gn_solve");
        if (r) return r;
        bool seen = false;
        if (spin) {
            const auto t0 = std::chrono::steady_clock::now();
            while (!(seen = __atomic_load_n(c->flag_h, __ATOMIC_ACQUIRE) != -1) &&
                   std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(HOST_SPIN_US)) {
            }
        }
        if (!seen) {
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) return hip_fail(c, e, "gn_solve");
        }
        *out = __atomic_load_n(c->flag_h, __ATOMIC_ACQUIRE);
        return 0;
    };
    const bool gn_spec = gn_pred && every > 0 && cg_mode != 1 && *gn_pred > 0;
    bool gn_unread = false;  // an earlier GN iteration's stop flag has not been read
    // graphs: the GN phases and the CG chunks replay as captured graphs (a
    // context without its own stream cannot capture)
    const bool graphs = B.graphs && every > 0 && c->stream && cg_mode != 1;
    for (int it = 0; it < iters; it++) {
        rc = check_launch(c, graphs ? solver_gn_begin_graph(B, it, c->stream) : solver_gn_begin(B, it, c->stream),
                          "gn_solve rhs");
        if (rc) return rc;
        const bool predict = pred && every > 0 && it < GN_MAX_ITERS;
        if (cg_mode == 1) {
            rc = check_launch(c, solver_cg_persistent(*B.w[0], c->stream), "gn_solve cg");
            if (rc) return rc;
            // the grid barrier gives up after a bounded spin: report it instead
            // of handing an unfinished step to the line search
            int bar_err = 0;
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) return hip_fail(c, e, "gn_solve cg");
            if ((rc = check_launch(c, solver_barrier_error(*B.w[0], &bar_err), "gn_solve cg"))) return rc;
            if (bar_err) return fail(c, -6, "gn_solve: the persistent CG kernel's grid barrier timed out");
        } else {
            // every scene's own cap is enforced on the device (cg_done); loop to the largest
            int cap = 0;
            for (const SolverWork* w : B.w) cap = std::max(cap, w->cg_max_iters);
            if (graphs) {
                // chunk j is one graph launch; the host reads chunk j-1's flag
                // while chunk j is already queued, so the stream never waits for
                // the host (a chunk queued past convergence exits kernel by kernel).
                // The first chunk is the predicted count (or `every`), then `every`.
                int* const slots[2] = {c->flag_h + 4, c->flag_h + 5};
                bool seen_done = false;
                int k0 = 0;
                for (int j = 0; k0 < cap; j++) {
                    const int L = std::min(cap - k0, j == 0 && predict && pred[it] > 0 ? pred[it] : every);
                    rc = check_launch(c, solver_cg_chunk(B, k0, L, c->stream, slots[j & 1], c->flag_h + 3),
                                      "gn_solve cg graph");
                    if (rc) return rc;
                    k0 += L;
                    hipError_t e = hipEventRecord(c->chk_ev[j & 1], c->stream);
                    if (e != hipSuccess) return hip_fail(c, e, "gn_solve cg graph");
                    if (j > 0) {
                        e = hipEventSynchronize(c->chk_ev[(j - 1) & 1]);
                        if (e != hipSuccess) return hip_fail(c, e, "gn_solve cg graph");
                        if (*(volatile int*)slots[(j - 1) & 1]) {
                            seen_done = true;
                            break;
                        }
                    }
                }
                if (!seen_done) {
                    // the last chunk's flag (its count is final too)
                    hipError_t e = hipStreamSynchronize(c->stream);
                    if (e != hipSuccess) return hip_fail(c, e, "gn_solve cg graph");
                    seen_done = true;
                }
                if (predict) pred[it] = std::max(1, (int)*(volatile int*)(c->flag_h + 3));
            } else {
                bool seen_done = false;
                int next_check = predict && pred[it] > 0 ? std::min(pred[it], cap) : every;
                for (int k = 0; k < cap; k++) {
                    rc = check_launch(c, solver_cg_iteration(B, k, c->stream), "gn_solve cg");
                    if (rc) return rc;
                    if (every > 0 && k + 1 == next_check) {
                        int done = 0;
                        if ((rc = host_flag(0, &done, k + 1, gn_unread ? gn_slot : nullptr))) return rc;
                        if (gn_unread) {
                            gn_unread = false;
                            if (gn_slot[0]) {
                                // the loop stopped before this iteration: its kernels did nothing
                                *gn_pred = std::max(1, gn_slot[1]);
                                return 0;
                            }
                        }
                        if (done) {
                            seen_done = true;
                            break;
                        }
                        next_check += every;
                    }
                }
                // the two-kernel form tests convergence at the start of the next
                // apply: record the test after the last iteration queued
                if (B.fused && !seen_done) {
                    rc = check_launch(c, solver_batch_flag(B, 0, c->stream, c->flag_h + 6, cap), "gn_solve cg");
                    if (rc) return rc;
                }
                if (predict) pred[it] = seen_done ? std::max(1, (int)*(volatile int*)(c->flag_h + 3)) : 0;
            }
        }
        if (graphs) {
            rc = check_launch(c, solver_gn_end_graph(B, it, stop_rule, c->stream, c->flag_h + 7),
                              "gn_solve line search");
            if (rc) return rc;
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) return hip_fail(c, e, "gn_solve line search");
            if (*(volatile int*)(c->flag_h + 7)) break;
            continue;
        }
        rc = check_launch(c, solver_gn_end(B, it, stop_rule, c->stream), "gn_solve line search");
        if (rc) return rc;
        if (every > 0) {
            if (gn_spec && it + 1 < *gn_pred && it + 1 < iters) {
                gn_unread = true;
                continue;
            }
            int done = 0;
            if ((rc = host_flag(1, &done, 0, gn_pred ? gn_slot : nullptr))) return rc;
            gn_unread = false;
            if (gn_pred) *gn_pred = std::max(1, gn_slot[1]);
            if (done) break;
        }
    }
    return 0;
}
}  // namespace

struct arapgs_solver_batch {
    int device = 0;
    SolverBatch b;
    int host_check_every = 8;
    int coarse_refresh = 8;
    int64_t solves = 0;
};

extern "C" {

int arapgs_gn_solve(arapgs_ctx* c, arapgs_solver* s, const float* node_pos, const float* aims, int iters,
                    int stop_rule, float* x_out, double* energies, int* iters_done) {
    CHECK_CTX(c);
    if (!s || !node_pos || !aims || !energies || iters < 0) return fail(c, -1, "gn_solve: bad arguments");
    if (iters > GN_MAX_ITERS) return fail(c, -1, "gn_solve: at most %d Gauss-Newton iterations", GN_MAX_ITERS);
    int rc = check_launch(c, solver_batch_set_frame(s->one, &node_pos, &aims, c->stream), "gn_solve");
    if (rc) return rc;
    const bool refresh = (s->solves % std::max(1, s->coarse_refresh)) == 0;
    s->solves++;
    if ((rc = run_gn(c, s->one, s->host_check_every, s->cg_mode, refresh, iters, stop_rule,
                     s->cg_predict ? s->cg_pred : nullptr, s->host_spin, s->cg_predict ? &s->gn_pred : nullptr)))
        return rc;
    SolverOutPtrs out = {};
    out.x[0] = x_out;
    out.energies[0] = energies;
    out.iters_done[0] = iters_done;
    return check_launch(c, solver_batch_finish(s->one, out, c->stream), "gn_solve");
}

int arapgs_solver_batch_create(arapgs_ctx* c, arapgs_solver* const* solvers, int n, arapgs_solver_batch** out) {
    CHECK_CTX(c);
    if (!solvers || !out || n < 1 || n > SOLVER_BATCH_MAX)
        return fail(c, -1, "solver_batch_create: need 1..%d solvers", SOLVER_BATCH_MAX);
    *out = nullptr;
    std::vector<SolverWork*> w;
    for (int i = 0; i < n; i++) {
        if (!solvers[i]) return fail(c, -1, "solver_batch_create: solver %d is null", i);
        if (solvers[i]->device != c->device) return fail(c, -1, "solver_batch_create: solver %d is on another device", i);
        for (int j = 0; j < i; j++)
            if (solvers[j] == solvers[i]) return fail(c, -1, "solver_batch_create: solver %d appears twice", i);
        w.push_back(&solvers[i]->w);
    }
    (void)hipSetDevice(c->device);
    arapgs_solver_batch* b = new arapgs_solver_batch();
    b->device = c->device;
    std::string err;
    const int rc = solver_batch_init(b->b, w, err);
    if (rc) {
        delete b;
        return fail(c, rc, "%s", err.c_str());
    }
    *out = b;
    return 0;
}

void arapgs_solver_batch_destroy(arapgs_solver_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    solver_batch_free(b->b);
    delete b;
}

int arapgs_solver_batch_set_host_checks(arapgs_solver_batch* b, int every) {
    if (!b || every < 0) return -1;
    b->host_check_every = every;
    return 0;
}

int arapgs_solver_batch_set_async_refresh(arapgs_solver_batch* b, int on) {
    if (!b || on < 0 || on > 1) return -1;
    b->b.async = on != 0;
    return 0;
}

int arapgs_solver_batch_set_cg_graphs(arapgs_solver_batch* b, int on) {
    if (!b || on < 0 || on > 1) return -1;
    b->b.graphs = on != 0;
    return 0;
}

int arapgs_solver_batch_set_coarse_refresh(arapgs_solver_batch* b, int frames) {
    if (!b || frames < 1) return -1;
    b->coarse_refresh = frames;
    b->solves = 0;
    return 0;
}

int arapgs_solver_batch_solve(arapgs_ctx* c, arapgs_solver_batch* b, const float* const* node_pos, const float* const* aims,
                              int iters, int stop_rule, float* const* x_out, double* const* energies,
                              int* const* iters_done, arapgs_node_xf* const* tables) {
    CHECK_CTX(c);
    if (!b || !node_pos || !aims || iters < 0) return fail(c, -1, "solver_batch_solve: bad arguments");
    if (iters > GN_MAX_ITERS) return fail(c, -1, "solver_batch_solve: at most %d Gauss-Newton iterations", GN_MAX_ITERS);
    if (b->device != c->device) return fail(c, -1, "solver_batch_solve: the batch is on another device");
    const int n = (int)b->b.w.size();
    for (int i = 0; i < n; i++)
        if (!node_pos[i] || !aims[i]) return fail(c, -1, "solver_batch_solve: scene %d lacks node positions or aims", i);
    if (!solver_batch_warm_uniform(b->b))
        return fail(c, -1, "solver_batch_solve: every solver of a batch needs the same warm-start setting");
    int rc = check_launch(c, solver_batch_set_frame(b->b, node_pos, aims, c->stream), "solver_batch_solve");
    if (rc) return rc;
    const bool refresh = (b->solves % std::max(1, b->coarse_refresh)) == 0;
    b->solves++;
    if ((rc = run_gn(c, b->b, b->host_check_every, 0, refresh, iters, stop_rule))) return rc;
    SolverOutPtrs out = {};
    for (int i = 0; i < n; i++) {
        out.x[i] = x_out ? x_out[i] : nullptr;
        out.energies[i] = energies ? energies[i] : nullptr;
        out.iters_done[i] = iters_done ? iters_done[i] : nullptr;
        out.table[i] = tables ? tables[i] : nullptr;
    }
    return check_launch(c, solver_batch_finish(b->b, out, c->stream), "solver_batch_solve");
}

int arapgs_gn_energy(arapgs_ctx* c, arapgs_solver* s, const float* node_pos, const float* aims, const float* x,
                     double* energy) {
    CHECK_CTX(c);
    if (!s || !node_pos || !aims || !x || !energy) return fail(c, -1, "gn_energy: bad arguments");
    int rc = check_launch(c, solver_batch_set_frame(s->one, &node_pos, &aims, c->stream), "gn_energy");
    if (rc) return rc;
    return check_launch(c, solver_energy(s->one, x, energy, c->stream), "gn_energy");
}

int arapgs_put_free_inputs(arapgs_ctx* c, arapgs_solver* s, const float* node_pos, const float* x,
                           arapgs_node_xf* table) {
    CHECK_CTX(c);
    if (!s || !node_pos || !x || !table) return fail(c, -1, "put_free_inputs: bad arguments");
    return check_launch(c, launch_build_node_table(s->w.dev.M, node_pos, x, s->w.dev.free_id, (NodeXf*)table,
                                                   c->stream), "put_free_inputs");
}

// ------------------------------------------------------------ deformation
int arapgs_blend_points(arapgs_ctx* c, const float* pts, int64_t p, const uint32_t* idx, const float* w, int k,
                        const arapgs_node_xf* table, int m, const uint8_t* skip, float* out) {
    CHECK_CTX(c);
    if (!pts || !idx || !w || !table || !out || p < 0 || k < 1 || k > KNN_MAX_T)
        return fail(c, -1, "blend_points: bad arguments");
    return check_launch(c, launch_blend_points(pts, p, idx, w, k, (const NodeXf*)table, m, skip, out, c->stream,
                                               c->node_cache),
                        "blend_points");
}

int arapgs_six_point_update(arapgs_ctx* c, int64_t n, int k, const arapgs_node_xf* table, int m,
                            const uint32_t* ends_idx, const float* ends_w, float* ends, const float* scale0,
                            const uint8_t* is_static, float pad, float coeff, float* pos, float* rot, float* scale,
                            float* sh) {
    CHECK_CTX(c);
    if (!table || !ends_idx || !ends_w || !ends || !scale0 || !pos || !rot || !scale || !sh || n < 0 || k < 1 ||
        k > KNN_MAX_T)
        return fail(c, -1, "six_point_update: bad arguments");
    if (m < 1) return fail(c, -1, "six_point_update: need m >= 1 nodes");
    if ((uintptr_t)rot % 16) return fail(c, -1, "six_point_update: rot must be 16-byte aligned");
    if ((uintptr_t)sh % 16) return fail(c, -1, "six_point_update: sh must be 16-byte aligned");
    return check_launch(c, launch_six_point(n, k, (const NodeXf*)table, m, ends_idx, ends_w, ends, scale0, is_static,
                                            pad, coeff, pos, rot, scale, sh, c->shrot_dev, c->stream, c->node_cache),
                        "six_point_update");
}

int arapgs_six_point_update_from(arapgs_ctx* c, int64_t n, int k, const arapgs_node_xf* table, int m,
                                 const uint32_t* ends_idx, const float* ends_w, float* ends, const float* scale0,
                                 const uint8_t* is_static, float pad, float coeff, const float* pos_in,
                                 const float* rot_in, const float* scale_in, const float* sh_in, float* pos,
                                 float* rot, float* scale, float* sh) {
    CHECK_CTX(c);
    if (!table || !ends_idx || !ends_w || !ends || !scale0 || !pos || !rot || !scale || !sh || !pos_in || !rot_in ||
        !scale_in || !sh_in || n < 0 || k < 1 || k > KNN_MAX_T)
        return fail(c, -1, "six_point_update_from: bad arguments");
    if (m < 1) return fail(c, -1, "six_point_update_from: need m >= 1 nodes");
    if ((uintptr_t)rot % 16 || (uintptr_t)rot_in % 16 || (uintptr_t)sh % 16 || (uintptr_t)sh_in % 16)
        return fail(c, -1, "six_point_update_from: rot and sh must be 16-byte aligned");
    // the two sets either coincide or do not overlap
    auto overlap = [](const float* a, const float* b, int64_t cnt) {
        return a != b && a < b + cnt && b < a + cnt;
    };
    if (overlap(pos_in, pos, 3 * n) || overlap(rot_in, rot, 4 * n) || overlap(scale_in, scale, 3 * n) ||
        overlap(sh_in, sh, 48 * n))
        return fail(c, -1, "six_point_update_from: input and output sets overlap");
    const SixPointIn in{pos_in, rot_in, scale_in, sh_in};
    return check_launch(c, launch_six_point(n, k, (const NodeXf*)table, m, ends_idx, ends_w, ends, scale0, is_static,
                                            pad, coeff, pos, rot, scale, sh, c->shrot_dev, c->stream, c->node_cache,
                                            &in),
                        "six_point_update_from");
}

int arapgs_node_quats(arapgs_ctx* c, const arapgs_node_xf* table, int m, float* q) {
    CHECK_CTX(c);
    if (!table || !q || m < 0) return fail(c, -1, "node_quats: bad arguments");
    return check_launch(c, launch_node_quats(m, (const NodeXf*)table, q, c->stream), "node_quats");
}

int arapgs_sample_sh(arapgs_ctx* c, int64_t s, int k, const float* w, const uint32_t* idx, const float* node_q,
                     const uint8_t* is_static, float* feat) {
    CHECK_CTX(c);
    if (!w || !idx || !node_q || !feat || s < 0 || k < 1 || k > KNN_MAX_T) return fail(c, -1, "sample_sh: bad arguments");
    return check_launch(c, launch_sample_sh(s, k, w, idx, node_q, is_static, feat, c->shrot_dev, c->stream), "sample_sh");
}

int arapgs_samples_frames(arapgs_ctx* c, int64_t s, int k, int nf, const arapgs_node_xf* const* tables_h,
                          const float* const* node_q_h, const uint32_t* idx, const float* w, const uint8_t* is_static,
                          float* samples, float* feat) {
    CHECK_CTX(c);
    if (!idx || !w || !samples || !feat || s < 0 || nf < 0 || (nf > 0 && (!tables_h || !node_q_h)))
        return fail(c, -1, "samples_frames: bad arguments");
    if (k != 8) return fail(c, -1, "samples_frames: k must be 8");
    if (nf > SAMPLE_FRAMES_MAX) return fail(c, -1, "samples_frames: at most %d frames per call", SAMPLE_FRAMES_MAX);
    if ((uintptr_t)idx % 16 || (uintptr_t)w % 16 || (uintptr_t)feat % 16)
        return fail(c, -1, "samples_frames: idx, w and feat must be 16-byte aligned");
    SampleFrames F{};
    F.n = nf;
    for (int f = 0; f < nf; f++) {
        if (!tables_h[f] || !node_q_h[f]) return fail(c, -1, "samples_frames: frame %d lacks a table or quaternions", f);
        F.table[f] = (const NodeXf*)tables_h[f];
        F.q[f] = node_q_h[f];
    }
    return check_launch(c, launch_samples_frames(s, k, F, idx, w, is_static, samples, feat, c->shrot_dev, c->stream,
                                                 c->sh_compose),
                        "samples_frames");
}

int arapgs_set_host_spin(arapgs_ctx* c, int on) {
    CHECK_CTX(c);
    if (on < 0 || on > 1) return fail(c, -1, "set_host_spin: on must be 0 or 1");
    c->host_spin = on != 0;
    return 0;
}

int arapgs_set_tile_cull(arapgs_ctx* c, int on) {
    CHECK_CTX(c);
    if (on < 0 || on > 1) return fail(c, -1, "set_tile_cull: on must be 0 or 1");
    c->tile_cull = on != 0;
    return 0;
}

int arapgs_set_sh_compose(arapgs_ctx* c, int on) {
    CHECK_CTX(c);
    if (on < 0 || on > 1) return fail(c, -1, "set_sh_compose: on must be 0 or 1");
    c->sh_compose = on != 0;
    return 0;
}

int arapgs_sh_rotate(arapgs_ctx* c, int64_t n, const float* r9, float* sh) {
    CHECK_CTX(c);
    if (!r9 || !sh || n < 0) return fail(c, -1, "sh_rotate: bad arguments");
    return check_launch(c, launch_sh_rotate(n, r9, sh, c->shrot_dev, c->stream), "sh_rotate");
}

int arapgs_polar(arapgs_ctx* c, int64_t n, const double* m9, double* r9, double* sdiag) {
    CHECK_CTX(c);
    if (!m9 || !r9 || !sdiag || n < 0) return fail(c, -1, "polar: bad arguments");
    return check_launch(c, launch_polar(n, m9, r9, sdiag, c->stream), "polar");
}

int arapgs_tile_test_host(int64_t n, const float* xy, const float* conic, const float* box, uint8_t* keep) {
    if (n < 0 || (n > 0 && (!xy || !conic || !box || !keep))) return -1;
    for (int64_t i = 0; i < n; i++) {
        const float* b = box + 4 * i;
        keep[i] = tile_may_contribute(make_float2(xy[2 * i], xy[2 * i + 1]),
                                      make_float4(conic[4 * i], conic[4 * i + 1], conic[4 * i + 2], conic[4 * i + 3]),
                                      b[0], b[1], b[2], b[3])
                      ? 1
                      : 0;
    }
    return 0;
}

}  // extern "C"

namespace {
// The preprocess's counts (instances, depth-bit range) into pinned host memory.
// dst[0..2] = the counts, then dst[3] = 1 stored with release semantics at
// system scope (a host polling dst[3] then sees the counts).
__global__ void counts_to_host_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
    if (threadIdx.x < 3) dst[threadIdx.x] = src[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(dst + 3, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
}  // namespace

extern "C" {

// seg_h (segment mode, host int64[V+1] or null): view v draws Gaussians
// [seg_h[v], seg_h[v+1]) only.
// allow_split: deep tiles may be blended in segments (the chain, whose blend
// backward re-runs the sequential blend, renders without).
static int render_views(arapgs_ctx* c, const arapgs_render_desc* dv, int V, int* num_rendered_h,
                        const int64_t* seg_h = nullptr, bool allow_split = true) {
    CHECK_CTX(c);
    if (!dv || V < 1 || V > RASTER_MAX_VIEWS) return fail(c, -1, "render: need 1..%d views", RASTER_MAX_VIEWS);
    const arapgs_render_desc* d = dv;
    if (!d->background || !d->means3D || !d->shs || !d->opacities || !d->scales || !d->rotations || d->P < 0 ||
        d->width <= 0 || d->height <= 0)
        return fail(c, -1, "render: bad arguments");
    if (d->D < 0 || d->D > 3) return fail(c, -1, "render: SH degree must be 0..3");
    const int tile = d->tile ? d->tile : 16;
    if (tile != 16 && tile != 8) return fail(c, -1, "render: tile must be 8 or 16");
    for (int v = 0; v < V; v++) {
        const arapgs_render_desc* e = dv + v;
        if (!e->viewmatrix || !e->projmatrix || !e->cam_pos || !e->out_color)
            return fail(c, -1, "render: view %d: bad arguments", v);
        if (e->colors_precomp || e->cov3D_precomp)
            return fail(c, -1, "render: precomputed colours/covariances are not supported");
        if (e->prefiltered) return fail(c, -1, "render: prefiltered mode is not supported");
        if (e->to_ortho && !(e->ortho_scale > 0.0f))
            return fail(c, -1, "render: view %d: orthographic mode needs ortho_scale > 0", v);
        if (v > 0 && (e->P != d->P || e->D != d->D || e->width != d->width || e->height != d->height ||
                      (e->tile ? e->tile : 16) != tile || e->means3D != d->means3D || e->shs != d->shs ||
                      e->opacities != d->opacities || e->scales != d->scales || e->rotations != d->rotations ||
                      e->background != d->background || e->scale_modifier != d->scale_modifier ||
                      (e->rects != nullptr) != (d->rects != nullptr) || e->boxmin_h != d->boxmin_h ||
                      e->boxmax_h != d->boxmax_h))
            return fail(c, -1, "render: view %d differs from view 0 in a shared argument", v);
    }
    if (seg_h) {
        if (seg_h[0] != 0 || seg_h[V] != d->P) return fail(c, -1, "render_segments: segments must run from 0 to P");
        for (int v = 0; v < V; v++)
            if (seg_h[v + 1] < seg_h[v]) return fail(c, -1, "render_segments: segment %d ends before it starts", v);
    }
    (void)hipSetDevice(c->device);
    RasterParams rp;
    rp.seg = nullptr;
    rp.P = d->P; rp.D = d->D; rp.W = d->width; rp.H = d->height; rp.tile = tile; rp.V = V;
    rp.gx = (d->width + tile - 1) / tile; rp.gy = (d->height + tile - 1) / tile;
    rp.bg = d->background; rp.means = d->means3D; rp.shs = d->shs; rp.opac = d->opacities; rp.scales = d->scales;
    rp.rots = d->rotations; rp.scale_mod = d->scale_modifier;
    rp.use_rects = d->rects != nullptr;
    rp.use_box = (d->boxmin_h && d->boxmax_h) ? 1 : 0;
    rp.unroll = c->blend_unroll;
    // the split composite does not mark contributors: renders that ask for them blend tile by tile
    bool any_surface = false;
    for (int v = 0; v < V; v++) any_surface = any_surface || dv[v].surface != nullptr;
    rp.split = (allow_split && !any_surface && tile == 16) ? c->blend_split : 0;
    rp.cull = c->tile_cull ? 1 : 0;
    for (int i = 0; i < 3; i++) {
        rp.boxmin[i] = rp.use_box ? d->boxmin_h[i] : 0.0f;
        rp.boxmax[i] = rp.use_box ? d->boxmax_h[i] : 0.0f;
    }
    RasterView* hv = c->r_views_h;
    for (int v = 0; v < V; v++) {
        const arapgs_render_desc* e = dv + v;
        hv[v].view = e->viewmatrix; hv[v].proj = e->projmatrix; hv[v].campos = e->cam_pos;
        hv[v].tanfovx = e->tan_fovx; hv[v].tanfovy = e->tan_fovy;
        // orthographic (setOrthoCam(ortho_scale, ortho_scale), GaussianView.cpp:1018-1019):
        // x_ndc = x_view / ortho_scale, so the EWA Jacobian is the constant
        // pixels-per-unit diag(W, H) / (2 ortho_scale)
        hv[v].ortho = e->to_ortho ? 1 : 0;
        hv[v].focal_x = e->to_ortho ? (float)e->width / (2.0f * e->ortho_scale) : (float)e->width / (2.0f * e->tan_fovx);
        hv[v].focal_y = e->to_ortho ? (float)e->height / (2.0f * e->ortho_scale)
                                    : (float)e->height / (2.0f * e->tan_fovy);
        hv[v].out = e->out_color; hv[v].surface = e->surface; hv[v].radii = e->radii; hv[v].rects = e->rects;
        hv[v].dout = nullptr;
    }
    // the staging copy completes before this call returns (the stream is
    // synchronised for the instance count below), so the pinned buffer is free again
    hipError_t e = hipMemcpyAsync(c->r_views, hv, sizeof(RasterView) * V, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "render views");
    rp.views = c->r_views;
    if (seg_h) {
        for (int v = 0; v <= V; v++) c->r_seg_h[v] = (int)seg_h[v];
        e = hipMemcpyAsync(c->r_seg, c->r_seg_h, sizeof(int) * (V + 1), hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "render segments");
        rp.seg = c->r_seg;
    }
    const int64_t PV = seg_h ? (int64_t)std::max(d->P, 1) : (int64_t)std::max(d->P, 1) * V;
    int rc = grow(c, &c->r_geo, &c->r_geo_cap, raster_geo_bytes(PV));
    if (!rc) rc = grow(c, &c->r_scan, &c->r_scan_cap, sizeof(uint32_t) * scan_workspace_words(PV));
    const int ntiles = rp.gx * rp.gy;
    if (!rc) rc = grow(c, &c->r_ranges, &c->r_ranges_cap, sizeof(uint2) * (size_t)ntiles * V);
    if (!rc) rc = grow(c, &c->r_rec, &c->r_rec_cap, raster_record_bytes(PV));
    if (rc) return rc;
    rc = check_launch(c, raster_preprocess(rp, c->r_geo, (uint32_t*)c->r_scan, c->r_total, c->stream),
                      "render preprocess");
    if (rc) return rc;
    uint32_t cnt[3] = {0, 0, 0};
    uint32_t* cnt_h = reinterpret_cast<uint32_t*>(c->flag_h + 10);
    __atomic_store_n(cnt_h + 3, 0u, __ATOMIC_SEQ_CST);
    hipLaunchKernelGGL(counts_to_host_kernel, dim3(1), dim3(64), 0, c->stream, (const uint32_t*)c->r_total, cnt_h);
    e = hipGetLastError();
    bool seen = false;
    if (e == hipSuccess && c->host_spin) {
        // poll the ready word (bounded; then the stream wait, which reports a failed launch)
        const auto t0 = std::chrono::steady_clock::now();
        while (!(seen = __atomic_load_n(cnt_h + 3, __ATOMIC_ACQUIRE) != 0u) &&
               std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(50)) {
        }
    }
    if (e == hipSuccess && !seen) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "render");
    for (int i = 0; i < 3; i++) cnt[i] = ((volatile uint32_t*)cnt_h)[i];
    const uint32_t n = cnt[0];
    const uint32_t dmin = n > 0 ? cnt[1] : 0u;
    int dbits = 1;
    while (n > 0 && dbits < 32 && ((uint64_t)(cnt[2] - cnt[1]) >> dbits) != 0) dbits++;
    if (n > 0x7fffffffu) return fail(c, -1, "render: too many tile instances (%u)", n);
    const size_t nn = std::max<size_t>(n, 1);
    rc = grow(c, &c->r_bin, &c->r_bin_cap, nn * 4 * sizeof(uint32_t) + 1024);
    if (!rc) rc = grow(c, &c->r_sort, &c->r_sort_cap, sizeof(uint32_t) * radix_workspace_words((int64_t)nn));
    if (rc) return rc;
    // r_bin: keys (tile ids), keys_alt, vals (record ids), vals_alt; BIN_VALS_AT is the vals offset
    char* b = (char*)c->r_bin;
    uint32_t* keys = (uint32_t*)b;
    uint32_t* keys_alt = (uint32_t*)(b + nn * 4);
    uint32_t* vals = (uint32_t*)(b + nn * BIN_VALS_AT);
    uint32_t* vals_alt = (uint32_t*)(b + nn * 12);
    if (rp.split > 0)
        rc = grow(c, &c->r_split, &c->r_split_cap, raster_split_bytes((int64_t)ntiles * V, n, rp.split));
    if (rc) return rc;
    rc = check_launch(c, raster_bin_and_blend(rp, c->r_geo, n, c->r_total, dmin, dbits, keys, vals, keys_alt,
                                              vals_alt, (uint32_t*)c->r_sort, (uint2*)c->r_ranges, c->r_rec,
                                              (uint32_t*)c->r_scan, rp.split > 0 ? c->r_split : nullptr, c->stream),
                      "render");
    if (rc) return rc;
    c->last_rp = rp;
    c->last_dmin = dmin;
    c->last_dbits = dbits;
    c->last_rendered = n;
    c->last_tiles = ntiles * V;
    if (num_rendered_h) *num_rendered_h = (int)n;
    return 0;
}

int arapgs_set_blend_split(arapgs_ctx* c, int entries) {
    CHECK_CTX(c);
    if (entries != 0 && (entries < 256 || entries % 256 != 0))
        return fail(c, -1, "set_blend_split: 0 or a multiple of 256 (at least 256)");
    c->blend_split = entries;
    return 0;
}

int arapgs_set_blend_unroll(arapgs_ctx* c, int unroll) {
    CHECK_CTX(c);
    if (unroll != 2 && unroll != 4 && unroll != 8) return fail(c, -1, "set_blend_unroll: 2, 4 or 8");
    c->blend_unroll = unroll;
    return 0;
}

int arapgs_render(arapgs_ctx* c, const arapgs_render_desc* d, int* num_rendered_h) {
    return render_views(c, d, 1, num_rendered_h);
}

int arapgs_render_views(arapgs_ctx* c, const arapgs_render_desc* descs, int n_views, int* num_rendered_h) {
    return render_views(c, descs, n_views, num_rendered_h);
}

int arapgs_render_segments(arapgs_ctx* c, const arapgs_render_desc* descs, int n_views, const int64_t* seg_h,
                           int* num_rendered_h) {
    if (!seg_h) {
        CHECK_CTX(c);
        return fail(c, -1, "render_segments: segments required");
    }
    return render_views(c, descs, n_views, num_rendered_h, seg_h);
}

int arapgs_render_debug(arapgs_ctx* c, uint64_t* keys_h, uint32_t* vals_h, int64_t cap, uint32_t* ranges_h,
                        int64_t ranges_cap) {
    CHECK_CTX(c);
    const size_t nn = std::max<size_t>(c->last_rendered, 1);
    const int64_t n = std::min<int64_t>(cap, c->last_rendered);
    char* b = (char*)c->r_bin;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && keys_h && n > 0) {
        // the sorted keys hold the tile id; the reference layout is tile << 32 | depth bits of the value's record
        std::vector<uint32_t> v((size_t)n), tile((size_t)n);
        const int64_t PV = c->last_rp.seg ? (int64_t)std::max(c->last_rp.P, 1)
                                          : (int64_t)std::max(c->last_rp.P, 1) * c->last_rp.V;
        std::vector<uint32_t> depth((size_t)PV);
        e = hipMemcpy(tile.data(), b, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(v.data(), b + nn * BIN_VALS_AT, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(depth.data(), c->r_geo, sizeof(uint32_t) * PV, hipMemcpyDeviceToHost);
        for (int64_t i = 0; e == hipSuccess && i < n; i++) keys_h[i] = ((uint64_t)tile[i] << 32) | depth[v[i]];
    }
    if (e == hipSuccess && vals_h && n > 0)
        e = hipMemcpy(vals_h, b + nn * BIN_VALS_AT, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
    const int64_t nr = std::min<int64_t>(ranges_cap, c->last_tiles);
    if (e == hipSuccess && ranges_h && nr > 0) e = hipMemcpy(ranges_h, c->r_ranges, sizeof(uint2) * nr, hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : hip_fail(c, e, "render_debug");
}

}  // extern "C"

struct arapgs_grid {
    int G = 0;
    int padding = 1;
    int64_t version = 0;  // incremented by every re-binning
    int64_t n_pairs = 0, n_valid = 0;
    int64_t n_gauss = 0;  // Gaussians the grid was built for
    int64_t* prefix = nullptr;
    int* list = nullptr;
    int* valid = nullptr;
    float* samples = nullptr;
    // adaptive low-pass filter per valid cell (6 floats each; null: lpf I)
    float* ada = nullptr;
    bool ada_on = false;
    int device = 0;
};

namespace {
// The density filter of a grid: its adaptive matrices when set, else lpf I.
DensityFilter grid_filter(const arapgs_grid* g, float lpf) {
    return DensityFilter{lpf, (g && g->ada_on) ? g->ada : nullptr};
}
}  // namespace

extern "C" {

int arapgs_scene_aabb(arapgs_ctx* c, const float* pos, int64_t n, int grid_num, float* out7_h) {
    CHECK_CTX(c);
    if (!pos || n <= 0 || grid_num <= 0 || !out7_h) return fail(c, -1, "scene_aabb: bad arguments");
    (void)hipSetDevice(c->device);
    uint32_t* ws = nullptr;
    float* d6 = nullptr;
    hipError_t e = hipMalloc(&ws, 64);
    if (e == hipSuccess) e = hipMalloc(&d6, 64);
    if (e != hipSuccess) return hip_fail(c, e, "scene_aabb");
    int rc = check_launch(c, launch_minmax(pos, n, ws, d6, c->stream), "scene_aabb");
    float mm[6];
    if (!rc) {
        e = hipMemcpyAsync(mm, d6, sizeof(mm), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(c, e, "scene_aabb");
    }
    (void)hipFree(ws);
    (void)hipFree(d6);
    if (rc) return rc;
    // expansion of getOverallAABB in float, then the step of getGridSamples
    float step = 0.0f;
    for (int k = 0; k < 3; k++) {
        const float lo0 = mm[k], hi0 = mm[3 + k];
        const float mean = (hi0 + lo0) / 2.0f;
        const float lo = (lo0 - mean) * 1.1f + mean;
        const float hi = ((hi0 - mean) * 1.1f + mean - lo) * 1.0078125f + lo;
        out7_h[k] = std::min(lo, -0.75f);
        out7_h[3 + k] = std::max(hi, 0.75f);
        step = std::max(step, (out7_h[3 + k] - out7_h[k]) / (float)grid_num);
    }
    out7_h[6] = step;
    return 0;
}

int arapgs_grid_reorder(arapgs_ctx* c, const float* pos, int64_t n, const float* aabb7_h, int grid_num,
                        uint32_t* order) {
    CHECK_CTX(c);
    if (!pos || !aabb7_h || !order || n < 0 || grid_num <= 0) return fail(c, -1, "grid_reorder: bad arguments");
    if (n == 0) return 0;
    GridGeom gg{{aabb7_h[0], aabb7_h[1], aabb7_h[2]}, aabb7_h[6], grid_num, 0};
    void* buf = nullptr;
    const size_t words = radix_workspace_words(n);
    hipError_t e = hipMalloc(&buf, (size_t)n * 24 + words * 4);
    if (e != hipSuccess) return hip_fail(c, e, "grid_reorder");
    uint64_t* keys = (uint64_t*)buf;
    uint64_t* keys_alt = keys + n;
    uint32_t* vals_alt = (uint32_t*)(keys_alt + n);
    uint32_t* ws = vals_alt + n;
    int bits = 1;
    while ((1ll << bits) < (int64_t)grid_num * grid_num * grid_num) bits++;
    int rc = check_launch(c, launch_center_cells(pos, n, gg, keys, order, c->stream), "grid_reorder");
    if (!rc) rc = check_launch(c, launch_radix_sort(keys, order, keys_alt, vals_alt, n, bits, ws, c->stream), "grid_reorder");
    e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    if (!rc && e != hipSuccess) rc = hip_fail(c, e, "grid_reorder");
    return rc;
}

int arapgs_gaussian_aabbs(arapgs_ctx* c, const float* pos, const float* rot, const float* scale, const float* opacity,
                          int64_t n, float* aabbs) {
    CHECK_CTX(c);
    if (!pos || !rot || !scale || !opacity || !aabbs || n < 0) return fail(c, -1, "gaussian_aabbs: bad arguments");
    return check_launch(c, launch_aabbs(pos, rot, scale, opacity, n, aabbs, c->stream), "gaussian_aabbs");
}

}  // extern "C"

namespace {
// The cell lists of GetBoxesGsGrid (cudakdtree.cu:425-451) for Gaussian boxes
// `aabbs` on the grid gg: prefix (int64[G^3], inclusive per-cell counts) and
// flags (u32[G^3], 1 where a cell holds any Gaussian) are caller buffers; the
// list (int[pairs], every cell's Gaussians ascending) is allocated here and
// returned in *list_out. The pairs are emitted and sorted in chunks of whole
// x-slabs (cells x*G^2 .. (x+1)*G^2-1) whose pair counts fit 32-bit offsets.
int build_cell_lists(arapgs_ctx* c, const char* what, const float* aabbs, int64_t n, const GridGeom& gg,
                     int64_t* prefix, uint32_t* flags, int** list_out, uint64_t* npairs_out) {
    *list_out = nullptr;
    *npairs_out = 0;
    const int G = gg.G;
    const int64_t G2 = (int64_t)G * G;
    const int64_t gc = G2 * G;
    std::vector<void*> tmp;
    auto dmal = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) return nullptr;
        tmp.push_back(p);
        return p;
    };
    auto cleanup = [&]() { for (void* p : tmp) (void)hipFree(p); };
    auto sync = [&]() -> int {
        hipError_t e = hipStreamSynchronize(c->stream);
        return e == hipSuccess ? 0 : hip_fail(c, e, what);
    };
    // 1. pairs per x-slab, and from them the chunks of whole slabs whose pairs
    //    fit the 32-bit offsets of one emit + sort pass
    unsigned long long* slab = (unsigned long long*)dmal(8 * (size_t)G);
    uint32_t* cell_count = (uint32_t*)dmal(4 * (size_t)gc);
    uint32_t* total = (uint32_t*)dmal(8);
    uint32_t* prefix_ws = (uint32_t*)dmal(4 * cell_prefix_workspace_words(gc));
    uint32_t* scan_ws = (uint32_t*)dmal(4 * scan_workspace_words(n));
    if (!slab || !cell_count || !total || !prefix_ws || !scan_ws) {
        cleanup();
        return fail(c, -2, "%s: allocation failed", what);
    }
    int rc = check_launch(c, launch_slab_counts(aabbs, n, gg, slab, c->stream), what);
    std::vector<unsigned long long> slab_h((size_t)G, 0ull);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(slab_h.data(), slab, 8 * (size_t)G, hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) rc = hip_fail(c, e, what);
        else rc = sync();
    }
    std::vector<std::pair<int, int>> chunks;  // [x0, x1)
    uint64_t npairs = 0, max_chunk = 0;
    const uint64_t chunk_pairs = c->grid_chunk_pairs ? c->grid_chunk_pairs : GRID_CHUNK_PAIRS;
    if (!rc) {
        uint64_t cur = 0;
        int x0 = 0;
        for (int x = 0; x < G; x++) {
            const uint64_t p = slab_h[x];
            if (p > GRID_CHUNK_PAIRS_MAX) {
                rc = fail(c, -1, "%s: slab %d holds %llu cell-Gaussian pairs (at most %llu per slab)", what, x,
                          (unsigned long long)p, (unsigned long long)GRID_CHUNK_PAIRS_MAX);
                break;
            }
            if (x > x0 && cur + p > chunk_pairs) {
                chunks.push_back({x0, x});
                max_chunk = std::max(max_chunk, cur);
                x0 = x;
                cur = 0;
            }
            cur += p;
            npairs += p;
        }
        if (!rc) {
            chunks.push_back({x0, G});
            max_chunk = std::max(max_chunk, cur);
        }
    }
    // 2. per chunk: counts, offsets, pairs keyed by the cell within the chunk,
    //    a stable radix sort (pairs leave the emitter in Gaussian order, so every
    //    cell's list comes out ascending) and the values appended to the list
    int* list = nullptr;
    if (!rc) {
        const size_t np = std::max<uint64_t>(max_chunk, 1);
        uint32_t* cnt = (uint32_t*)dmal(4 * (size_t)n);
        uint32_t* offs = (uint32_t*)dmal(4 * (size_t)n);
        uint32_t* keys = (uint32_t*)dmal(4 * np * 2);
        uint32_t* vals = (uint32_t*)dmal(4 * np * 2);
        uint32_t* sort_ws = (uint32_t*)dmal(4 * radix_workspace_words((int64_t)np));
        if (!cnt || !offs || !keys || !vals || !sort_ws ||
            hipMalloc(&list, 4 * std::max<uint64_t>(npairs, 1)) != hipSuccess) {
            list = nullptr;
            rc = fail(c, -2, "%s: allocation failed (%llu cell-Gaussian pairs)", what, (unsigned long long)npairs);
        }
        if (!rc) rc = check_launch(c, launch_zero_cells(cell_count, gc, c->stream), what);
        uint64_t base = 0;
        for (size_t k = 0; !rc && k < chunks.size(); k++) {
            const int x0 = chunks[k].first, x1 = chunks[k].second;
            rc = check_launch(c, launch_box_count_slabs(aabbs, n, gg, x0, x1, cnt, c->stream), what);
            if (!rc) rc = check_launch(c, launch_exclusive_scan(cnt, n, offs, total, scan_ws, c->stream), what);
            uint32_t cp = 0;
            if (!rc) {
                hipError_t e = hipMemcpyAsync(&cp, total, 4, hipMemcpyDeviceToHost, c->stream);
                if (e != hipSuccess) rc = hip_fail(c, e, what);
                else rc = sync();
            }
            uint64_t want = 0;
            for (int x = x0; x < x1; x++) want += slab_h[x];
            if (!rc && (uint64_t)cp != want)
                rc = fail(c, -1, "%s: chunk %d..%d counts %u pairs, slabs %llu", what, x0, x1, cp,
                          (unsigned long long)want);
            if (!rc && cp > 0) {
                int bits = 1;
                while ((1ll << bits) < (int64_t)(x1 - x0) * G2) bits++;
                rc = check_launch(c, launch_box_emit_slabs(aabbs, n, gg, x0, x1, offs, keys, vals, cell_count,
                                                           c->stream), what);
                if (!rc)
                    rc = check_launch(c, launch_radix_sort32(keys, vals, keys + np, vals + np, cp, bits, sort_ws,
                                                             c->stream), what);
                if (!rc) {
                    hipError_t e = hipMemcpyAsync(list + base, vals, 4 * (size_t)cp, hipMemcpyDeviceToDevice,
                                                  c->stream);
                    if (e != hipSuccess) rc = hip_fail(c, e, what);
                }
            }
            base += cp;
        }
    }
    // 3. the 64-bit per-cell prefix and the occupancy flags
    if (!rc) rc = check_launch(c, launch_cell_prefix(cell_count, gc, prefix, flags, prefix_ws, c->stream), what);
    if (!rc) rc = sync();
    cleanup();
    if (rc) {
        if (list) (void)hipFree(list);
        return rc;
    }
    *list_out = list;
    *npairs_out = npairs;
    return 0;
}
}  // namespace

extern "C" {

int arapgs_grid_build(arapgs_ctx* c, const float* aabbs, int64_t n, const float* aabb7_h, int grid_num, int padding,
                      arapgs_grid** out) {
    CHECK_CTX(c);
    if (!aabbs || !aabb7_h || !out || n <= 0 || grid_num <= 0 || grid_num > 1024)
        return fail(c, -1, "grid_build: bad arguments");
    if (n > 0x7fffffffLL) return fail(c, -1, "grid_build: at most 2^31-1 Gaussians");
    *out = nullptr;
    (void)hipSetDevice(c->device);
    const int64_t gc = (int64_t)grid_num * grid_num * grid_num;
    GridGeom gg{{aabb7_h[0], aabb7_h[1], aabb7_h[2]}, aabb7_h[6], grid_num, padding};
    arapgs_grid* g = new arapgs_grid();
    g->G = grid_num;
    g->padding = padding;
    g->n_gauss = n;
    g->device = c->device;
    uint32_t *flags = nullptr, *flag_ex = nullptr, *total = nullptr, *scan_ws = nullptr;
    auto cleanup = [&]() {
        for (void* p : {(void*)flags, (void*)flag_ex, (void*)total, (void*)scan_ws})
            if (p) (void)hipFree(p);
    };
    if (hipMalloc(&flags, 4 * (size_t)gc) != hipSuccess || hipMalloc(&flag_ex, 4 * (size_t)gc) != hipSuccess ||
        hipMalloc(&total, 8) != hipSuccess || hipMalloc(&scan_ws, 4 * scan_workspace_words(gc)) != hipSuccess ||
        hipMalloc(&g->prefix, 8 * (size_t)gc) != hipSuccess) {
        cleanup();
        arapgs_grid_destroy(g);
        return fail(c, -2, "grid_build: allocation failed");
    }
    uint64_t npairs = 0;
    int rc = build_cell_lists(c, "grid_build", aabbs, n, gg, g->prefix, flags, &g->list, &npairs);
    // valid cells and their samples
    if (!rc) rc = check_launch(c, launch_exclusive_scan(flags, gc, flag_ex, total + 1, scan_ws, c->stream), "grid_build");
    uint32_t nvalid = 0;
    if (!rc) {
        hipError_t e = hipMemcpyAsync(&nvalid, total + 1, 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(c, e, "grid_build");
    }
    if (!rc) {
        if (hipMalloc(&g->valid, 4 * std::max<size_t>(nvalid, 1)) != hipSuccess ||
            hipMalloc(&g->samples, 12 * 64 * std::max<size_t>(nvalid, 1)) != hipSuccess)
            rc = fail(c, -2, "grid_build: allocation failed");
        if (!rc) rc = check_launch(c, launch_compact_valid(flags, flag_ex, gc, g->valid, c->stream), "grid_build");
        if (!rc) rc = check_launch(c, launch_samples(g->valid, nvalid, gg, g->samples, c->stream), "grid_build");
        hipError_t e = hipStreamSynchronize(c->stream);
        if (!rc && e != hipSuccess) rc = hip_fail(c, e, "grid_build");
    }
    cleanup();
    if (rc) {
        arapgs_grid_destroy(g);
        return rc;
    }
    g->n_pairs = (int64_t)npairs;
    g->n_valid = nvalid;
    *out = g;
    return 0;
}

int arapgs_grid_rebin(arapgs_ctx* c, arapgs_grid* g, const float* aabbs, int64_t n, const float* aabb7_h) {
    CHECK_CTX(c);
    if (!g || !aabbs || !aabb7_h || n != g->n_gauss) return fail(c, -1, "grid_rebin: bad arguments");
    if (g->device != c->device) return fail(c, -1, "grid_rebin: grid of another device");
    (void)hipSetDevice(c->device);
    const int64_t gc = (int64_t)g->G * g->G * g->G;
    GridGeom gg{{aabb7_h[0], aabb7_h[1], aabb7_h[2]}, aabb7_h[6], g->G, g->padding};
    uint32_t* flags = nullptr;
    int64_t* prefix = nullptr;
    if (hipMalloc(&flags, 4 * (size_t)gc) != hipSuccess || hipMalloc(&prefix, 8 * (size_t)gc) != hipSuccess) {
        if (flags) (void)hipFree(flags);
        return fail(c, -2, "grid_rebin: allocation failed");
    }
    int* list = nullptr;
    uint64_t npairs = 0;
    // new prefix and list; the grid keeps the old ones until both are complete
    int rc = build_cell_lists(c, "grid_rebin", aabbs, n, gg, prefix, flags, &list, &npairs);
    (void)hipFree(flags);
    if (rc) {
        (void)hipFree(prefix);
        return rc;
    }
    (void)hipFree(g->list);
    (void)hipFree(g->prefix);
    g->list = list;
    g->prefix = prefix;
    g->n_pairs = (int64_t)npairs;
    g->version++;
    return 0;
}

void arapgs_grid_destroy(arapgs_grid* g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    for (void* p : {(void*)g->prefix, (void*)g->list, (void*)g->valid, (void*)g->samples, (void*)g->ada})
        if (p) (void)hipFree(p);
    delete g;
}

int arapgs_grid_set_ada_lpf(arapgs_ctx* c, arapgs_grid* g, const float* samples, float lpf_param) {
    CHECK_CTX(c);
    if (!g || !(lpf_param >= 0.0f)) return fail(c, -1, "grid_set_ada_lpf: bad arguments");
    if (lpf_param == 0.0f) {
        g->ada_on = false;
        return 0;
    }
    (void)hipSetDevice(c->device);
    if (!g->ada && g->n_valid > 0) {
        // the stream may still run density calls that read the scalar filter
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMalloc((void**)&g->ada, sizeof(float) * 6 * (size_t)g->n_valid);
        if (e != hipSuccess) return hip_fail(c, e, "grid_set_ada_lpf");
    }
    int rc = check_launch(c, launch_ada_lpf(g->n_valid, samples ? samples : g->samples, lpf_param, g->ada, c->stream),
                          "grid_set_ada_lpf");
    if (!rc) g->ada_on = true;
    return rc;
}

int arapgs_grid_ada_lpf(const arapgs_grid* g, const float** mats) {
    if (!g || !mats) return -1;
    *mats = g->ada_on ? g->ada : nullptr;
    return 0;
}

int arapgs_grid_info(const arapgs_grid* g, int64_t* n_pairs, int64_t* n_valid, const int64_t** prefix, const int** list,
                     const int** valid, const float** samples) {
    if (!g) return -1;
    if (n_pairs) *n_pairs = g->n_pairs;
    if (n_valid) *n_valid = g->n_valid;
    if (prefix) *prefix = g->prefix;
    if (list) *list = g->list;
    if (valid) *valid = g->valid;
    if (samples) *samples = g->samples;
    return 0;
}

int arapgs_density_backward(arapgs_ctx* c, const arapgs_grid* g, const float* samples, const float* pos,
                            const float* rot, const float* scale, const float* opacity, const float* sh, float lpf,
                            const float* dfeat, const float* dopac, float* dpos, float* drot, float* dscale,
                            float* dopacity, float* dsh, float* dsamples) {
    CHECK_CTX(c);
    if (!g || !pos || !rot || !scale || !opacity || !sh || !dpos || !drot || !dscale || !dopacity || !dsh)
        return fail(c, -1, "density_backward: bad arguments");
    if ((uintptr_t)sh % 16) return fail(c, -1, "density_backward: sh must be 16-byte aligned");
    (void)hipSetDevice(c->device);
    const int64_t N = g->n_gauss;  // the Gaussians the grid was built for
    int rc = grow(c, &c->db_ws, &c->db_cap, sizeof(float) * 58 * (size_t)std::max<int64_t>(N, 1));
    if (rc) return rc;
    float* gd = (float*)c->db_ws;
    rc = check_launch(c, launch_zero_f32(gd, 58 * N, c->stream), "density_backward");
    if (!rc && (dfeat || dopac))
        rc = check_launch(c, launch_density_backward(g->n_valid, g->valid, g->prefix, g->list,
                                                     samples ? samples : g->samples, pos, rot, scale, opacity, sh,
                                                     grid_filter(g, lpf), dfeat, dopac, dsamples, gd, N, c->stream),
                          "density_backward");
    else if (!rc && dsamples)
        rc = check_launch(c, launch_zero_f32(dsamples, 3 * 64 * g->n_valid, c->stream), "density_backward");
    if (!rc)
        rc = check_launch(c, launch_density_attr_grad(N, rot, scale, gd, grid_filter(g, lpf), dpos, drot, dscale,
                                                      dopacity, dsh, c->stream),
                          "density_backward");
    return rc;
}

int arapgs_density_sample(arapgs_ctx* c, const arapgs_grid* g, const float* pos, const float* rot, const float* scale,
                          const float* opacity, const float* sh, float lpf, float* feat, float* opac) {
    CHECK_CTX(c);
    if (!g || !pos || !rot || !scale || !opacity || !sh || !feat || !opac) return fail(c, -1, "density_sample: bad arguments");
    return check_launch(c, launch_density(g->n_valid, g->valid, g->prefix, g->list, g->samples, pos, rot, scale,
                                          opacity, sh, grid_filter(g, lpf), feat, opac, c->stream), "density_sample");
}

}  // extern "C"

extern "C" {

int arapgs_aims_translate(arapgs_ctx* c, const float* node_pos, int m, const int* active, int n_active,
                          const float* delta_h, float* aims) {
    CHECK_CTX(c);
    if (!node_pos || !aims || !delta_h || m <= 0 || (n_active > 0 && !active)) return fail(c, -1, "aims_translate: bad arguments");
    return check_launch(c, launch_aims_translate(node_pos, m, active, n_active, delta_h, aims, c->stream), "aims_translate");
}

int arapgs_aims_twist(arapgs_ctx* c, const float* node_pos, int m, const int* active, int n_active, const float* axis_h,
                      float radian, float* aims) {
    CHECK_CTX(c);
    if (!node_pos || !aims || !axis_h || m <= 0 || n_active <= 0 || !active) return fail(c, -1, "aims_twist: bad arguments");
    return check_launch(c, launch_aims_twist(node_pos, m, active, n_active, axis_h, radian, aims, c->stream), "aims_twist");
}

int arapgs_aims_scale(arapgs_ctx* c, const float* node_pos, int m, const int* active, int n_active, float s,
                      float* aims) {
    CHECK_CTX(c);
    if (!node_pos || !aims || m <= 0 || n_active <= 0 || !active) return fail(c, -1, "aims_scale: bad arguments");
    return check_launch(c, launch_aims_scale(node_pos, m, active, n_active, s, aims, c->stream), "aims_scale");
}

int arapgs_aims_set(arapgs_ctx* c, const float* node_pos, int m, const int* idx, int n, const float* vals,
                    float* aims) {
    CHECK_CTX(c);
    if (!node_pos || !aims || m <= 0 || n < 0 || (n > 0 && (!idx || !vals))) return fail(c, -1, "aims_set: bad arguments");
    return check_launch(c, launch_aims_set(node_pos, m, idx, n, vals, aims, c->stream), "aims_set");
}

int arapgs_static_rows(arapgs_ctx* c, const uint8_t* node_static, const uint32_t* idx, int64_t n, int len,
                       uint8_t* out) {
    CHECK_CTX(c);
    if (n < 0 || len <= 0 || (n > 0 && (!node_static || !idx || !out))) return fail(c, -1, "static_rows: bad arguments");
    return check_launch(c, launch_static_rows(node_static, idx, n, len, out, c->stream), "static_rows");
}

namespace {
// Start / end events of one profiled launch (null when profiling is off).
std::pair<hipEvent_t, hipEvent_t>* prof_begin(arapgs_ctx* c, int which) {
    if (!c->prof) return nullptr;
    auto& list = c->prof_events[which];
    size_t& used = c->prof_used[which];
    if (used == list.size() && list.size() < 4096) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) list.push_back({a, b});
    }
    std::pair<hipEvent_t, hipEvent_t>* ev = used < list.size() ? &list[used++] : nullptr;
    if (ev) (void)hipEventRecord(ev->first, c->stream);
    return ev;
}

void prof_end(arapgs_ctx* c, std::pair<hipEvent_t, hipEvent_t>* ev) {
    if (ev) (void)hipEventRecord(ev->second, c->stream);
}
}  // namespace

int arapgs_deform_frame(arapgs_ctx* c, arapgs_solver* s, const arapgs_frame_buffers* fb, const float* aims,
                        int gn_iters, int stop_rule, double* energies, int* iters_done) {
    CHECK_CTX(c);
    if (!s || !fb || !aims || !energies) return fail(c, -1, "deform_frame: bad arguments");
    const int m = s->w.dev.M, k = s->w.dev.K;
    if (fb->m != m || fb->k != k) return fail(c, -1, "deform_frame: buffers do not match the solver (m, k)");
    int rc = arapgs_gn_solve(c, s, fb->node_pos, aims, gn_iters, stop_rule, nullptr, energies, iters_done);
    if (rc) return rc;
    rc = check_launch(c, launch_build_node_table(m, fb->node_pos, s->w.x, s->w.dev.free_id, s->table, c->stream),
                      "deform_frame table");
    // samples first: they use the current node positions (GaussianView.cpp:1500-1504)
    std::pair<hipEvent_t, hipEvent_t>* ev = nullptr;
    if (!rc && fb->s > 0) {
        ev = prof_begin(c, 1);
        rc = check_launch(c, launch_blend_points(fb->samples, fb->s, fb->sample_idx, fb->sample_w, k, s->table, m,
                                                 fb->sample_static, fb->samples, c->stream, c->node_cache),
                          "deform_frame samples");
        prof_end(c, ev);
    }
    ev = rc ? nullptr : prof_begin(c, 0);
    if (!rc)
        rc = check_launch(c, launch_six_point(fb->n, k, s->table, m, fb->ends_idx, fb->ends_w, fb->ends, fb->scale0,
                                              fb->gs_static, fb->axis_padding, fb->end_coeff, fb->pos, fb->rot,
                                              fb->scale, fb->sh, c->shrot_dev, c->stream, c->node_cache),
                          "deform_frame six-point");
    prof_end(c, ev);
    if (!rc)
        rc = check_launch(c, launch_blend_points(fb->node_pos, m, fb->node_src_idx, fb->node_src_w, k, s->table, m,
                                                 nullptr, s->node_tmp, c->stream, c->node_cache),
                          "deform_frame nodes");
    if (!rc) rc = check_launch(c, launch_copy_f32(fb->node_pos, s->node_tmp, 3 * (int64_t)m, c->stream), "deform_frame");
    if (!rc && fb->s > 0 && fb->sample_feat) {
        rc = check_launch(c, launch_node_quats(m, s->table, s->node_q, c->stream), "deform_frame quats");
        ev = rc ? nullptr : prof_begin(c, 2);
        if (!rc)
            rc = check_launch(c, launch_sample_sh(fb->s, k, fb->sample_w, fb->sample_idx, s->node_q, fb->sample_static,
                                                  fb->sample_feat, c->shrot_dev, c->stream), "deform_frame sample SH");
        prof_end(c, ev);
    }
    return rc;
}

}  // extern "C"

extern "C" {

int arapgs_profile_enable(arapgs_ctx* c, int on) {
    CHECK_CTX(c);
    c->prof = on != 0;
    for (size_t& u : c->prof_used) u = 0;
    return 0;
}

int arapgs_profile_read_kernel(arapgs_ctx* c, int which, float* ms_h, int cap) {
    CHECK_CTX(c);
    if (which < 0 || which > 2 || (cap > 0 && !ms_h)) return fail(c, -1, "profile_read: bad arguments");
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "profile_read");
    int n = 0;
    auto& list = c->prof_events[which];
    for (size_t i = 0; i < c->prof_used[which] && n < cap; i++) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, list[i].first, list[i].second) == hipSuccess) ms_h[n++] = ms;
    }
    c->prof_used[which] = 0;
    return n;
}

int arapgs_profile_read(arapgs_ctx* c, float* ms_h, int cap) { return arapgs_profile_read_kernel(c, 0, ms_h, cap); }

}  // extern "C"

// ---------------------------------------------------------------- adjoint chain
struct arapgs_chain {
    int device = 0;
    int64_t N = 0, S = 0, nv = 0;
    int K = 0, M = 0, V = 0, W = 0, H = 0, D = 3;
    arapgs_chain_desc d{};
    std::vector<const float*> view, proj, campos;
    std::vector<float> tfx, tfy;
    const arapgs_grid* grid = nullptr;  // its lists are read at every call (a re-binning replaces them)
    const float* samples0 = nullptr;
    // deformed state and its inputs
    float *ends0 = nullptr, *ends = nullptr, *pos = nullptr, *rot = nullptr, *scale = nullptr, *sh = nullptr;
    float* samples = nullptr;
    float* img = nullptr;  // V*3*H*W render scratch of the backward
    // gradients
    float *grec = nullptr, *gout = nullptr, *gdens = nullptr, *gends = nullptr, *gsamp = nullptr;
    // node -> (point, slot) inverse index
    uint32_t* entries = nullptr;
    uint2* rng = nullptr;
    int64_t n_end_entries = 0, n_entries = 0;
    std::vector<void*> allocs;
};

namespace {

template <typename T>
int chain_alloc(arapgs_ctx* c, arapgs_chain* ch, T** p, size_t count) {
    hipError_t e = hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1));
    if (e != hipSuccess) { *p = nullptr; return hip_fail(c, e, "chain allocation"); }
    ch->allocs.push_back((void*)*p);
    return 0;
}

void chain_free(arapgs_chain* ch) {
    if (!ch) return;
    (void)hipSetDevice(ch->device);
    for (void* p : ch->allocs) (void)hipFree(p);
    delete ch;
}

// Deformed state of the chain from (table, q0, s0): end points, copies of the
// rest attributes, the fused blend + six-point update and the sample blend.
int chain_deform(arapgs_ctx* c, arapgs_chain* ch, const arapgs_node_xf* table, const float* rot0,
                 const float* scale0) {
    const arapgs_chain_desc& d = ch->d;
    hipStream_t st = c->stream;
    int rc = check_launch(c, launch_end_points(d.pos0, rot0, scale0, ch->N, d.axis_padding, d.end_coeff, ch->ends0, st),
                          "chain end points");
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(ch->ends, ch->ends0, sizeof(float) * 18 * ch->N, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ch->pos, d.pos0, sizeof(float) * 3 * ch->N, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ch->rot, rot0, sizeof(float) * 4 * ch->N, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ch->scale, scale0, sizeof(float) * 3 * ch->N, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ch->sh, d.sh0, sizeof(float) * 48 * ch->N, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return hip_fail(c, e, "chain copies");
    rc = check_launch(c, launch_six_point(ch->N, ch->K, (const NodeXf*)table, ch->M, d.ends_idx, d.ends_w, ch->ends,
                                          scale0, nullptr, d.axis_padding, d.end_coeff, ch->pos, ch->rot, ch->scale,
                                          ch->sh, c->shrot_dev, st, c->node_cache),
                      "chain six-point");
    if (rc) return rc;
    if (ch->S > 0)
        rc = check_launch(c, launch_blend_points(ch->samples0, ch->S, d.sample_idx, d.sample_w, ch->K,
                                                 (const NodeXf*)table, ch->M, nullptr, ch->samples, st, c->node_cache),
                          "chain sample blend");
    return rc;
}

int chain_render(arapgs_ctx* c, arapgs_chain* ch, float* const* outs) {
    std::vector<arapgs_render_desc> rd((size_t)ch->V);
    for (int v = 0; v < ch->V; v++) {
        arapgs_render_desc& r = rd[v];
        r = arapgs_render_desc{};
        r.P = (int)ch->N; r.D = ch->D; r.M = 16; r.background = ch->d.background;
        r.width = ch->W; r.height = ch->H;
        r.means3D = ch->pos; r.shs = ch->sh; r.opacities = ch->d.opacity; r.scales = ch->scale;
        r.scale_modifier = 1.0f; r.rotations = ch->rot;
        r.viewmatrix = ch->view[v]; r.projmatrix = ch->proj[v]; r.cam_pos = ch->campos[v];
        r.tan_fovx = ch->tfx[v]; r.tan_fovy = ch->tfy[v];
        r.out_color = outs[v];
        r.ortho_scale = 1.0f; r.tile = 16;
    }
    return render_views(c, rd.data(), ch->V, nullptr, nullptr, false);
}

}  // namespace

extern "C" {

int arapgs_chain_create(arapgs_ctx* c, const arapgs_chain_desc* d, arapgs_chain** out) {
    CHECK_CTX(c);
    if (!d || !out) return fail(c, -1, "chain_create: bad arguments");
    *out = nullptr;
    if (d->n < 1 || d->k < 1 || d->k > KNN_MAX_T || d->m < 1 || !d->pos0 || !d->sh0 || !d->opacity || !d->ends_idx ||
        !d->ends_w || !d->background)
        return fail(c, -1, "chain_create: bad arguments");
    if (d->n > 0x7fffffff) return fail(c, -1, "chain_create: too many Gaussians");
    if (d->n_views < 1 || d->n_views > RASTER_MAX_VIEWS || d->width < 1 || d->height < 1 || d->sh_degree < 0 ||
        d->sh_degree > 3 || !d->viewmatrix_h || !d->projmatrix_h || !d->cam_pos_h || !d->tan_fovx_h ||
        !d->tan_fovy_h)
        return fail(c, -1, "chain_create: bad camera arguments");
    if (d->grid && (!d->sample_idx || !d->sample_w)) return fail(c, -1, "chain_create: a grid needs sample ids/weights");
    if ((uintptr_t)d->sh0 % 16) return fail(c, -1, "chain_create: sh0 must be 16-byte aligned");
    (void)hipSetDevice(c->device);
    arapgs_chain* ch = new arapgs_chain();
    ch->device = c->device;
    ch->d = *d;
    ch->N = d->n; ch->K = d->k; ch->M = d->m; ch->V = d->n_views; ch->W = d->width; ch->H = d->height;
    ch->D = d->sh_degree;
    for (int v = 0; v < ch->V; v++) {
        if (!d->viewmatrix_h[v] || !d->projmatrix_h[v] || !d->cam_pos_h[v]) {
            chain_free(ch);
            return fail(c, -1, "chain_create: view %d: bad camera pointers", v);
        }
        ch->view.push_back(d->viewmatrix_h[v]);
        ch->proj.push_back(d->projmatrix_h[v]);
        ch->campos.push_back(d->cam_pos_h[v]);
        ch->tfx.push_back(d->tan_fovx_h[v]);
        ch->tfy.push_back(d->tan_fovy_h[v]);
    }
    ch->d.viewmatrix_h = nullptr; ch->d.projmatrix_h = nullptr; ch->d.cam_pos_h = nullptr;
    ch->d.tan_fovx_h = nullptr; ch->d.tan_fovy_h = nullptr;
    if (d->grid) {
        ch->nv = d->grid->n_valid;
        ch->S = 64 * ch->nv;
        ch->grid = d->grid;
        ch->samples0 = d->grid->samples;
    }
    ch->n_end_entries = 6 * ch->N * ch->K;
    ch->n_entries = ch->n_end_entries + ch->S * ch->K;
    if (ch->n_entries > 0xffffffffll) {
        chain_free(ch);
        return fail(c, -1, "chain_create: too many (point, neighbour) entries");
    }
    const size_t N = (size_t)ch->N, S = (size_t)ch->S, V = (size_t)ch->V;
    int rc = 0;
    rc = rc ? rc : chain_alloc(c, ch, &ch->ends0, 18 * N);
    rc = rc ? rc : chain_alloc(c, ch, &ch->ends, 18 * N);
    rc = rc ? rc : chain_alloc(c, ch, &ch->pos, 3 * N);
    rc = rc ? rc : chain_alloc(c, ch, &ch->rot, 4 * N);
    rc = rc ? rc : chain_alloc(c, ch, &ch->scale, 3 * N);
    rc = rc ? rc : chain_alloc(c, ch, &ch->sh, 48 * N);
    rc = rc ? rc : chain_alloc(c, ch, &ch->samples, 3 * S);
    rc = rc ? rc : chain_alloc(c, ch, &ch->img, V * 3 * (size_t)ch->W * ch->H);
    rc = rc ? rc : chain_alloc(c, ch, &ch->grec, V * N * 8);
    rc = rc ? rc : chain_alloc(c, ch, &ch->gout, 58 * N);
    rc = rc ? rc : chain_alloc(c, ch, &ch->gdens, 58 * N);
    rc = rc ? rc : chain_alloc(c, ch, &ch->gends, 18 * N);
    rc = rc ? rc : chain_alloc(c, ch, &ch->gsamp, 3 * S);
    rc = rc ? rc : chain_alloc(c, ch, &ch->entries, (size_t)ch->n_entries);
    rc = rc ? rc : chain_alloc(c, ch, &ch->rng, (size_t)ch->M);
    uint32_t* ws = nullptr;
    if (!rc) {
        hipError_t e = hipMalloc((void**)&ws, sizeof(uint32_t) * inverse_index_words(ch->n_entries));
        if (e != hipSuccess) rc = hip_fail(c, e, "chain allocation");
    }
    if (!rc)
        rc = check_launch(c, build_inverse_index(ch->n_end_entries, ch->n_entries, ch->M, d->ends_idx,
                                                 d->sample_idx, ws, ch->entries, ch->rng, c->stream),
                          "chain inverse index");
    if (ws) {
        hipError_t e = hipStreamSynchronize(c->stream);
        (void)hipFree(ws);
        if (!rc && e != hipSuccess) rc = hip_fail(c, e, "chain inverse index");
    }
    if (rc) {
        chain_free(ch);
        return rc;
    }
    *out = ch;
    return 0;
}

void arapgs_chain_destroy(arapgs_chain* ch) { chain_free(ch); }

int arapgs_chain_forward(arapgs_ctx* c, arapgs_chain* ch, const arapgs_node_xf* table, const float* rot0,
                         const float* scale0, float* const* images_h, float* feat, float* dens) {
    CHECK_CTX(c);
    if (!ch || !table || !rot0 || !scale0 || !images_h) return fail(c, -1, "chain_forward: bad arguments");
    if (ch->device != c->device) return fail(c, -1, "chain_forward: chain of another device");
    if ((uintptr_t)rot0 % 16) return fail(c, -1, "chain_forward: rot0 must be 16-byte aligned");
    for (int v = 0; v < ch->V; v++)
        if (!images_h[v]) return fail(c, -1, "chain_forward: image %d is NULL", v);
    if (ch->S > 0 && (!feat || !dens)) return fail(c, -1, "chain_forward: the grid needs feat and dens outputs");
    (void)hipSetDevice(c->device);
    int rc = chain_deform(c, ch, table, rot0, scale0);
    if (rc) return rc;
    rc = chain_render(c, ch, images_h);
    if (rc) return rc;
    if (ch->S > 0)
        rc = check_launch(c, launch_density(ch->nv, ch->grid->valid, ch->grid->prefix, ch->grid->list, ch->samples, ch->pos,
                                            ch->rot, ch->scale, ch->d.opacity, ch->sh, grid_filter(ch->grid, ch->d.lpf), feat, dens,
                                            c->stream),
                          "chain density");
    return rc;
}

int arapgs_chain_backward(arapgs_ctx* c, arapgs_chain* ch, const arapgs_node_xf* table, const float* rot0,
                          const float* scale0, const float* const* dimages_h, const float* dfeat,
                          const float* ddens, arapgs_node_xf* dtable, float* drot0, float* dscale0) {
    CHECK_CTX(c);
    if (!ch || !table || !rot0 || !scale0 || !dtable || !drot0 || !dscale0)
        return fail(c, -1, "chain_backward: bad arguments");
    if (ch->device != c->device) return fail(c, -1, "chain_backward: chain of another device");
    if ((uintptr_t)rot0 % 16) return fail(c, -1, "chain_backward: rot0 must be 16-byte aligned");
    (void)hipSetDevice(c->device);
    hipStream_t st = c->stream;
    int rc = chain_deform(c, ch, table, rot0, scale0);
    if (rc) return rc;
    // renderer: recompute the binning, then the blend and preprocess adjoints
    std::vector<float*> outs((size_t)ch->V);
    const size_t img = 3 * (size_t)ch->W * ch->H;
    for (int v = 0; v < ch->V; v++) outs[v] = ch->img + v * img;
    rc = chain_render(c, ch, outs.data());
    if (rc) return rc;
    rc = check_launch(c, launch_zero_f32(ch->grec, (int64_t)ch->V * ch->N * 8, st), "chain backward");
    if (rc) return rc;
    if (dimages_h) {
        RasterView* hv = c->r_views_h;
        for (int v = 0; v < ch->V; v++) hv[v].dout = dimages_h[v];
        hipError_t e = hipMemcpyAsync(c->r_views, hv, sizeof(RasterView) * ch->V, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);  // the pinned staging is reused by the next render
        if (e != hipSuccess) return hip_fail(c, e, "chain backward views");
        RasterParams rp = c->last_rp;
        rp.views = c->r_views;
        rc = check_launch(c, raster_blend_backward(rp, c->r_geo, (const uint32_t*)((char*)c->r_bin +
                                                                                     std::max<size_t>(c->last_rendered, 1) *
                                                                                         BIN_VALS_AT),
                                                   (const uint2*)c->r_ranges, ch->grec, st),
                          "chain blend backward");
        if (rc) return rc;
    }
    {
        RasterParams rp = c->last_rp;
        rp.views = c->r_views;
        rc = check_launch(c, raster_preprocess_backward(rp, c->r_geo, ch->grec, ch->gout, st),
                          "chain preprocess backward");
        if (rc) return rc;
    }
    // density
    const float* gd = nullptr;
    if (ch->S > 0) {
        rc = check_launch(c, launch_zero_f32(ch->gdens, 58 * ch->N, st), "chain backward");
        if (!rc && (dfeat || ddens))
            rc = check_launch(c, launch_density_backward(ch->nv, ch->grid->valid, ch->grid->prefix, ch->grid->list, ch->samples,
                                                         ch->pos, ch->rot, ch->scale, ch->d.opacity, ch->sh,
                                                         grid_filter(ch->grid, ch->d.lpf), dfeat, ddens, ch->gsamp,
                                                         ch->gdens, ch->N,
                                                         st),
                              "chain density backward");
        else if (!rc)
            rc = check_launch(c, launch_zero_f32(ch->gsamp, 3 * ch->S, st), "chain backward");
        if (rc) return rc;
        gd = ch->gdens;
    }
    rc = check_launch(c, launch_six_point_backward(ch->N, ch->K, (const NodeXf*)table, ch->d.ends_idx, ch->d.ends_w,
                                                   ch->ends, rot0, scale0, ch->d.sh0, ch->rot, ch->scale, ch->gout,
                                                   gd, grid_filter(ch->grid, ch->d.lpf), ch->d.axis_padding,
                                                   ch->d.end_coeff, c->shrot_dev,
                                                   ch->gends, drot0, dscale0, st),
                      "chain six-point backward");
    if (rc) return rc;
    return check_launch(c, launch_node_gather(ch->M, ch->K, ch->n_end_entries, ch->rng, ch->entries,
                                              (const NodeXf*)table, ch->d.ends_w, ch->ends0, ch->gends,
                                              ch->d.sample_w, ch->samples0, ch->gsamp, (NodeXf*)dtable, st),
                        "chain node gather");
}

int arapgs_chain_state(const arapgs_chain* ch, const float** pos, const float** rot, const float** scale,
                       const float** sh, const float** samples) {
    if (!ch) return -1;
    if (pos) *pos = ch->pos;
    if (rot) *rot = ch->rot;
    if (scale) *scale = ch->scale;
    if (sh) *sh = ch->sh;
    if (samples) *samples = ch->samples;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- realignment
struct arapgs_realign {
    int device = 0;
    const arapgs_grid* grid = nullptr;
    int64_t N = 0, S = 0;
    arapgs_realign_desc d{};
    int64_t t = 0;  // Adam step
    float *scale = nullptr, *opac = nullptr, *feat = nullptr, *dens = nullptr, *dfeat = nullptr, *ddens = nullptr;
    float *gd = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    // options (GaussianView.cpp:1695-1733, 4191-4207, 4452-4478)
    const uint8_t* moved = nullptr;    // only these Gaussians are updated (null: all)
    const float* max_scale = nullptr;  // upper bound of the activated scales (null: none)
    float conv_tol = 0.0f;             // per-cell convergence (0: off)
    // per valid cell: static (no moved Gaussian), skip (static or converged), loss, reference loss
    uint8_t *stat = nullptr, *skip = nullptr, *conv = nullptr;
    float *cell_loss = nullptr, *cell_ref = nullptr;
    uint8_t* cell_map = nullptr;  // G^3 bytes for ExpandOpRange (on first use)
    int64_t seen_version = -1;  // grid version the static flags were computed for
    bool restart = true;        // the next step records the reference losses and clears convergence
    std::vector<void*> allocs;
};

namespace {
void realign_free(arapgs_realign* r) {
    if (!r) return;
    (void)hipSetDevice(r->device);
    for (void* p : r->allocs) (void)hipFree(p);
    delete r;
}
}  // namespace

extern "C" {

int arapgs_realign_create(arapgs_ctx* c, const arapgs_grid* grid, int64_t n, const arapgs_realign_desc* d,
                          arapgs_realign** out) {
    CHECK_CTX(c);
    if (!grid || !d || !out || n < 1 || n < grid->n_gauss) return fail(c, -1, "realign_create: bad arguments");
    *out = nullptr;
    (void)hipSetDevice(c->device);
    arapgs_realign* r = new arapgs_realign();
    r->device = c->device;
    r->grid = grid;
    r->N = n;
    r->S = 64 * grid->n_valid;
    r->d = *d;
    if (r->d.beta1 == 0.0f) r->d.beta1 = 0.9f;
    if (r->d.beta2 == 0.0f) r->d.beta2 = 0.999f;
    if (r->d.eps == 0.0f) r->d.eps = 1e-15f;
    const size_t N = (size_t)n, S = (size_t)std::max<int64_t>(r->S, 1);
    const size_t NV = (size_t)std::max<int64_t>(grid->n_valid, 1);
    struct { float** p; size_t count; } bufs[] = {
        {&r->scale, 3 * N}, {&r->opac, N}, {&r->feat, 48 * S}, {&r->dens, S}, {&r->dfeat, 48 * S}, {&r->ddens, S},
        {&r->gd, 58 * N}, {&r->grad, 59 * N}, {&r->m, 59 * N}, {&r->v, 59 * N},
        {&r->cell_loss, NV}, {&r->cell_ref, NV}, {(float**)&r->stat, (NV + 3) / 4}, {(float**)&r->skip, (NV + 3) / 4},
        {(float**)&r->conv, (NV + 3) / 4}};
    for (auto& b : bufs) {
        hipError_t e = hipMalloc((void**)b.p, sizeof(float) * b.count);
        if (e != hipSuccess) {
            realign_free(r);
            return hip_fail(c, e, "realign allocation");
        }
        r->allocs.push_back((void*)*b.p);
    }
    int rc = check_launch(c, launch_zero_f32(r->m, 59 * (int64_t)N, c->stream), "realign_create");
    if (!rc) rc = check_launch(c, launch_zero_f32(r->v, 59 * (int64_t)N, c->stream), "realign_create");
    if (rc) {
        realign_free(r);
        return rc;
    }
    *out = r;
    return 0;
}

void arapgs_realign_destroy(arapgs_realign* r) { realign_free(r); }

int arapgs_realign_step(arapgs_ctx* c, arapgs_realign* r, const float* samples, const float* aim_feat,
                        const float* aim_dens, float lpf, float* pos, float* rot, float* log_scale,
                        float* logit_opacity, float* sh, double* loss) {
    CHECK_CTX(c);
    if (!r || !samples || !aim_feat || !aim_dens || !pos || !rot || !log_scale || !logit_opacity || !sh)
        return fail(c, -1, "realign_step: bad arguments");
    if (r->device != c->device) return fail(c, -1, "realign_step: object of another device");
    if ((uintptr_t)sh % 16) return fail(c, -1, "realign_step: sh must be 16-byte aligned");
    (void)hipSetDevice(c->device);
    hipStream_t st = c->stream;
    const arapgs_grid* g = r->grid;
    const int64_t N = r->N;
    int rc = check_launch(c, launch_activate(N, log_scale, logit_opacity, r->scale, r->opac, st), "realign activate");
    // a new binning (or new options) resets the static cells and the convergence record
    const bool cells_new = r->seen_version != g->version;
    if (!rc && cells_new) {
        if (r->moved)
            rc = check_launch(c, launch_static_cells(g->n_valid, g->valid, g->prefix, g->list, r->moved, r->stat, st),
                              "realign static cells");
        r->seen_version = g->version;
    }
    if (!rc)
        rc = check_launch(c, launch_density(g->n_valid, g->valid, g->prefix, g->list, samples, pos, rot, r->scale,
                                            r->opac, sh, grid_filter(g, lpf), r->feat, r->dens, st),
                          "realign density");
    if (!rc && loss) rc = check_launch(c, launch_zero_f32((float*)loss, 2, st), "realign loss");  // 0.0 as f64
    if (!rc)
        rc = check_launch(c, launch_l1_grad(r->S, r->feat, r->dens, aim_feat, aim_dens, r->d.w_density, r->dfeat,
                                            r->ddens, loss, r->cell_loss, st),
                          "realign loss");
    if (!rc)
        rc = check_launch(c, launch_converge(g->n_valid, r->cell_loss, r->cell_ref, r->moved ? r->stat : nullptr,
                                             r->conv_tol, r->restart ? 1 : 0, r->conv, r->skip, st),
                          "realign convergence");
    r->restart = false;
    if (!rc) rc = check_launch(c, launch_zero_f32(r->gd, 58 * N, st), "realign backward");
    if (!rc)
        rc = check_launch(c, launch_density_backward(g->n_valid, g->valid, g->prefix, g->list, samples, pos, rot,
                                                     r->scale, r->opac, sh, grid_filter(g, lpf), r->dfeat, r->ddens, nullptr, r->gd,
                                                     N, st, r->skip),
                          "realign backward");
    float* dpos = r->grad;
    float* drot = dpos + 3 * N;
    float* dscale = drot + 4 * N;
    float* dopac = dscale + 3 * N;
    float* dsh = dopac + N;
    if (!rc)
        rc = check_launch(c, launch_density_attr_grad(N, rot, r->scale, r->gd, grid_filter(g, lpf), dpos, drot, dscale,
                                                      dopac, dsh, st),
                          "realign backward");
    if (rc) return rc;
    r->t++;
    const float lr[5] = {r->d.lr_pos, r->d.lr_rot, r->d.lr_scale, r->d.lr_opacity, r->d.lr_sh};
    const float bc1 = 1.0f - (float)std::pow((double)r->d.beta1, (double)r->t);
    const float bc2 = 1.0f - (float)std::pow((double)r->d.beta2, (double)r->t);
    return check_launch(c, launch_adam(N, dpos, drot, dscale, dopac, dsh, r->scale, r->opac, pos, rot, log_scale,
                                       logit_opacity, sh, r->m, r->v, lr, r->d.beta1, r->d.beta2, r->d.eps, bc1, bc2,
                                       r->moved, r->max_scale, st),
                        "realign adam");
}

int arapgs_realign_set_options(arapgs_ctx* c, arapgs_realign* r, const arapgs_realign_options* o) {
    CHECK_CTX(c);
    if (!r || !o || o->conv_tol < 0.0f) return fail(c, -1, "realign_set_options: bad arguments");
    r->moved = o->moved;
    r->max_scale = o->max_scale;
    r->conv_tol = o->conv_tol;
    r->seen_version = -1;  // static cells and convergence start over
    r->restart = true;
    return 0;
}

int arapgs_realign_expand(arapgs_ctx* c, arapgs_realign* r) {
    CHECK_CTX(c);
    if (!r) return fail(c, -1, "realign_expand: bad arguments");
    if (r->restart) return 0;  // nothing converged yet
    (void)hipSetDevice(c->device);
    const arapgs_grid* g = r->grid;
    if (!r->cell_map) {
        const int64_t gc = (int64_t)g->G * g->G * g->G;
        hipError_t e = hipMalloc((void**)&r->cell_map, (size_t)((gc + 3) / 4) * 4);
        if (e != hipSuccess) return hip_fail(c, e, "realign_expand");
        r->allocs.push_back(r->cell_map);
    }
    return check_launch(c, launch_expand_op_range(g->n_valid, g->valid, g->G, r->moved ? r->stat : nullptr, r->conv,
                                                  r->skip, r->cell_map, c->stream),
                        "realign_expand");
}

int arapgs_realign_cells(const arapgs_realign* r, int64_t* n_valid, const uint8_t** skip, const float** cell_loss) {
    if (!r) return -1;
    if (n_valid) *n_valid = r->grid->n_valid;
    if (skip) *skip = r->skip;
    if (cell_loss) *cell_loss = r->cell_loss;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- PLY codec
extern "C" {

int arapgs_ply_info(const char* path, int64_t* count_h, int* has_index_h) {
    if (!path) return -1;
    PlyHeader h;
    std::string err;
    if (ply_parse_header(path, &h, &err)) return -1;
    if (count_h) *count_h = h.count;
    if (has_index_h) *has_index_h = h.has_index;
    return 0;
}

int arapgs_ply_load(arapgs_ctx* c, const char* path, int morton, float* pos, float* rot, float* scale, float* opacity,
                    float* sh, int* index, float* minmax6_h) {
    CHECK_CTX(c);
    if (!path) return fail(c, -1, "ply_load: bad arguments");
    PlyHeader h;
    std::string err;
    if (ply_parse_header(path, &h, &err)) return fail(c, -1, "ply_load: %s", err.c_str());
    const int64_t n = h.count;
    if (n == 0) return 0;
    if (!pos || !rot || !scale || !opacity || !sh) return fail(c, -1, "ply_load: bad arguments");
    if (n > 0xffffffffll) return fail(c, -1, "ply_load: too many vertices");
    std::vector<float> recs((size_t)n * h.stride);
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(c, -1, "ply_load: cannot open %s", path);
    std::fseek(f, h.data_offset, SEEK_SET);
    const size_t got = std::fread(recs.data(), sizeof(float) * h.stride, (size_t)n, f);
    std::fclose(f);
    if (got != (size_t)n) return fail(c, -1, "ply_load: %s is truncated (%zu of %lld vertices)", path, got, (long long)n);
    (void)hipSetDevice(c->device);
    float* drec = nullptr;
    uint32_t* ws = nullptr;
    float* mm = nullptr;
    hipError_t e = hipMalloc(&drec, sizeof(float) * recs.size());
    if (e == hipSuccess) e = hipMalloc(&ws, sizeof(uint32_t) * ply_workspace_words(n));
    if (e == hipSuccess) e = hipMalloc(&mm, sizeof(float) * 8);
    if (e == hipSuccess) e = hipMemcpyAsync(drec, recs.data(), sizeof(float) * recs.size(), hipMemcpyHostToDevice, c->stream);
    int rc = 0;
    if (e != hipSuccess) rc = hip_fail(c, e, "ply_load");
    if (!rc)
        rc = check_launch(c, launch_ply_unpack(n, h.stride, drec, morton, ws, mm, pos, rot, scale, opacity, sh, index,
                                               c->stream),
                          "ply_load");
    float mm_h[6] = {0, 0, 0, 0, 0, 0};
    if (!rc && morton) {
        e = hipMemcpyAsync(mm_h, mm, sizeof(mm_h), hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) rc = hip_fail(c, e, "ply_load");
    }
    e = hipStreamSynchronize(c->stream);
    if (!rc && e != hipSuccess) rc = hip_fail(c, e, "ply_load");
    if (!rc && minmax6_h) std::memcpy(minmax6_h, mm_h, sizeof(mm_h));
    (void)hipFree(drec);
    (void)hipFree(ws);
    (void)hipFree(mm);
    return rc;
}

int arapgs_ply_save(arapgs_ctx* c, const char* path, int64_t n, const float* pos, const float* rot, const float* scale,
                    const float* opacity, const float* sh, const float* boxmin_h, const float* boxmax_h,
                    const uint8_t* too_small, int64_t* written_h) {
    CHECK_CTX(c);
    if (!path || n < 0 || (n > 0 && (!pos || !rot || !scale || !opacity || !sh)))
        return fail(c, -1, "ply_save: bad arguments");
    if (n > 0xffffffffll) return fail(c, -1, "ply_save: too many Gaussians");
    (void)hipSetDevice(c->device);
    float* drec = nullptr;
    uint32_t* ws = nullptr;
    uint32_t* dcount = nullptr;
    uint32_t count = 0;
    hipError_t e = hipSuccess;
    int rc = 0;
    std::vector<float> recs;
    if (n > 0) {
        e = hipMalloc(&drec, sizeof(float) * 62 * (size_t)n);
        if (e == hipSuccess) e = hipMalloc(&ws, sizeof(uint32_t) * ply_pack_workspace_words(n));
        if (e == hipSuccess) e = hipMalloc(&dcount, sizeof(uint32_t) * 4);
        if (e != hipSuccess) rc = hip_fail(c, e, "ply_save");
        if (!rc)
            rc = check_launch(c, launch_ply_pack(n, pos, rot, scale, opacity, sh, boxmin_h, boxmax_h, too_small, ws,
                                                 drec, dcount, c->stream),
                              "ply_save");
        if (!rc) {
            e = hipMemcpyAsync(&count, dcount, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = hip_fail(c, e, "ply_save");
        }
        if (!rc) {
            recs.resize((size_t)count * 62);
            e = hipMemcpy(recs.data(), drec, sizeof(float) * recs.size(), hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = hip_fail(c, e, "ply_save");
        }
        (void)hipFree(drec);
        (void)hipFree(ws);
        (void)hipFree(dcount);
        if (rc) return rc;
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(c, -1, "ply_save: cannot create %s", path);
    std::fprintf(f, "ply\nformat binary_little_endian 1.0\nelement vertex %u\n", count);
    const char* head[] = {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"};
    for (const char* p : head) std::fprintf(f, "property float %s\n", p);
    for (int i = 0; i < 45; i++) std::fprintf(f, "property float f_rest_%d\n", i);
    const char* tail[] = {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"};
    for (const char* p : tail) std::fprintf(f, "property float %s\n", p);
    std::fprintf(f, "end_header\n");
    const size_t put = recs.empty() ? 0 : std::fwrite(recs.data(), sizeof(float) * 62, count, f);
    std::fclose(f);
    if (put != count) return fail(c, -1, "ply_save: short write to %s", path);
    if (written_h) *written_h = count;
    return 0;
}

}  // extern "C"
