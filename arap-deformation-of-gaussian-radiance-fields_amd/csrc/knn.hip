// Graph build on gfx950: farthest-point node sampling and the point -> node
// K-nearest search with the reference's exact tie semantics.
//
// The reference KNN (DeformGraph::findNearestNodes, Deform.hpp:153-185) is a
// selection sort over the first k slots of a position array: at step i it keeps
// slot i unless a strictly smaller distance exists, otherwise takes the
// highest-position minimum of the descending scan, then swaps slot i with the
// winner. This file reproduces the resulting index sequence bit for bit:
//   1. one pass over the nodes (staged in LDS) keeps the L smallest squared
//      distances per query (exact float arithmetic, no contraction);
//   2. if every node outside that list has a strictly larger sqrt-distance than
//      the k-th selected one, the selection sort is replayed on the candidates
//      only (positions start at the node index; non-candidates never win);
//   3. otherwise the query is flagged and a wave-cooperative kernel replays the
//      full selection sort over all M positions.
#pragma clang fp contract(off)
#include "common.hpp"

#include <mutex>

namespace arapgs {

namespace {

constexpr int L_TOP = 16;   // candidate list length (>= KNN_MAX + 1 + slack)

// Exact distance of Deform.hpp:164-165: float differences, products summed left
// to right without contraction (this file is compiled with contraction off; the
// header intrinsics __fmul_rn/__fadd_rn are plain contractible ops here, so the
// expression is written out), correctly rounded sqrt.
__device__ __forceinline__ float knn_d2(float px, float py, float pz, float nx, float ny, float nz) {
    const float a = px - nx, b = py - ny, c = pz - nz;
    return (a * a + b * b) + c * c;
}

// Weights of DeformGraph::computeWeights (Deform.hpp:187-208) in double from the
// float distances of the selected slots; written as float.
__device__ __forceinline__ void write_weights(const float* dsel, int k, float* w) {
    const double dmax = (double)dsel[k];
    double wd[KNN_MAX_T];
    double sum = 0.0;
    for (int j = 0; j < k; j++) {
        const double x = 1.0 - (double)dsel[j] / dmax;
        wd[j] = x * x;
        sum += wd[j];
    }
    if (k == 1) { w[0] = 1.0f; return; }
    for (int j = 0; j < k; j++) w[j] = (float)(wd[j] / sum);
}

template <int K>
__global__ void __launch_bounds__(256) knn_main_kernel(const float* __restrict__ nodes, int M,
                                                       const float* __restrict__ queries, int64_t Q,
                                                       uint32_t* __restrict__ out_idx, float* __restrict__ out_w,
                                                       uint32_t* __restrict__ fallback_list,
                                                       uint32_t* __restrict__ fallback_count) {
    constexpr int KK = K + 1;  // slots selected by the reference (k_nearest + 1)
    extern __shared__ float s_nodes[];  // 3*M
    for (int i = threadIdx.x; i < 3 * M; i += blockDim.x) s_nodes[i] = nodes[i];
    __syncthreads();
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Q) return;
    const float px = queries[3 * q], py = queries[3 * q + 1], pz = queries[3 * q + 2];
    float td[L_TOP];
    int ti[L_TOP];
#pragma unroll
    for (int s = 0; s < L_TOP; s++) { td[s] = __builtin_inff(); ti[s] = -1; }
    for (int j = 0; j < M; j++) {
        const float d2 = knn_d2(px, py, pz, s_nodes[3 * j], s_nodes[3 * j + 1], s_nodes[3 * j + 2]);
        if (d2 < td[L_TOP - 1]) {
            float v = d2;
            int id = j;
#pragma unroll
            for (int s = 0; s < L_TOP; s++) {
                if (v < td[s]) {
                    const float tv = td[s]; td[s] = v; v = tv;
                    const int tid = ti[s]; ti[s] = id; id = tid;
                }
            }
        }
    }
    // sqrt distances of the kept candidates: ascending d2, so non-decreasing
    float sd[L_TOP];
#pragma unroll
    for (int s = 0; s < L_TOP; s++) sd[s] = ti[s] >= 0 ? __builtin_sqrtf(td[s]) : __builtin_inff();
    const float T = sd[KK - 1];
    const bool complete = (M <= L_TOP) || (sd[L_TOP - 1] > T);
    if (!complete) {
        const uint32_t slot = atomicAdd(fallback_count, 1u);
        fallback_list[slot] = (uint32_t)q;
        return;
    }
    // Replay the selection sort on the candidates (value <= T); every node not
    // kept has a larger value, so it can neither win nor tie.
    int pos[L_TOP];
    bool alive[L_TOP];
#pragma unroll
    for (int s = 0; s < L_TOP; s++) { pos[s] = ti[s]; alive[s] = (ti[s] >= 0) && (sd[s] <= T); }
    float dsel[KK];
    uint32_t* oq = out_idx + q * KK;
#pragma unroll
    for (int i = 0; i < KK; i++) {
        float mv = __builtin_inff();
#pragma unroll
        for (int s = 0; s < L_TOP; s++)
            if (alive[s] && sd[s] < mv) mv = sd[s];
        int at_i = -1, win = -1, best_pos = -1;
#pragma unroll
        for (int s = 0; s < L_TOP; s++) {
            if (!alive[s]) continue;
            if (pos[s] == i) at_i = s;
            if (sd[s] == mv && pos[s] > best_pos) { best_pos = pos[s]; win = s; }
        }
        // slot i keeps its element unless something is strictly smaller
        if (at_i >= 0 && sd[at_i] == mv) win = at_i;
        int wpos = 0, wid = 0;
        float wv = 0.0f;
#pragma unroll
        for (int s = 0; s < L_TOP; s++)
            if (s == win) { wpos = pos[s]; wid = ti[s]; wv = sd[s]; }
        oq[i] = (uint32_t)wid;
        dsel[i] = wv;
#pragma unroll
        for (int s = 0; s < L_TOP; s++) {
            if (s == win) { pos[s] = i; alive[s] = false; }
            else if (s == at_i) pos[s] = wpos;  // element at slot i moves to the winner's slot
        }
    }
    if (out_w) {
        const double dmax = (double)dsel[K];
        double wd[K];
        double sum = 0.0;
#pragma unroll
        for (int j = 0; j < K; j++) {
            const double x = 1.0 - (double)dsel[j] / dmax;
            wd[j] = x * x;
            sum += wd[j];
        }
        float* wq = out_w + q * K;
        if (K == 1) wq[0] = 1.0f;
        else {
#pragma unroll
            for (int j = 0; j < K; j++) wq[j] = (float)(wd[j] / sum);
        }
    }
}

// Exact fallback: one wave per flagged query replays the full selection sort.
// Swapped positions are tracked in a small override table (<= 2(k+1) entries).
__global__ void __launch_bounds__(64) knn_fallback_kernel(const float* __restrict__ nodes, int M,
                                                          const float* __restrict__ queries, int k,
                                                          uint32_t* __restrict__ out_idx, float* __restrict__ out_w,
                                                          const uint32_t* __restrict__ fallback_list,
                                                          const uint32_t* __restrict__ fallback_count) {
    const uint32_t n_fb = *fallback_count;
    const int kk = k + 1;
    __shared__ int ov_pos[2 * (KNN_MAX_T + 1)], ov_el[2 * (KNN_MAX_T + 1)];
    __shared__ int n_ov;
    __shared__ float red_v[64];
    __shared__ int red_p[64];
    for (uint32_t f = blockIdx.x; f < n_fb; f += gridDim.x) {
        const int64_t q = fallback_list[f];
        const float px = queries[3 * q], py = queries[3 * q + 1], pz = queries[3 * q + 2];
        if (threadIdx.x == 0) n_ov = 0;
        __syncthreads();
        float dsel[KNN_MAX_T + 1];
        for (int i = 0; i < kk; i++) {
            // element currently at position p
            auto elem_at = [&](int p) {
                int e = p;
                for (int o = 0; o < n_ov; o++)
                    if (ov_pos[o] == p) e = ov_el[o];
                return e;
            };
            float bv = __builtin_inff();
            int bp = -1;  // priority: position i beats all, then higher positions
            for (int p = i + (int)threadIdx.x; p < M; p += 64) {
                const int e = elem_at(p);
                const float v = __builtin_sqrtf(knn_d2(px, py, pz, nodes[3 * e], nodes[3 * e + 1], nodes[3 * e + 2]));
                const int pr = (p == i) ? 0x7fffffff : p;
                if (v < bv || (v == bv && pr > bp)) { bv = v; bp = pr; }
            }
            red_v[threadIdx.x] = bv;
            red_p[threadIdx.x] = bp;
            __syncthreads();
            if (threadIdx.x == 0) {
                float v = red_v[0];
                int pr = red_p[0];
                for (int l = 1; l < 64; l++)
                    if (red_v[l] < v || (red_v[l] == v && red_p[l] > pr)) { v = red_v[l]; pr = red_p[l]; }
                const int wpos = (pr == 0x7fffffff) ? i : pr;
                const int we = elem_at(wpos), ie = elem_at(i);
                out_idx[q * kk + i] = (uint32_t)we;
                red_v[0] = v;
                // swap elements at positions i and wpos
                int have_i = -1, have_w = -1;
                for (int o = 0; o < n_ov; o++) {
                    if (ov_pos[o] == i) have_i = o;
                    if (ov_pos[o] == wpos) have_w = o;
                }
                if (have_i < 0) { have_i = n_ov; ov_pos[n_ov++] = i; }
                ov_el[have_i] = we;
                if (wpos != i) {
                    if (have_w < 0) { have_w = n_ov; ov_pos[n_ov++] = wpos; }
                    ov_el[have_w] = ie;
                }
            }
            __syncthreads();
            dsel[i] = red_v[0];
            __syncthreads();
        }
        if (out_w && threadIdx.x == 0) write_weights(dsel, k, out_w + q * k);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- FPS
// farthest_control_points_sampling (helper.cpp:139-177): distances as
// pts_distance (helper.cpp:60-63) — float differences, squares summed in double,
// double sqrt rounded to float; running min with strict '<'; argmax keeps the
// first maximum (std::max_element).

// FetchFirstNodeIdx (helper.cpp:180-195): argmax of x+y+z (float, first max,
// initial -100000).
__global__ void fps_first_kernel(const float* __restrict__ pos, int64_t N, unsigned long long* __restrict__ best) {
    // packed key: larger sum first, then smaller index
    __shared__ unsigned long long red[256];
    unsigned long long local = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const float s = (0.0f + pos[3 * i]) + pos[3 * i + 1] + pos[3 * i + 2];
        if (!(s > -100000.0f)) continue;
        uint32_t b = __float_as_uint(s);
        b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);  // order-preserving map
        const unsigned long long key = ((unsigned long long)b << 32) | (uint32_t)(0xffffffffu - (uint32_t)i);
        local = local > key ? local : key;
    }
    red[threadIdx.x] = local;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + s] ? red[threadIdx.x] : red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(best, red[0]);
}

__global__ void fps_init_kernel(float* __restrict__ dist, int64_t N, const unsigned long long* __restrict__ best,
                                int* __restrict__ out, unsigned int* __restrict__ counter) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        dist[i] = 3.402823466e+38f;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned long long b = *best;
        out[0] = b == 0 ? 0 : (int)(0xffffffffu - (uint32_t)(b & 0xffffffffu));
        *counter = 0;
    }
}

// One FPS step: fold the last chosen point into the running minimum and find
// the first maximum; the last block to finish writes out[step+1].
__global__ void __launch_bounds__(256) fps_step_kernel(const float* __restrict__ pos, int64_t N, float* __restrict__ dist,
                                                       int step, int* __restrict__ out,
                                                       unsigned long long* __restrict__ partial,
                                                       unsigned int* __restrict__ counter) {
    const int last = out[step];
    const float lx = pos[3 * (int64_t)last], ly = pos[3 * (int64_t)last + 1], lz = pos[3 * (int64_t)last + 2];
    unsigned long long local = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const float dx = lx - pos[3 * i], dy = ly - pos[3 * i + 1], dz = lz - pos[3 * i + 2];
        const double s = ((double)dx * (double)dx + (double)dy * (double)dy) + (double)dz * (double)dz;
        const float d = (float)__builtin_sqrt(s);
        float cur = dist[i];
        if (d < cur) { cur = d; dist[i] = d; }
        // distances are >= 0: float bits order like unsigned ints
        const unsigned long long key =
            ((unsigned long long)__float_as_uint(cur) << 32) | (uint32_t)(0xffffffffu - (uint32_t)i);
        local = local > key ? local : key;
    }
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = local;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + s] ? red[threadIdx.x] : red[threadIdx.x + s];
        __syncthreads();
    }
    __shared__ bool is_last;
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = red[0];
        __threadfence();
        const unsigned int done = atomicAdd(counter, 1u);
        is_last = (done == gridDim.x - 1);
    }
    __syncthreads();
    if (is_last) {
        __threadfence();
        unsigned long long v = 0;
        for (unsigned int b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
            const unsigned long long p = __hip_atomic_load(&partial[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v = v > p ? v : p;
        }
        red[threadIdx.x] = v;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + s] ? red[threadIdx.x] : red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            out[step + 1] = (int)(0xffffffffu - (uint32_t)(red[0] & 0xffffffffu));
            *counter = 0;
        }
    }
}

__global__ void clear_count_kernel(uint32_t* __restrict__ count) {
    if (threadIdx.x == 0) *count = 0u;
}

}  // namespace

// ---------------------------------------------------------------- launchers
int launch_knn(const float* nodes, int M, const float* queries, int64_t Q, int k, uint32_t* out_idx, float* out_w,
               uint32_t* scratch_list, uint32_t* scratch_count, hipStream_t st) {
    if (Q == 0) return 0;
    const size_t lds = sizeof(float) * 3 * (size_t)M;
    // a kernel, not a memset, clears the count: callers may capture the launches
    hipLaunchKernelGGL(clear_count_kernel, dim3(1), dim3(64), 0, st, scratch_count);
    const int64_t blocks = (Q + 255) / 256;
    switch (k) {
#define ARAPGS_KNN_CASE(KV)                                                                                     \
    case KV:                                                                                                    \
        hipLaunchKernelGGL(knn_main_kernel<KV>, dim3((unsigned)blocks), dim3(256), lds, st, nodes, M, queries, Q, \
                           out_idx, out_w, scratch_list, scratch_count);                                        \
        break;
        ARAPGS_KNN_CASE(1) ARAPGS_KNN_CASE(2) ARAPGS_KNN_CASE(3) ARAPGS_KNN_CASE(4) ARAPGS_KNN_CASE(5)
        ARAPGS_KNN_CASE(6) ARAPGS_KNN_CASE(7) ARAPGS_KNN_CASE(8) ARAPGS_KNN_CASE(9) ARAPGS_KNN_CASE(10)
        ARAPGS_KNN_CASE(11) ARAPGS_KNN_CASE(12)
#undef ARAPGS_KNN_CASE
        default:
            return -1;
    }
    HIP_RET(hipGetLastError());
    hipLaunchKernelGGL(knn_fallback_kernel, dim3(1024), dim3(64), 0, st, nodes, M, queries, k, out_idx, out_w,
                       scratch_list, scratch_count);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_fps(const float* pos, int64_t N, int m_steps, int* out, float* dist, unsigned long long* partial,
               int n_partial, unsigned int* counter, unsigned long long* best, hipStream_t st) {
    HIP_RET(hipMemsetAsync(best, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(fps_first_kernel, dim3(512), dim3(256), 0, st, pos, N, best);
    hipLaunchKernelGGL(fps_init_kernel, dim3(512), dim3(256), 0, st, dist, N, best, out, counter);
    for (int s = 0; s < m_steps; s++)
        hipLaunchKernelGGL(fps_step_kernel, dim3(n_partial), dim3(256), 0, st, pos, N, dist, s, out, partial, counter);
    HIP_RET(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- no-allocation KNN
// The scratch of arapgs_knn_noalloc: word 0 is the fallback count, words
// 64 .. 64 + KNN_NOALLOC_MAX_QUERIES - 1 the fallback list. Every GPU that
// loads the library holds its own copy until the process ends. In host code
// `scratch` names a host-side stand-in, not device memory: host code uses only
// the per-device address that hipGetSymbolAddress returned.
static __device__ __align__(256) uint32_t scratch[64 + KNN_NOALLOC_MAX_QUERIES];

namespace {
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[KNN_NOALLOC_MAX_DEVICES] = {};
}  // namespace

int knn_noalloc_scratch(int device, uint32_t** out) {
    *out = nullptr;
    if (device < 0 || device >= KNN_NOALLOC_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        HIP_RET(hipGetDevice(&cur));
        if (cur != device) return -1;  // the caller makes the device current
        void* p = nullptr;
        HIP_RET(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return -1;
        g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[device];
    return 0;
}

}  // namespace arapgs
