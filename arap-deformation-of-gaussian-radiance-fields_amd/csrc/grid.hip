// G^3 spatial grid over the Gaussians (getGridSamples, GaussianView.cpp:3870-4149;
// GetGsGrid / GetBoxesGsGrid + scans, cudakdtree.cu:403-451,508-553) and the
// density sample on its cells (the un-vendored forward3d_grid, called at
// GaussianView.cpp:4159-4186), for gfx950.
//
//   * scene AABB: device min/max reduction, then the float expansion rule on the host
//   * reorder: stable radix sort of the centre cells (the rank rule of
//     GaussianView.cpp:3904-3953 is exactly a stable sort by cell)
//   * lists: every Gaussian emits its (cell, id) pairs at a scanned offset, a
//     stable radix sort by cell yields per-cell lists in ascending id order —
//     the order of the reference's serial fill loop (GaussianView.cpp:4076-4100)
//   * density: one workgroup per valid cell, its Gaussians staged through LDS,
//     64 samples per cell (one lane each)
#pragma clang fp contract(off)
#include "common.hpp"
#include "grid.hpp"
#include "sort.hpp"

namespace arapgs {

namespace {

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ uint32_t ordered_bits(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float from_ordered(uint32_t b) {
    return __uint_as_float((b & 0x80000000u) ? (b & 0x7fffffffu) : ~b);
}

__global__ void minmax_kernel(const float* __restrict__ pos, int64_t n, uint32_t* __restrict__ mm) {
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const uint32_t b = ordered_bits(pos[3 * i + c]);
            lo[c] = min(lo[c], b);
            hi[c] = max(hi[c], b);
        }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        for (int o = 32; o > 0; o >>= 1) {
            lo[c] = min(lo[c], (uint32_t)__shfl_xor((int)lo[c], o));
            hi[c] = max(hi[c], (uint32_t)__shfl_xor((int)hi[c], o));
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMin(&mm[c], lo[c]);
            atomicMax(&mm[3 + c], hi[c]);
        }
    }
}

__global__ void minmax_init_kernel(uint32_t* __restrict__ mm) {
    if (threadIdx.x < 3) mm[threadIdx.x] = 0xffffffffu;
    else if (threadIdx.x < 6) mm[threadIdx.x] = 0u;
}

__global__ void decode_minmax_kernel(const uint32_t* __restrict__ mm, float* __restrict__ out) {
    if (threadIdx.x < 6) out[threadIdx.x] = from_ordered(mm[threadIdx.x]);
}

__device__ __forceinline__ int cell_coord(float p, float mn, float step) { return (int)floorf((p - mn) / step); }

__global__ void center_cell_kernel(const float* __restrict__ pos, int64_t n, GridGeom gg, uint64_t* __restrict__ keys,
                                   uint32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = cell_coord(pos[3 * i], gg.mn[0], gg.step);
    const int y = cell_coord(pos[3 * i + 1], gg.mn[1], gg.step);
    const int z = cell_coord(pos[3 * i + 2], gg.mn[2], gg.step);
    keys[i] = (uint64_t)(uint32_t)(x * gg.G * gg.G + y * gg.G + z);
    vals[i] = (uint32_t)i;
}

// alpha-cutoff AABB (GaussianView.cpp:3960-4019)
__global__ void aabb_kernel(const float* __restrict__ pos, const float* __restrict__ rot,
                            const float* __restrict__ scale, const float* __restrict__ opacity, int64_t n,
                            float* __restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    float q[4] = {rot[4 * g], rot[4 * g + 1], rot[4 * g + 2], rot[4 * g + 3]};
    quat_normalize(q);
    float R[3][3];
    quat_to_mat(q[0], q[1], q[2], q[3], R);
    const float cutoff = 1.0f / 255.0f;
    float s3[3] = {0.0f, 0.0f, 0.0f};
    if (opacity[g] > cutoff) {
        const float lg = (float)log((double)(cutoff / opacity[g]));
        const float f = __builtin_sqrtf(-2.0f * lg);
#pragma unroll
        for (int i = 0; i < 3; i++) s3[i] = f * (scale[3 * g + i] + 0.0f);
    }
    float lo[3], hi[3];
#pragma unroll
    for (int c = 0; c < 3; c++) lo[c] = hi[c] = pos[3 * g + c];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float off = R[c][d] * s3[d];
            const float a = pos[3 * g + c] + off, b = pos[3 * g + c] + (-off);
            lo[c] = fminf(lo[c], fminf(a, b));
            hi[c] = fmaxf(hi[c], fmaxf(a, b));
        }
#pragma unroll
    for (int c = 0; c < 3; c++) { out[6 * g + c] = lo[c]; out[6 * g + 3 + c] = hi[c]; }
}

__device__ __forceinline__ void box_range(const float* a, const GridGeom& gg, int lo[3], int hi[3]) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
        lo[c] = max((int)(floorf((a[c] - gg.mn[c]) / gg.step) - (float)gg.padding), 0);
        hi[c] = min((int)(floorf((a[3 + c] - gg.mn[c]) / gg.step) + (float)gg.padding), gg.G - 1);
    }
}

__global__ void box_count_kernel(const float* __restrict__ aabbs, int64_t n, GridGeom gg, uint32_t* __restrict__ cnt) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int lo[3], hi[3];
    box_range(aabbs + 6 * g, gg, lo, hi);
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    cnt[g] = (nx > 0 && ny > 0 && nz > 0) ? (uint32_t)(nx * ny * nz) : 0u;
}

// 64-bit sum of the per-Gaussian cell counts (the total number of
// cell-Gaussian pairs, which may exceed 2^32 on large grids).
__global__ void sum_counts_kernel(const uint32_t* __restrict__ cnt, int64_t n, unsigned long long* __restrict__ out) {
    unsigned long long s = 0ull;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += cnt[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

// Pairs per x-slab of the grid (cells x*G^2 .. (x+1)*G^2-1): a Gaussian whose
// box spans ny*nz cells in y and z adds ny*nz to every slab its box crosses.
__global__ void slab_count_kernel(const float* __restrict__ aabbs, int64_t n, GridGeom gg,
                                  unsigned long long* __restrict__ slab) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int lo[3], hi[3];
    box_range(aabbs + 6 * g, gg, lo, hi);
    const int ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    if (hi[0] < lo[0] || ny <= 0 || nz <= 0) return;
    for (int x = lo[0]; x <= hi[0]; x++) atomicAdd(&slab[x], (unsigned long long)(ny * nz));
}

// Cells of each Gaussian's box inside the slabs [x0, x1).
__global__ void box_count_slabs_kernel(const float* __restrict__ aabbs, int64_t n, GridGeom gg, int x0, int x1,
                                       uint32_t* __restrict__ cnt) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int lo[3], hi[3];
    box_range(aabbs + 6 * g, gg, lo, hi);
    const int nx = min(hi[0], x1 - 1) - max(lo[0], x0) + 1;
    const int ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    cnt[g] = (nx > 0 && ny > 0 && nz > 0) ? (uint32_t)(nx * ny * nz) : 0u;
}

// The pairs of the slabs [x0, x1): key = cell - x0*G^2 (32-bit), value = the
// Gaussian, written from offs[g] in ascending cell order; every pair counts
// into its cell.
__global__ void box_emit_slabs_kernel(const float* __restrict__ aabbs, int64_t n, GridGeom gg, int x0, int x1,
                                      const uint32_t* __restrict__ offs, uint32_t* __restrict__ keys,
                                      uint32_t* __restrict__ vals, uint32_t* __restrict__ cell_count) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int lo[3], hi[3];
    box_range(aabbs + 6 * g, gg, lo, hi);
    const int xa = max(lo[0], x0), xb = min(hi[0], x1 - 1);
    const int G2 = gg.G * gg.G;
    const int c0 = x0 * G2;
    uint32_t o = offs[g];
    for (int x = xa; x <= xb; x++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int z = lo[2]; z <= hi[2]; z++) {
                const int c = x * G2 + y * gg.G + z;
                keys[o] = (uint32_t)(c - c0);
                vals[o] = (uint32_t)g;
                o++;
                atomicAdd(&cell_count[c], 1u);
            }
}

// 64-bit inclusive scan of the per-cell counts (tiles of 2048 cells): block
// sums, one-block scan of the sums, then each tile scanned with its offset.
constexpr int SCAN64_THREADS = 256;
constexpr int SCAN64_ITEMS = 8;
constexpr int SCAN64_TILE = SCAN64_THREADS * SCAN64_ITEMS;

__device__ __forceinline__ unsigned long long block_exclusive_scan64(unsigned long long v, unsigned long long* sh,
                                                                     unsigned long long* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            const unsigned long long c = sh[w];
            sh[w] = run;
            run += c;
        }
        sh[16] = run;
    }
    __syncthreads();
    const unsigned long long r = sh[wave] + x - v;
    if (total) *total = sh[16];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(SCAN64_THREADS) scan64_reduce_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                                       unsigned long long* __restrict__ sums) {
    __shared__ unsigned long long sh[17];
    const int64_t base = (int64_t)blockIdx.x * SCAN64_TILE + (int64_t)threadIdx.x * SCAN64_ITEMS;
    unsigned long long s = 0;
#pragma unroll
    for (int i = 0; i < SCAN64_ITEMS; i++)
        if (base + i < n) s += in[base + i];
    unsigned long long tot;
    block_exclusive_scan64(s, sh, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN64_THREADS) scan64_blocks_kernel(unsigned long long* __restrict__ sums,
                                                                       int64_t nb) {
    __shared__ unsigned long long sh[17];
    unsigned long long carry = 0;
    for (int64_t base = 0; base < nb; base += SCAN64_THREADS) {
        const int64_t i = base + threadIdx.x;
        const unsigned long long v = i < nb ? sums[i] : 0ull;
        unsigned long long tot;
        const unsigned long long ex = block_exclusive_scan64(v, sh, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
}

// prefix[c] = pairs in cells 0..c (inclusive), flags[c] = cell c holds any.
__global__ void __launch_bounds__(SCAN64_THREADS) scan64_apply_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                                      const unsigned long long* __restrict__ block_off,
                                                                      int64_t* __restrict__ prefix,
                                                                      uint32_t* __restrict__ flags) {
    __shared__ unsigned long long sh[17];
    const int64_t base = (int64_t)blockIdx.x * SCAN64_TILE + (int64_t)threadIdx.x * SCAN64_ITEMS;
    uint32_t v[SCAN64_ITEMS];
    unsigned long long s = 0;
#pragma unroll
    for (int i = 0; i < SCAN64_ITEMS; i++) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        s += v[i];
    }
    unsigned long long run = block_exclusive_scan64(s, sh, nullptr) + block_off[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN64_ITEMS; i++) {
        run += v[i];
        if (base + i < n) {
            prefix[base + i] = (int64_t)run;
            flags[base + i] = v[i] ? 1u : 0u;
        }
    }
}

__global__ void compact_valid_kernel(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos_ex,
                                     int64_t n, int* __restrict__ valid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flags[i]) valid[pos_ex[i]] = (int)i;
}

__global__ void samples_kernel(const int* __restrict__ valid, int64_t nv, GridGeom gg, float* __restrict__ samples) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv * 64) return;
    const int64_t v = t / 64;
    const int s = (int)(t % 64);
    const int c = valid[v];
    const int G = gg.G;
    const int xi = c / (G * G), yi = (c - xi * G * G) / G, zi = c % G;
    const float interval = gg.step / 4;
    const float x0 = (float)((double)(gg.mn[0] + xi * gg.step) + 0.5 * (double)interval);
    const float y0 = (float)((double)(gg.mn[1] + yi * gg.step) + 0.5 * (double)interval);
    const float z0 = (float)((double)(gg.mn[2] + zi * gg.step) + 0.5 * (double)interval);
    const int sx = s / 16, sy = (s - sx * 16) / 4, sz = s % 4;
    samples[3 * t] = x0 + sx * interval;
    samples[3 * t + 1] = y0 + sy * interval;
    samples[3 * t + 2] = z0 + sz * interval;
}

// Density sample: a_g(x) = o_g exp(-1/2 d^T (Sigma_g + L)^-1 d), L = lpf I or
// the cell's adaptive filter, feature = sum a_g sh_g, opacity = sum a_g. One
// workgroup (64 lanes) per cell.
__global__ void __launch_bounds__(64) density_kernel(int64_t nv, const int* __restrict__ valid,
                                                     const int64_t* __restrict__ prefix, const int* __restrict__ list,
                                                     const float* __restrict__ samples, const float* __restrict__ pos,
                                                     const float* __restrict__ rot, const float* __restrict__ scale,
                                                     const float* __restrict__ opacity, const float* __restrict__ sh,
                                                     DensityFilter flt, float* __restrict__ feat,
                                                     float* __restrict__ opac) {
#pragma clang fp contract(fast)
    constexpr int B = 64;
    __shared__ float s_q[B][10];   // mu(3), inverse covariance (6), opacity
    __shared__ float s_sh[B][48];
    const int64_t v = blockIdx.x;
    if (v >= nv) return;
    const int c = valid[v];
    const int64_t beg = c == 0 ? 0 : prefix[c - 1], end = prefix[c];
    const int64_t si = v * 64 + threadIdx.x;
    const float x0 = samples[3 * si], x1 = samples[3 * si + 1], x2 = samples[3 * si + 2];
    float acc[48];
#pragma unroll
    for (int k = 0; k < 48; k++) acc[k] = 0.0f;
    float aop = 0.0f;
    for (int64_t b0 = beg; b0 < end; b0 += B) {
        const int64_t li = b0 + threadIdx.x;
        if (li < end) {
            const int g = list[li];
            float q[4] = {rot[4 * g], rot[4 * g + 1], rot[4 * g + 2], rot[4 * g + 3]};
            quat_normalize(q);
            float R[3][3];
            quat_to_mat(q[0], q[1], q[2], q[3], R);
            float S[3][3];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    float a = 0.0f;
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        const float sk = scale[3 * g + k];
                        a += R[i][k] * sk * sk * R[j][k];
                    }
                    S[i][j] = a;
                }
            density_filter_add(flt, v, S);
            const float c00 = S[1][1] * S[2][2] - S[1][2] * S[2][1];
            const float c01 = S[1][2] * S[2][0] - S[1][0] * S[2][2];
            const float c02 = S[1][0] * S[2][1] - S[1][1] * S[2][0];
            const float det = S[0][0] * c00 + S[0][1] * c01 + S[0][2] * c02;
            const float id = 1.0f / det;
            s_q[threadIdx.x][0] = pos[3 * g];
            s_q[threadIdx.x][1] = pos[3 * g + 1];
            s_q[threadIdx.x][2] = pos[3 * g + 2];
            s_q[threadIdx.x][3] = c00 * id;
            s_q[threadIdx.x][4] = c01 * id;
            s_q[threadIdx.x][5] = c02 * id;
            s_q[threadIdx.x][6] = (S[0][0] * S[2][2] - S[0][2] * S[2][0]) * id;
            s_q[threadIdx.x][7] = (S[0][2] * S[1][0] - S[0][0] * S[1][2]) * id;
            s_q[threadIdx.x][8] = (S[0][0] * S[1][1] - S[0][1] * S[1][0]) * id;
            s_q[threadIdx.x][9] = opacity[g];
            const float4* s4 = reinterpret_cast<const float4*>(sh + 48 * (int64_t)g);
#pragma unroll
            for (int k = 0; k < 12; k++) {
                const float4 t4 = s4[k];
                s_sh[threadIdx.x][4 * k] = t4.x; s_sh[threadIdx.x][4 * k + 1] = t4.y;
                s_sh[threadIdx.x][4 * k + 2] = t4.z; s_sh[threadIdx.x][4 * k + 3] = t4.w;
            }
        }
        __syncthreads();
        const int m = (int)min((int64_t)B, end - b0);
        for (int j = 0; j < m; j++) {
            const float d0 = x0 - s_q[j][0], d1 = x1 - s_q[j][1], d2 = x2 - s_q[j][2];
            const float qf = s_q[j][3] * d0 * d0 + s_q[j][6] * d1 * d1 + s_q[j][8] * d2 * d2 +
                             2.0f * (s_q[j][4] * d0 * d1 + s_q[j][5] * d0 * d2 + s_q[j][7] * d1 * d2);
            const float a = s_q[j][9] * __expf(-0.5f * qf);
            aop += a;
#pragma unroll
            for (int k = 0; k < 48; k++) acc[k] += a * s_sh[j][k];
        }
        __syncthreads();
    }
    float4* o4 = reinterpret_cast<float4*>(feat + 48 * si);
#pragma unroll
    for (int k = 0; k < 12; k++) o4[k] = make_float4(acc[4 * k], acc[4 * k + 1], acc[4 * k + 2], acc[4 * k + 3]);
    opac[si] = aop;
}

// GetAdaLpfRatio (GaussianView.cpp:4670-4749), one thread per valid cell. The
// cell's samples are a 4x4x4 block (sample x*16 + y*4 + z); corner c has
// offsets q_c = (c&1, (c>>1)&1, (c>>2)&1) and the sample at 48 q_x + 12 q_y +
// 3 q_z. With p_c = (x_c - x_0) / 3, the reference solves the 24 x 9 least
// squares min sum_c |M q_c - p_c|^2 by the normal equations; they decouple
// per row of M, with Q Q^T = 2 I + 2 11^T over the cube's corners, whose
// inverse is 3/8 on the diagonal and -1/8 off it, so M = (P Q^T)(Q Q^T)^-1
// in closed form (evaluated in double). out = lpf_param * M M^T.
__global__ void __launch_bounds__(256) ada_lpf_kernel(int64_t nv, const float* __restrict__ samples,
                                                      float lpf_param, float* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const float* x = samples + 3 * 64 * v;
    double PQ[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};  // P Q^T
    for (int c = 1; c < 8; c++) {
        const int qx = c & 1, qy = (c >> 1) & 1, qz = (c >> 2) & 1;
        const float* xc = x + 3 * (48 * qx + 12 * qy + 3 * qz);
        for (int j = 0; j < 3; j++) {
            const double pj = (double)((xc[j] - x[j]) / 3.0f);
            if (qx) PQ[j][0] += pj;
            if (qy) PQ[j][1] += pj;
            if (qz) PQ[j][2] += pj;
        }
    }
    double M[3][3];
    for (int j = 0; j < 3; j++) {
        const double sum = PQ[j][0] + PQ[j][1] + PQ[j][2];
        for (int k = 0; k < 3; k++) M[j][k] = 0.5 * PQ[j][k] - 0.125 * sum;
    }
    double L[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] = (double)lpf_param * (M[i][0] * M[j][0] + M[i][1] * M[j][1] + M[i][2] * M[j][2]);
    float* o = out + 6 * v;
    o[0] = (float)L[0][0]; o[1] = (float)L[1][1]; o[2] = (float)L[2][2];
    o[3] = (float)L[0][1]; o[4] = (float)L[0][2]; o[5] = (float)L[1][2];
}

__global__ void zero_words_kernel(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

}  // namespace

int launch_minmax(const float* pos, int64_t n, uint32_t* mm_ws, float* out6, hipStream_t st) {
    hipLaunchKernelGGL(minmax_init_kernel, dim3(1), dim3(64), 0, st, mm_ws);
    hipLaunchKernelGGL(minmax_kernel, dim3(std::min<unsigned>(nblk(n, 256), 1024)), dim3(256), 0, st, pos, n, mm_ws);
    hipLaunchKernelGGL(decode_minmax_kernel, dim3(1), dim3(64), 0, st, mm_ws, out6);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_center_cells(const float* pos, int64_t n, const GridGeom& gg, uint64_t* keys, uint32_t* vals,
                        hipStream_t st) {
    hipLaunchKernelGGL(center_cell_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, pos, n, gg, keys, vals);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_aabbs(const float* pos, const float* rot, const float* scale, const float* opacity, int64_t n, float* out,
                 hipStream_t st) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(aabb_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, pos, rot, scale, opacity, n, out);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_box_count(const float* aabbs, int64_t n, const GridGeom& gg, uint32_t* cnt, hipStream_t st) {
    hipLaunchKernelGGL(box_count_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, aabbs, n, gg, cnt);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_sum_counts(const uint32_t* cnt, int64_t n, unsigned long long* out, hipStream_t st) {
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t*>(out), (int64_t)2);
    if (n > 0)
        hipLaunchKernelGGL(sum_counts_kernel, dim3(std::min<unsigned>(nblk(n, 256), 1024u)), dim3(256), 0, st, cnt, n,
                           out);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_slab_counts(const float* aabbs, int64_t n, const GridGeom& gg, unsigned long long* slab, hipStream_t st) {
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(256), 0, st, reinterpret_cast<uint32_t*>(slab),
                       (int64_t)2 * gg.G);
    if (n > 0) hipLaunchKernelGGL(slab_count_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, aabbs, n, gg, slab);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_zero_cells(uint32_t* cell_count, int64_t gc, hipStream_t st) {
    hipLaunchKernelGGL(zero_words_kernel, dim3(std::min<unsigned>(nblk(gc, 256), 4096)), dim3(256), 0, st, cell_count,
                       gc);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_box_count_slabs(const float* aabbs, int64_t n, const GridGeom& gg, int x0, int x1, uint32_t* cnt,
                           hipStream_t st) {
    hipLaunchKernelGGL(box_count_slabs_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, aabbs, n, gg, x0, x1, cnt);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_box_emit_slabs(const float* aabbs, int64_t n, const GridGeom& gg, int x0, int x1, const uint32_t* offs,
                          uint32_t* keys, uint32_t* vals, uint32_t* cell_count, hipStream_t st) {
    hipLaunchKernelGGL(box_emit_slabs_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, aabbs, n, gg, x0, x1, offs, keys,
                       vals, cell_count);
    HIP_RET(hipGetLastError());
    return 0;
}

size_t cell_prefix_workspace_words(int64_t gc) { return 2 * (size_t)((gc + SCAN64_TILE - 1) / SCAN64_TILE) + 2; }

int launch_cell_prefix(const uint32_t* cell_count, int64_t gc, int64_t* prefix, uint32_t* flags, uint32_t* ws,
                       hipStream_t st) {
    const int64_t nb = (gc + SCAN64_TILE - 1) / SCAN64_TILE;
    if (nb == 0) return 0;
    unsigned long long* sums = reinterpret_cast<unsigned long long*>(ws);
    hipLaunchKernelGGL(scan64_reduce_kernel, dim3((unsigned)nb), dim3(SCAN64_THREADS), 0, st, cell_count, gc, sums);
    hipLaunchKernelGGL(scan64_blocks_kernel, dim3(1), dim3(SCAN64_THREADS), 0, st, sums, nb);
    hipLaunchKernelGGL(scan64_apply_kernel, dim3((unsigned)nb), dim3(SCAN64_THREADS), 0, st, cell_count, gc,
                       (const unsigned long long*)sums, prefix, flags);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_compact_valid(const uint32_t* flags, const uint32_t* pos_ex, int64_t gc, int* valid, hipStream_t st) {
    hipLaunchKernelGGL(compact_valid_kernel, dim3(nblk(gc, 256)), dim3(256), 0, st, flags, pos_ex, gc, valid);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_samples(const int* valid, int64_t nv, const GridGeom& gg, float* samples, hipStream_t st) {
    if (nv == 0) return 0;
    hipLaunchKernelGGL(samples_kernel, dim3(nblk(nv * 64, 256)), dim3(256), 0, st, valid, nv, gg, samples);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_density(int64_t nv, const int* valid, const int64_t* prefix, const int* list, const float* samples,
                   const float* pos, const float* rot, const float* scale, const float* opacity, const float* sh,
                   DensityFilter flt, float* feat, float* opac, hipStream_t st) {
    if (nv == 0) return 0;
    hipLaunchKernelGGL(density_kernel, dim3((unsigned)nv), dim3(64), 0, st, nv, valid, prefix, list, samples, pos, rot,
                       scale, opacity, sh, flt, feat, opac);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_ada_lpf(int64_t nv, const float* samples, float lpf_param, float* out, hipStream_t st) {
    if (nv == 0) return 0;
    hipLaunchKernelGGL(ada_lpf_kernel, dim3(nblk(nv, 256)), dim3(256), 0, st, nv, samples, lpf_param, out);
    HIP_RET(hipGetLastError());
    return 0;
}

}  // namespace arapgs
