// Device-wide exclusive scan of u32 counts and a stable LSD radix sort of
// (u64 key, u32 value) pairs for gfx950 (64-wide wavefronts).
//
// Radix sort of u32 or u64 keys, 8-bit digits, three kernels per pass:
//   1. histogram: every block counts the digits of its 4096-key tile (LDS
//      atomics), written digit-major hist[digit * nblocks + block];
//   2. exclusive scan of hist (digit-major order = the stable global order);
//   3. scatter: each wave ranks its 1024 keys 64 at a time; lanes with equal
//      digits are found with 8 ballots (one per digit bit), the rank is the
//      popcount of equal lanes below plus the wave's running digit count; wave
//      counts are prefix-summed across the block so the tile keeps its order.
//      The tile is then reordered by digit in LDS and written out in that
//      order, so consecutive lanes store consecutive slots of a digit's run
//      (about 16 keys per digit and tile) instead of 64 scattered lines.
#include "common.hpp"
#include "sort.hpp"

namespace arapgs {

namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;                       // keys per thread
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS; // 4096 keys per block
constexpr int WAVES = SORT_THREADS / 64;

// ---------------------------------------------------------------- scan
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    // wave inclusive scan via shuffles
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            const uint32_t c = sh[w];
            sh[w] = run;
            run += c;
        }
        sh[16] = run;
    }
    __syncthreads();
    const uint32_t r = sh[wave] + x - v;
    if (total) *total = sh[16];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(SCAN_THREADS) scan_reduce_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                                   uint32_t* __restrict__ block_sums) {
    __shared__ uint32_t sh[17];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++)
        if (base + i < n) s += in[base + i];
    uint32_t tot;
    block_exclusive_scan(s, sh, &tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// single block: exclusive scan of the block sums in place; writes the total
__global__ void __launch_bounds__(SCAN_THREADS) scan_blocks_kernel(uint32_t* __restrict__ sums, int nb,
                                                                   uint32_t* __restrict__ total) {
    __shared__ uint32_t sh[17];
    uint32_t carry = 0;
    for (int base = 0; base < nb; base += SCAN_THREADS) {
        const int i = base + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, sh, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(SCAN_THREADS) scan_apply_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                                  const uint32_t* __restrict__ block_off,
                                                                  uint32_t* __restrict__ out) {
    __shared__ uint32_t sh[17];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        s += v[i];
    }
    uint32_t run = block_exclusive_scan(s, sh, nullptr) + block_off[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
}

template <typename KeyT>
__global__ void __launch_bounds__(SORT_THREADS) radix_hist_kernel(const KeyT* __restrict__ keys, int64_t n,
                                                                  int shift, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
    for (int i = threadIdx.x; i < SORT_TILE; i += SORT_THREADS) {
        const int64_t k = base + i;
        if (k < n) atomicAdd(&h[(uint32_t)(keys[k] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

template <typename KeyT>
__global__ void __launch_bounds__(SORT_THREADS) radix_scatter_kernel(const KeyT* __restrict__ keys_in,
                                                                     const uint32_t* __restrict__ vals_in, int64_t n,
                                                                     int shift, const uint32_t* __restrict__ offs,
                                                                     KeyT* __restrict__ keys_out,
                                                                     uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t wave_cnt[WAVES][256];
    __shared__ uint32_t base_off[256];
    __shared__ uint32_t local_start[256];
    __shared__ uint32_t scan_sh[17];
    __shared__ KeyT s_key[SORT_TILE];
    __shared__ uint32_t s_val[SORT_TILE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < WAVES * 256; d += SORT_THREADS) (&wave_cnt[0][0])[d] = 0;
    base_off[threadIdx.x] = offs[(int64_t)threadIdx.x * gridDim.x + blockIdx.x];
    __syncthreads();
    const int64_t bbase = (int64_t)blockIdx.x * SORT_TILE;
    const int64_t wbase = bbase + (int64_t)wave * (SORT_ITEMS * 64);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    KeyT key[SORT_ITEMS];
    uint32_t val[SORT_ITEMS], rank[SORT_ITEMS];
    // ranks within the wave (stable, in key order)
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; it++) {
        const int64_t k = wbase + it * 64 + lane;
        const bool valid = k < n;
        key[it] = valid ? keys_in[k] : KeyT(0);
        val[it] = valid ? vals_in[k] : 0u;
        const uint32_t dg = (uint32_t)(key[it] >> shift) & 0xffu;
        uint64_t eq = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bal = __ballot((dg >> b) & 1u);
            eq &= ((dg >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t below = (uint32_t)__popcll(eq & lt_mask);
        const uint32_t prev = valid ? wave_cnt[wave][dg] : 0u;
        rank[it] = prev + below;
        // the highest lane of each digit group publishes the new count
        const bool leader = valid && ((eq >> lane) == 1ull);
        __builtin_amdgcn_wave_barrier();
        if (leader) wave_cnt[wave][dg] = prev + (uint32_t)__popcll(eq);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: exclusive prefix over the waves, then the tile's digit starts
    uint32_t tot = 0;
    {
        const int d = threadIdx.x;
#pragma unroll
        for (int w = 0; w < WAVES; w++) {
            const uint32_t c = wave_cnt[w][d];
            wave_cnt[w][d] = tot;
            tot += c;
        }
    }
    local_start[threadIdx.x] = block_exclusive_scan(tot, scan_sh, nullptr);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; it++) {
        const int64_t k = wbase + it * 64 + lane;
        if (k >= n) continue;
        const uint32_t dg = (uint32_t)(key[it] >> shift) & 0xffu;
        const uint32_t lp = local_start[dg] + wave_cnt[wave][dg] + rank[it];
        s_key[lp] = key[it];
        s_val[lp] = val[it];
    }
    __syncthreads();
    const int nloc = (int)min((int64_t)SORT_TILE, n - bbase);
    for (int i = threadIdx.x; i < nloc; i += SORT_THREADS) {
        const KeyT kk = s_key[i];
        const uint32_t dg = (uint32_t)(kk >> shift) & 0xffu;
        const uint32_t dst = base_off[dg] + (uint32_t)i - local_start[dg];
        keys_out[dst] = kk;
        vals_out[dst] = s_val[i];
    }
}

}  // namespace

size_t scan_workspace_words(int64_t n) { return (size_t)((n + SCAN_TILE - 1) / SCAN_TILE) + 1; }

int launch_exclusive_scan(const uint32_t* in, int64_t n, uint32_t* out, uint32_t* total, uint32_t* ws,
                          hipStream_t st) {
    const int nb = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
    if (nb == 0) {
        hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(SCAN_THREADS), 0, st, ws, 0, total);  // *total = 0
        HIP_RET(hipGetLastError());
        return 0;
    }
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(SCAN_THREADS), 0, st, in, n, ws);
    hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(SCAN_THREADS), 0, st, ws, nb, total);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(SCAN_THREADS), 0, st, in, n, ws, out);
    HIP_RET(hipGetLastError());
    return 0;
}

size_t radix_workspace_words(int64_t n) {
    const int64_t nb = (n + SORT_TILE - 1) / SORT_TILE;
    const int64_t h = 256 * std::max<int64_t>(nb, 1);
    return (size_t)(2 * h + scan_workspace_words(h) + 1);
}

template <typename KeyT>
static int radix_sort_impl(KeyT* keys, uint32_t* vals, KeyT* keys_alt, uint32_t* vals_alt, int64_t n, int bits,
                           uint32_t* ws, hipStream_t st) {
    if (n <= 1) return 0;
    const int64_t nb = (n + SORT_TILE - 1) / SORT_TILE;
    const int64_t h = 256 * nb;
    uint32_t* hist = ws;
    uint32_t* offs = ws + h;
    uint32_t* scan_ws = ws + 2 * h;
    uint32_t* scan_total = scan_ws + scan_workspace_words(h);
    KeyT* kin = keys;
    uint32_t* vin = vals;
    KeyT* kout = keys_alt;
    uint32_t* vout = vals_alt;
    const int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(radix_hist_kernel<KeyT>, dim3((unsigned)nb), dim3(SORT_THREADS), 0, st, kin, n, shift, hist);
        int rc = launch_exclusive_scan(hist, h, offs, scan_total, scan_ws, st);
        if (rc) return rc;
        hipLaunchKernelGGL(radix_scatter_kernel<KeyT>, dim3((unsigned)nb), dim3(SORT_THREADS), 0, st, kin, vin, n,
                           shift, offs, kout, vout);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    HIP_RET(hipGetLastError());
    if (kin != keys) {
        HIP_RET(hipMemcpyAsync(keys, kin, sizeof(KeyT) * n, hipMemcpyDeviceToDevice, st));
        HIP_RET(hipMemcpyAsync(vals, vin, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    }
    return 0;
}

// Sorts keys[0,n) by bits [0, bits) (stable); the result ends in (keys, vals)
// (the alternate buffers are scratch).
int launch_radix_sort(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, int64_t n, int bits,
                      uint32_t* ws, hipStream_t st) {
    return radix_sort_impl<uint64_t>(keys, vals, keys_alt, vals_alt, n, bits, ws, st);
}

int launch_radix_sort32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int64_t n, int bits,
                        uint32_t* ws, hipStream_t st) {
    return radix_sort_impl<uint32_t>(keys, vals, keys_alt, vals_alt, n, bits, ws, st);
}

}  // namespace arapgs
