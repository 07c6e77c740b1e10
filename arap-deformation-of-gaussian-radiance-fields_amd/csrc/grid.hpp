// Grid / density launchers (grid.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace arapgs {

struct GridGeom {
    float mn[3];
    float step;
    int G;
    int padding;
};

// The low-pass filter added to a density Gaussian's covariance: lpf I, or,
// when ada is set, the valid cell's own symmetric matrix ada[6v .. 6v+5]
// (xx, yy, zz, xy, xz, yz; the adaptive filter of GetAdaLpfRatio,
// GaussianView.cpp:4670-4749). With ada set, the density backward hands
// dL/dSigma (not dL/dSigma^-1) to the per-Gaussian kernels.
struct DensityFilter {
    float lpf;
    const float* ada;
};

// Adds the filter of valid cell v to the covariance S.
__device__ __forceinline__ void density_filter_add(const DensityFilter& f, int64_t v, float S[3][3]) {
    if (f.ada) {
        const float* L = f.ada + 6 * v;
        S[0][0] += L[0]; S[1][1] += L[1]; S[2][2] += L[2];
        S[0][1] += L[3]; S[1][0] += L[3];
        S[0][2] += L[4]; S[2][0] += L[4];
        S[1][2] += L[5]; S[2][1] += L[5];
    } else {
        S[0][0] += f.lpf; S[1][1] += f.lpf; S[2][2] += f.lpf;
    }
}

int launch_minmax(const float* pos, int64_t n, uint32_t* mm_ws, float* out6, hipStream_t st);
int launch_center_cells(const float* pos, int64_t n, const GridGeom& gg, uint64_t* keys, uint32_t* vals,
                        hipStream_t st);
int launch_aabbs(const float* pos, const float* rot, const float* scale, const float* opacity, int64_t n, float* out,
                 hipStream_t st);
int launch_box_count(const float* aabbs, int64_t n, const GridGeom& gg, uint32_t* cnt, hipStream_t st);
// *out (device, 8 bytes) = the 64-bit sum of cnt[0..n).
int launch_sum_counts(const uint32_t* cnt, int64_t n, unsigned long long* out, hipStream_t st);
// Pairs per emit + sort pass of the cell-list build (temporary buffers of
// 16 bytes per pair), and the most one x-slab may hold (32-bit offsets).
constexpr unsigned long long GRID_CHUNK_PAIRS = 1ull << 28;
constexpr unsigned long long GRID_CHUNK_PAIRS_MAX = 0x7fffffffull;

// Cell lists are built in chunks of whole x-slabs (cells x*G^2 .. (x+1)*G^2-1)
// so that each chunk's pairs fit 32-bit sort offsets; the per-cell prefix is
// 64-bit. slab: 2*G words (G u64 pair counts).
int launch_slab_counts(const float* aabbs, int64_t n, const GridGeom& gg, unsigned long long* slab, hipStream_t st);
int launch_zero_cells(uint32_t* cell_count, int64_t gc, hipStream_t st);
int launch_box_count_slabs(const float* aabbs, int64_t n, const GridGeom& gg, int x0, int x1, uint32_t* cnt,
                           hipStream_t st);
int launch_box_emit_slabs(const float* aabbs, int64_t n, const GridGeom& gg, int x0, int x1, const uint32_t* offs,
                          uint32_t* keys, uint32_t* vals, uint32_t* cell_count, hipStream_t st);
// prefix[c] (int64) = pairs in cells 0..c, flags[c] = 1 if cell c holds any;
// ws holds cell_prefix_workspace_words(gc) words.
size_t cell_prefix_workspace_words(int64_t gc);
int launch_cell_prefix(const uint32_t* cell_count, int64_t gc, int64_t* prefix, uint32_t* flags, uint32_t* ws,
                       hipStream_t st);
int launch_compact_valid(const uint32_t* flags, const uint32_t* pos_ex, int64_t gc, int* valid, hipStream_t st);
int launch_samples(const int* valid, int64_t nv, const GridGeom& gg, float* samples, hipStream_t st);
int launch_density(int64_t nv, const int* valid, const int64_t* prefix, const int* list, const float* samples,
                   const float* pos, const float* rot, const float* scale, const float* opacity, const float* sh,
                   DensityFilter flt, float* feat, float* opac, hipStream_t st);
// GetAdaLpfRatio (GaussianView.cpp:4670-4749) for nv valid cells: the linear
// map M that best carries the unit cube's corners to the cell's 8 corner
// samples (relative to sample 0, over the 3 sample intervals), and
// out[6v ..] = lpf_param * M M^T.
int launch_ada_lpf(int64_t nv, const float* samples, float lpf_param, float* out, hipStream_t st);

}  // namespace arapgs
