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

int launch_minmax(const float* pos, int64_t n, uint32_t* mm_ws, float* out6, hipStream_t st);
int launch_center_cells(const float* pos, int64_t n, const GridGeom& gg, uint64_t* keys, uint32_t* vals,
                        hipStream_t st);
int launch_aabbs(const float* pos, const float* rot, const float* scale, const float* opacity, int64_t n, float* out,
                 hipStream_t st);
int launch_box_count(const float* aabbs, int64_t n, const GridGeom& gg, uint32_t* cnt, hipStream_t st);
// *out (device, 8 bytes) = the 64-bit sum of cnt[0..n).
int launch_sum_counts(const uint32_t* cnt, int64_t n, unsigned long long* out, hipStream_t st);
int launch_box_emit(const float* aabbs, int64_t n, const GridGeom& gg, const uint32_t* offs, uint64_t* keys,
                    uint32_t* vals, uint32_t* cell_count, hipStream_t st);
int launch_cell_prefix(const uint32_t* cell_ex, const uint32_t* cell_count, int64_t gc, int* prefix, uint32_t* flags,
                       hipStream_t st);
int launch_compact_valid(const uint32_t* flags, const uint32_t* pos_ex, int64_t gc, int* valid, hipStream_t st);
int launch_samples(const int* valid, int64_t nv, const GridGeom& gg, float* samples, hipStream_t st);
int launch_density(int64_t nv, const int* valid, const int* prefix, const int* list, const float* samples,
                   const float* pos, const float* rot, const float* scale, const float* opacity, const float* sh,
                   float lpf, float* feat, float* opac, hipStream_t st);

}  // namespace arapgs
