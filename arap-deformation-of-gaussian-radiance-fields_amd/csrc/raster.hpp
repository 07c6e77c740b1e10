// Rasteriser launchers (raster.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace arapgs {

struct RasterParams {
    int P, D, W, H, tile, gx, gy;
    const float* bg;       // device, 3
    const float* means;    // P*3
    const float* shs;      // P*48
    const float* opac;     // P
    const float* scales;   // P*3
    const float* rots;     // P*4 (w,x,y,z)
    const float* view;     // device, 16 column-major
    const float* proj;     // device, 16 column-major
    const float* campos;   // device, 3
    float scale_mod, tanfovx, tanfovy, focal_x, focal_y;
    int use_rects, use_box;
    float boxmin[3], boxmax[3];
};

size_t raster_geo_bytes(int P);
// Preprocess + scan of tiles touched; *d_total (device) = instances to render.
int raster_preprocess(const RasterParams& rp, void* geo_ws, uint32_t* scan_ws, uint32_t* d_total, int* radii,
                      int* rects, hipStream_t st);
// Keys, sort, ranges and blend for n_rendered instances.
int raster_bin_and_blend(const RasterParams& rp, void* geo_ws, uint32_t n_rendered, const uint32_t* d_total,
                         uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t* sort_ws,
                         uint2* ranges, float* out, int* surface, hipStream_t st);

}  // namespace arapgs
