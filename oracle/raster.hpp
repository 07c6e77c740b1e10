// TEST INFRASTRUCTURE — the oracle (see common.hpp). Arguments of the
// oracle's 3DGS rasteriser (raster.cpp), shared with the chain (chain.cpp).
#pragma once
#include <cstdint>

extern "C" {

struct orc_raster_args {
    int P, D, W, H, tile;           // tile edge in pixels
    const float* bg;                // 3
    const float* means;             // P*3
    const float* shs;               // P*48 (interleaved coef*3+ch)
    const float* opac;              // P
    const float* scales;            // P*3
    float scale_mod;
    const float* rots;              // P*4 (w,x,y,z)
    const float* view;              // 16, column-major world->camera
    const float* proj;              // 16, column-major world->clip
    const float* campos;            // 3
    float tanfovx, tanfovy;
    int use_rects;                  // per-axis 3-sigma rectangles (fast culling)
    const float* boxmin;            // crop box or NULL
    const float* boxmax;
    int ortho;                      // orthographic camera (setOrthoCam, GaussianView.cpp:1018-1019)
    float ortho_scale;              // its half extent in view units (x_ndc = x_view / ortho_scale)
};

void orc_raster_preprocess(const orc_raster_args* a, float* depth, int* radius, float* xy, float* conic, float* rgb,
                           int* rect, int* touched);
long orc_raster_bin(const orc_raster_args* a, const float* depth, const int* rect, const int* touched,
                    uint64_t* keys, uint32_t* vals, int* ranges);

}  // extern "C"
