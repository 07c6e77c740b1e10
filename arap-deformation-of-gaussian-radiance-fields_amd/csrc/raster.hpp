// Rasteriser launchers (raster.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>

namespace arapgs {

// One camera of a (multi-view) render. Lives in device memory.
struct RasterView {
    const float* view;     // device, 16 column-major
    const float* proj;     // device, 16 column-major
    const float* campos;   // device, 3
    float tanfovx, tanfovy, focal_x, focal_y;  // orthographic: focal_* = pixels per view-space unit
    int ortho;             // 1: orthographic projection (constant EWA Jacobian)
    float* out;            // device, 3*H*W planar CHW
    int* surface;          // optional, P
    int* radii;            // optional, P
    int* rects;            // optional, 2*P
    const float* dout;     // backward only: dL/d out (3*H*W planar CHW)
};
constexpr int RASTER_MAX_VIEWS = 64;

// V views of one set of Gaussians at one resolution. Per-view records are laid
// out view-major (index v*P + g); keys carry the combined tile id v*ntiles + t.
// Segment mode (seg != null, device int[V+1] ascending Gaussian offsets with
// seg[0] = 0, seg[V] = P): view v draws only Gaussians [seg[v], seg[v+1]) —
// independent scenes stored back to back — and each Gaussian has one record
// (index g).
struct RasterParams {
    int P, D, W, H, tile, gx, gy, V;
    const float* bg;       // device, 3
    const float* means;    // P*3
    const float* shs;      // P*48
    const float* opac;     // P
    const float* scales;   // P*3
    const float* rots;     // P*4 (w,x,y,z)
    float scale_mod;
    int use_rects, use_box;
    int unroll;            // Gaussians per compositing step of the 16x16 blend (2, 4, 8)
    int split;             // 16x16 tiles holding more instances than this are blended in segments of
                           // this many by separate workgroups and composited (0: one workgroup per tile)
    int cull;              // 1: a record emits only the tiles of its rectangle that pass the conservative
                           // tile test (rectangles of at most RASTER_CULL_MAX_AREA tiles)
    float boxmin[3], boxmax[3];
    const RasterView* views;  // device, V
    const int* seg;           // device, V+1, or null
};

inline int64_t raster_records(const RasterParams& rp) { return rp.seg ? (int64_t)rp.P : (int64_t)rp.P * rp.V; }

// The view whose segment holds Gaussian g: the last v with seg[v] <= g.
__host__ __device__ inline int segment_of(const int* seg, int V, int g) {
    int lo = 0, hi = V - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Conservative test whether a Gaussian can reach alpha >= 1/255 at any pixel
// centre of the box [x0, x1] x [y0, y1]. The blend skips a pixel when
// o * exp(power) < 1/255 with power = -q/2, q = a dx^2 + 2 b dx dy + c dy^2
// (conic (a, b, c), d = mean - pixel). The minimum of the convex q over the box
// is 0 when the mean lies inside, otherwise it sits on one of the four edges at
// the clamped 1-D minimiser. A Gaussian is dropped only when that minimum
// exceeds the threshold by a margin (1e-3 + 1e-5 * sum of |terms|) far above
// the float rounding of the per-pixel evaluation, so every pixel would have
// skipped it too and the composite is unchanged bit for bit.
__host__ __device__ inline float cull_log_threshold(float opacity) { return logf(255.0f * opacity) + 1e-3f; }

__host__ __device__ inline bool box_may_contribute(float2 xy, float4 co, float lthr, float x0, float y0, float x1,
                                                   float y1) {
    const float a = co.x, b = co.y, c = co.z;
    const float dxl = xy.x - x1, dxh = xy.x - x0, dyl = xy.y - y1, dyh = xy.y - y0;
    if (dxl <= 0.0f && dxh >= 0.0f && dyl <= 0.0f && dyh >= 0.0f) return true;
    if (!(a > 0.0f && c > 0.0f && a * c - b * b > 0.0f)) return true;
    const float y_l = fminf(dyh, fmaxf(dyl, -b * dxl / c));
    const float y_h = fminf(dyh, fmaxf(dyl, -b * dxh / c));
    const float x_l = fminf(dxh, fmaxf(dxl, -b * dyl / a));
    const float x_h = fminf(dxh, fmaxf(dxl, -b * dyh / a));
    const float q0 = a * dxl * dxl + 2.0f * b * dxl * y_l + c * y_l * y_l;
    const float q1 = a * dxh * dxh + 2.0f * b * dxh * y_h + c * y_h * y_h;
    const float q2 = a * x_l * x_l + 2.0f * b * x_l * dyl + c * dyl * dyl;
    const float q3 = a * x_h * x_h + 2.0f * b * x_h * dyh + c * dyh * dyh;
    const float qmin = fminf(fminf(q0, q1), fminf(q2, q3));
    const float mx = fmaxf(fabsf(dxl), fabsf(dxh)), my = fmaxf(fabsf(dyl), fabsf(dyh));
    const float mag = a * mx * mx + c * my * my + 2.0f * fabsf(b) * mx * my;
    return !(0.5f * qmin > lthr + 1e-5f * mag);
}

__host__ __device__ inline bool tile_may_contribute(float2 xy, float4 co, float x0, float y0, float x1, float y1) {
    return box_may_contribute(xy, co, cull_log_threshold(co.w), x0, y0, x1, y1);
}

// Emission-time tile culling (RasterParams::cull): a record whose rectangle
// spans at most this many tiles emits only the tiles whose box (pixel centres
// inside the image) passes box_may_contribute. Every sub-box the blend tests
// (the 8x8 quadrants of a 16x16 tile, or the 8x8 tile itself) lies inside its
// tile's box, so a record dropped from a tile is one the blend would have
// dropped from every sub-box: the images are unchanged bit for bit, only the
// instance lists are shorter. Larger rectangles emit every tile.
constexpr int RASTER_CULL_MAX_AREA = 64;

size_t raster_geo_bytes(int64_t PV);
// Preprocess of all V views + scan of tiles touched. d_counts (device, 3 words):
// [0] instances to render, [1] / [2] min / max depth bits of the visible
// records. scan_ws sized for V*P.
int raster_preprocess(const RasterParams& rp, void* geo_ws, uint32_t* scan_ws, uint32_t* d_counts, hipStream_t st);
// Binning, ranges and blend for n_rendered instances: the V*P records are
// sorted by depth bits - dmin (dbits wide), the records' tiles are emitted in
// that order and the instances are stably sorted by tile id (keys = view *
// ntiles + tile). rec_ws holds raster_record_bytes(V*P) bytes, scan_ws is the
// preprocess's scan workspace (sized for V*P).
size_t raster_record_bytes(int64_t PV);
// Workspace of the deep-tile split (rp.split > 0) for nt tiles and n_rendered
// instances: per-tile segment counts and offsets, their scan and one
// (colour, transmittance) float4 per pixel of every segment.
size_t raster_split_bytes(int64_t nt, uint32_t n_rendered, int split);
int raster_bin_and_blend(const RasterParams& rp, void* geo_ws, uint32_t n_rendered, const uint32_t* d_total,
                         uint32_t dmin, int dbits, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                         uint32_t* vals_alt, uint32_t* sort_ws, uint2* ranges, void* rec_ws, uint32_t* scan_ws,
                         void* split_ws, hipStream_t st);

// Backward of the blend for every view of the last binning (geo, vals,
// ranges as left by raster_bin_and_blend): per record gv the gradient of
// (x, y, conic a, b, c, r, g, b) accumulates into grec, component-major
// (component k at k * V * P + gv; zeroed by the caller). Each tile re-runs its forward pass to recover every pixel's final
// transmittance and last contributor, then walks its list back to front.
int raster_blend_backward(const RasterParams& rp, void* geo_ws, const uint32_t* vals, const uint2* ranges,
                          float* grec, hipStream_t st);
// Backward of the preprocess: per Gaussian, over all views, from grec to the
// gradients of means (3), rotations (4, as given, unnormalised), scales (3)
// and SH (48), written (not accumulated) to gout[58*g..] in that order.
int raster_preprocess_backward(const RasterParams& rp, void* geo_ws, const float* grec, float* gout, hipStream_t st);

}  // namespace arapgs
