// 3D Gaussian Splatting forward rasteriser for gfx950 (the path behind
// CudaRasterizer::Rasterizer::forward, called at GaussianView.cpp:1099-1128).
//
//   preprocess  one thread per Gaussian: view/clip transform, EWA 2D covariance,
//               conic, 3-sigma extent, tile rectangle, view-dependent SH colour
//   bin         exclusive scan of tiles touched, (tile << 32 | depth bits) keys,
//               stable radix sort, per-tile [start, end) ranges
//   blend       one 16x16 workgroup per tile; Gaussians staged through LDS in
//               256-wide batches, front-to-back alpha compositing with early
//               termination; planar CHW float output
//
// Everything that feeds a key or a tile rectangle is evaluated with single
// float operations in a fixed order and no contraction, so keys, tile ranges
// and sort order are reproducible bit for bit on the CPU.
#pragma clang fp contract(off)
#include "common.hpp"
#include "raster.hpp"
#include "sort.hpp"

namespace arapgs {

namespace {

__constant__ float kShC0 = 0.28209479177387814f;
__constant__ float kShC1 = 0.4886025119029199f;

struct Geo {
    float* depth;      // P
    float2* xy;        // P
    float4* conic;     // P  (a, b, c of the inverse 2D covariance, opacity)
    float4* rgb;       // P  (r, g, b, pad)
    int4* rect;        // P  tile rectangle [x0, y0, x1, y1)
    uint32_t* touched; // P
};

// The Gaussian falloff exp(power), power <= 0, of the blend: the hardware
// base-2 exponential of power * log2(e) (a few ulp; the per-pixel test
// thresholds, 1/255 on alpha and the conservative tile test's 1e-3 margin, are
// far coarser). The forward kernels and the backward's forward re-run share
// it, so both see the same contributors.
__device__ __forceinline__ float blend_exp(float power) { return __expf(power); }

// The conservative test of record (xy, co) against tile (tx, ty)'s box of pixel
// centres inside the image.
__device__ __forceinline__ bool cull_tile_pass(const RasterParams& rp, float2 xy, float4 co, float lthr, int tx,
                                               int ty) {
    const int x0 = tx * rp.tile, y0 = ty * rp.tile;
    return box_may_contribute(xy, co, lthr, (float)x0, (float)y0, (float)min(x0 + rp.tile - 1, rp.W - 1),
                              (float)min(y0 + rp.tile - 1, rp.H - 1));
}

// One Gaussian in one view (record index gv = v*P + g, or g in segment mode);
// the view's optional per-Gaussian outputs are indexed by gl (g, or in segment
// mode the Gaussian's index within its view's segment).
__device__ __forceinline__ void project_one(const RasterParams& rp, const RasterView& vw, int g, int64_t gv, Geo& geo,
                                            const float* __restrict__ sh, int gl) {
    int* radii = vw.radii;
    int* rects_out = vw.rects;
    geo.touched[gv] = 0;
    if (radii) radii[gl] = 0;
    const float px = rp.means[3 * g], py = rp.means[3 * g + 1], pz = rp.means[3 * g + 2];
    if (rp.use_box && (px < rp.boxmin[0] || py < rp.boxmin[1] || pz < rp.boxmin[2] || px > rp.boxmax[0] ||
                       py > rp.boxmax[1] || pz > rp.boxmax[2]))
        return;
    const float* V = vw.view;
    const float* Pm = vw.proj;
    const float vx = V[0] * px + V[4] * py + V[8] * pz + V[12];
    const float vy = V[1] * px + V[5] * py + V[9] * pz + V[13];
    const float vz = V[2] * px + V[6] * py + V[10] * pz + V[14];
    if (vz <= 0.2f) return;
    const float hx = Pm[0] * px + Pm[4] * py + Pm[8] * pz + Pm[12];
    const float hy = Pm[1] * px + Pm[5] * py + Pm[9] * pz + Pm[13];
    const float hw = Pm[3] * px + Pm[7] * py + Pm[11] * pz + Pm[15];
    const float pw = 1.0f / (hw + 0.0000001f);
    const float ndx = hx * pw, ndy = hy * pw;
    const float qw = rp.rots[4 * g], qx = rp.rots[4 * g + 1], qy = rp.rots[4 * g + 2], qz = rp.rots[4 * g + 3];
    float R[3][3];
    R[0][0] = 1.0f - 2.0f * (qy * qy + qz * qz);
    R[0][1] = 2.0f * (qx * qy - qw * qz);
    R[0][2] = 2.0f * (qx * qz + qw * qy);
    R[1][0] = 2.0f * (qx * qy + qw * qz);
    R[1][1] = 1.0f - 2.0f * (qx * qx + qz * qz);
    R[1][2] = 2.0f * (qy * qz - qw * qx);
    R[2][0] = 2.0f * (qx * qz - qw * qy);
    R[2][1] = 2.0f * (qy * qz + qw * qx);
    R[2][2] = 1.0f - 2.0f * (qx * qx + qy * qy);
    float s2[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float s = rp.scale_mod * rp.scales[3 * g + k];
        s2[k] = s * s;
    }
    float S[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            S[i][j] = R[i][0] * s2[0] * R[j][0] + R[i][1] * s2[1] * R[j][1] + R[i][2] * s2[2] * R[j][2];
    float j00, j02, j11, j12;
    if (vw.ortho) {
        j00 = vw.focal_x; j02 = 0.0f;
        j11 = vw.focal_y; j12 = 0.0f;
    } else {
        const float limx = 1.3f * vw.tanfovx, limy = 1.3f * vw.tanfovy;
        const float txtz = vx / vz, tytz = vy / vz;
        const float tx = fminf(limx, fmaxf(-limx, txtz)) * vz;
        const float ty = fminf(limy, fmaxf(-limy, tytz)) * vz;
        j00 = vw.focal_x / vz; j02 = -(vw.focal_x * tx) / (vz * vz);
        j11 = vw.focal_y / vz; j12 = -(vw.focal_y * ty) / (vz * vz);
    }
    float T[2][3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        T[0][c] = j00 * V[c * 4 + 0] + j02 * V[c * 4 + 2];
        T[1][c] = j11 * V[c * 4 + 1] + j12 * V[c * 4 + 2];
    }
    float TS[2][3];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) TS[r][c] = T[r][0] * S[0][c] + T[r][1] * S[1][c] + T[r][2] * S[2][c];
    const float ca = TS[0][0] * T[0][0] + TS[0][1] * T[0][1] + TS[0][2] * T[0][2] + 0.3f;
    const float cb = TS[0][0] * T[1][0] + TS[0][1] * T[1][1] + TS[0][2] * T[1][2];
    const float cc = TS[1][0] * T[1][0] + TS[1][1] * T[1][1] + TS[1][2] * T[1][2] + 0.3f;
    const float det = ca * cc - cb * cb;
    if (det == 0.0f) return;
    const float dinv = 1.0f / det;
    const float mid = 0.5f * (ca + cc);
    const float disc = __builtin_sqrtf(fmaxf(0.1f, mid * mid - det));
    const float l1 = mid + disc, l2 = mid - disc;
    const int rad = (int)ceilf(3.0f * __builtin_sqrtf(fmaxf(l1, l2)));
    const float ix = ((ndx + 1.0f) * (float)rp.W - 1.0f) * 0.5f;
    const float iy = ((ndy + 1.0f) * (float)rp.H - 1.0f) * 0.5f;
    int ex = rad, ey = rad;
    if (rp.use_rects) {
        ex = (int)ceilf(3.0f * __builtin_sqrtf(ca));
        ey = (int)ceilf(3.0f * __builtin_sqrtf(cc));
    }
    const float tf = (float)rp.tile;
    const int r0x = min(rp.gx, max(0, (int)((ix - (float)ex) / tf)));
    const int r0y = min(rp.gy, max(0, (int)((iy - (float)ey) / tf)));
    const int r1x = min(rp.gx, max(0, (int)((ix + (float)ex + tf - 1.0f) / tf)));
    const int r1y = min(rp.gy, max(0, (int)((iy + (float)ey + tf - 1.0f) / tf)));
    const int area = (r1x - r0x) * (r1y - r0y);
    if (area == 0) return;
    // view-dependent colour (degree D)
    float dx = px - vw.campos[0], dy = py - vw.campos[1], dz = pz - vw.campos[2];
    const float dn = __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
    dx /= dn; dy /= dn; dz /= dn;
    float col[3];
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        float r = kShC0 * sh[ch];
        if (rp.D > 0) {
            r = r - kShC1 * dy * sh[3 + ch] + kShC1 * dz * sh[6 + ch] - kShC1 * dx * sh[9 + ch];
            if (rp.D > 1) {
                const float xx = dx * dx, yy = dy * dy, zz = dz * dz, xy2 = dx * dy, yz = dy * dz, xz = dx * dz;
                r = r + 1.0925484305920792f * xy2 * sh[12 + ch] + -1.0925484305920792f * yz * sh[15 + ch] +
                    0.31539156525252005f * (2.0f * zz - xx - yy) * sh[18 + ch] +
                    -1.0925484305920792f * xz * sh[21 + ch] + 0.5462742152960396f * (xx - yy) * sh[24 + ch];
                if (rp.D > 2) {
                    r = r + -0.5900435899266435f * dy * (3.0f * xx - yy) * sh[27 + ch] +
                        2.890611442640554f * xy2 * dz * sh[30 + ch] +
                        -0.4570457994644658f * dy * (4.0f * zz - xx - yy) * sh[33 + ch] +
                        0.3731763325901154f * dz * (2.0f * zz - 3.0f * xx - 3.0f * yy) * sh[36 + ch] +
                        -0.4570457994644658f * dx * (4.0f * zz - xx - yy) * sh[39 + ch] +
                        1.445305721320277f * dz * (xx - yy) * sh[42 + ch] +
                        -0.5900435899266435f * dx * (xx - 3.0f * yy) * sh[45 + ch];
                }
            }
        }
        col[ch] = fmaxf(r + 0.5f, 0.0f);
    }
    const float2 gxy = make_float2(ix, iy);
    const float4 gco = make_float4(cc * dinv, -cb * dinv, ca * dinv, rp.opac[g]);
    uint32_t count = (uint32_t)area;
    if (rp.cull && area <= RASTER_CULL_MAX_AREA) {
        // the tiles that pass the conservative test (emit_tiles_cull_kernel
        // repeats the same test on the same values)
        const float lthr = cull_log_threshold(gco.w);
        count = 0u;
        for (int ty = r0y; ty < r1y; ty++)
            for (int tx = r0x; tx < r1x; tx++) count += cull_tile_pass(rp, gxy, gco, lthr, tx, ty) ? 1u : 0u;
    }
    geo.depth[gv] = vz;
    geo.xy[gv] = gxy;
    geo.conic[gv] = gco;
    geo.rgb[gv] = make_float4(col[0], col[1], col[2], 0.0f);
    geo.rect[gv] = make_int4(r0x, r0y, r1x, r1y);
    geo.touched[gv] = count;
    if (radii) radii[gl] = rad;
    if (rects_out) { rects_out[2 * gl] = ex; rects_out[2 * gl + 1] = ey; }
}

// One thread per Gaussian for all V views: the 48 SH coefficients are read once.
// Also folds the depth-bit range of the visible records into dminmax[0..1]
// (bit patterns of positive floats order like the floats).
__global__ void __launch_bounds__(256) preprocess_kernel(RasterParams rp, Geo geo, uint32_t* __restrict__ dminmax) {
    __shared__ uint32_t s_mn[4], s_mx[4];
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t mn = 0xffffffffu, mx = 0u;
    if (g < rp.P) {
        float sh[48];
        const float4* s4 = reinterpret_cast<const float4*>(rp.shs + 48 * (int64_t)g);
#pragma unroll
        for (int i = 0; i < 12; i++) {
            const float4 q = s4[i];
            sh[4 * i] = q.x; sh[4 * i + 1] = q.y; sh[4 * i + 2] = q.z; sh[4 * i + 3] = q.w;
        }
        if (rp.seg) {
            // segment mode: one record per Gaussian, in the view whose segment holds it
            const int v = segment_of(rp.seg, rp.V, g);
            const RasterView vw = rp.views[v];
            project_one(rp, vw, g, g, geo, sh, g - rp.seg[v]);
            if (geo.touched[g]) {
                const uint32_t db = __float_as_uint(geo.depth[g]);
                mn = min(mn, db);
                mx = max(mx, db);
            }
        } else {
            for (int v = 0; v < rp.V; v++) {
                const RasterView vw = rp.views[v];
                const int64_t gv = (int64_t)v * rp.P + g;
                project_one(rp, vw, g, gv, geo, sh, g);
                if (geo.touched[gv]) {
                    const uint32_t db = __float_as_uint(geo.depth[gv]);
                    mn = min(mn, db);
                    mx = max(mx, db);
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor((int)mn, o));
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, o));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { s_mn[wv] = mn; s_mx[wv] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) { mn = min(mn, s_mn[w]); mx = max(mx, s_mx[w]); }
        if (mn <= mx) {
            atomicMin(&dminmax[0], mn);
            atomicMax(&dminmax[1], mx);
        }
    }
}

__global__ void depth_range_init_kernel(uint32_t* __restrict__ dminmax) {
    dminmax[0] = 0xffffffffu;
    dminmax[1] = 0u;
}

// Depth-first binning. The records are sorted once by depth (their key is the
// depth bit pattern minus the frame's minimum, so the order is the float order;
// the stable sort keeps equal depths in record order), then every record emits
// its tiles in that order and the instances are stably sorted by tile alone:
// each tile's list ends up ordered by (depth bits, record), exactly the order of
// a stable sort of (tile << 32 | depth bits) keys emitted in record order. The
// wide sort runs on the V*P records instead of the instances, and the instance
// sort needs only the tile-id bits.
__global__ void __launch_bounds__(256) record_keys_kernel(int64_t PV, Geo geo, uint32_t dmin,
                                                          uint32_t* __restrict__ rkeys, uint32_t* __restrict__ rvals) {
    const int64_t gv = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gv >= PV) return;
    rkeys[gv] = geo.touched[gv] ? __float_as_uint(geo.depth[gv]) - dmin : 0u;
    rvals[gv] = (uint32_t)gv;
}

__global__ void __launch_bounds__(256) gather_touched_kernel(int64_t PV, const uint32_t* __restrict__ rvals,
                                                             const uint32_t* __restrict__ touched,
                                                             uint32_t* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= PV) return;
    out[r] = touched[rvals[r]];
}

// The tiles of the depth-ordered records: the keys (view * ntiles + tile) and
// values (the record index) of record r fill slots [offsets[r], offsets[r] +
// count) in row-major rect order. One wave per 64 consecutive records writes
// their slots, which are consecutive, with consecutive lanes on consecutive
// slots: the lane of slot j finds its record (the last of the 64 whose first
// slot is <= j) by a binary search over the lanes' first slots.
__global__ void __launch_bounds__(256) emit_tiles_kernel(int64_t PV, int P, const int* __restrict__ seg, int V,
                                                         int gx, int ntiles, Geo geo,
                                                         const uint32_t* __restrict__ rvals,
                                                         const uint32_t* __restrict__ offsets,
                                                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = r < PV;
    uint32_t gv = 0, first = 0xffffffffu, cnt = 0, base = 0;
    int rx = 0, ry = 0, rw = 1;
    if (valid) {
        gv = rvals[r];
        first = offsets[r];
        cnt = geo.touched[gv];
        if (cnt) {
            const int4 rc = geo.rect[gv];
            rx = rc.x;
            ry = rc.y;
            rw = rc.z - rc.x;
            base = (seg ? (uint32_t)segment_of(seg, V, (int)gv) : gv / (uint32_t)P) * (uint32_t)ntiles;
        }
    }
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0ull) return;
    const int last = 63 - __builtin_clzll(vmask);
    const uint32_t begin = __shfl(first, 0);
    const uint32_t end = __shfl(first + cnt, last);
    // uniform trip count: every lane takes part in the shuffles (a shuffle from
    // a lane that left a divergent loop does not return that lane's value)
    for (uint32_t jb = begin; jb < end; jb += 64) {
        const uint32_t j = jb + (uint32_t)lane;
        int lo = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
            const uint32_t f = __shfl(first, lo + step);
            if (f <= j) lo += step;
        }
        const uint32_t f0 = __shfl(first, lo);
        const int w = __shfl(rw, lo);
        const int x0 = __shfl(rx, lo), y0 = __shfl(ry, lo);
        const uint32_t b = __shfl(base, lo);
        const uint32_t g = __shfl(gv, lo);
        if (j < end) {
            const int q = (int)(j - f0);
            const int y = y0 + q / w, x = x0 + q % w;
            keys[j] = b + (uint32_t)(y * gx + x);
            vals[j] = g;
        }
    }
}

// emit_tiles_kernel with emission-time culling: the wave walks the rectangle
// slots of its 64 records (a wave-local scan of the rectangle areas), tests
// each slot's tile as the preprocess counted it and writes the passing ones,
// in rect order, to [offsets[r], offsets[r] + touched) by their rank (the
// passing slots of the record in earlier 64-slot chunks, carried by the
// record's lane, plus those below the slot in this chunk). A write past the
// record's count is never made.
__global__ void __launch_bounds__(256) emit_tiles_cull_kernel(RasterParams rp, int64_t PV, int ntiles, Geo geo,
                                                              const uint32_t* __restrict__ rvals,
                                                              const uint32_t* __restrict__ offsets,
                                                              uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t gv = 0, first = 0, cnt = 0, area = 0, base = 0;
    int rx = 0, ry = 0, rw = 1, test = 0;
    float2 xy = make_float2(0.0f, 0.0f);
    float4 co = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float lthr = 0.0f;
    if (r < PV) {
        gv = rvals[r];
        first = offsets[r];
        cnt = geo.touched[gv];
        if (cnt) {
            const int4 rc = geo.rect[gv];
            rx = rc.x;
            ry = rc.y;
            rw = rc.z - rc.x;
            area = (uint32_t)(rw * (rc.w - rc.y));
            base = (rp.seg ? (uint32_t)segment_of(rp.seg, rp.V, (int)gv) : gv / (uint32_t)rp.P) * (uint32_t)ntiles;
            test = (rp.cull && (int)area <= RASTER_CULL_MAX_AREA) ? 1 : 0;
            xy = geo.xy[gv];
            co = geo.conic[gv];
            lthr = cull_log_threshold(co.w);
        }
    }
    // wave-inclusive scan of the areas: lane l's slots are [afirst, afirst + area)
    uint32_t incl = area;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    const uint32_t afirst = incl - area;
    const uint32_t total = __shfl(incl, 63);
    uint32_t carry = 0;  // passing slots of this lane's record in earlier chunks
    // uniform trip count: every lane takes part in every shuffle
    for (uint32_t jb = 0; jb < total; jb += 64) {
        const uint32_t j = jb + (uint32_t)lane;
        int lo = 0;  // the last lane whose first slot is <= j (its area is > 0 when j < total)
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
            const uint32_t f = __shfl(afirst, lo + step);
            if (f <= j) lo += step;
        }
        const uint32_t f0 = __shfl(afirst, lo);
        const int w = __shfl(rw, lo), x0 = __shfl(rx, lo), y0 = __shfl(ry, lo);
        const uint32_t b = __shfl(base, lo), g = __shfl(gv, lo), out0 = __shfl(first, lo), n_out = __shfl(cnt, lo);
        const uint32_t cr = __shfl(carry, lo);
        const int tst = __shfl(test, lo);
        const float2 rxy = make_float2(__shfl(xy.x, lo), __shfl(xy.y, lo));
        const float4 rco = make_float4(__shfl(co.x, lo), __shfl(co.y, lo), __shfl(co.z, lo), __shfl(co.w, lo));
        const float rl = __shfl(lthr, lo);
        bool pass = false;
        int x = 0, y = 0;
        if (j < total) {
            const int q = (int)(j - f0);
            y = y0 + q / w;
            x = x0 + q % w;
            pass = !tst || cull_tile_pass(rp, rxy, rco, rl, x, y);
        }
        const uint64_t pm = __ballot(pass);
        const int lstart = (int)(f0 > jb ? f0 - jb : 0u);  // the record's first lane in this chunk (<= lane)
        const uint64_t below = ((1ull << lane) - 1ull) & ~((1ull << lstart) - 1ull);
        if (pass) {
            const uint32_t rank = cr + (uint32_t)__popcll(pm & below);
            if (rank < n_out) {
                keys[out0 + rank] = b + (uint32_t)(y * rp.gx + x);
                vals[out0 + rank] = g;
            }
        }
        // this lane's record: the passing slots of this chunk inside its range
        const int64_t s0 = (int64_t)afirst - (int64_t)jb, e0 = s0 + (int64_t)area;
        const int ls = (int)max<int64_t>(0, min<int64_t>(64, s0)), le = (int)max<int64_t>(0, min<int64_t>(64, e0));
        if (le > ls) {
            const uint64_t hi = le == 64 ? ~0ull : ((1ull << le) - 1ull);
            carry += (uint32_t)__popcll(pm & hi & ~((1ull << ls) - 1ull));
        }
    }
}

__global__ void zero_u32_kernel(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void ranges_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ n_ptr,
                              uint2* __restrict__ ranges) {
    const uint32_t n = *n_ptr;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t t = keys[i];
        if (i == 0 || keys[i - 1] != t) ranges[t].x = i;
        if (i == n - 1 || keys[i + 1] != t) ranges[t].y = i + 1;
    }
}

// One workgroup per tile (tile x tile threads, tile in {8, 16}); each thread
// owns one pixel. Each batch of the tile's Gaussians is filtered with the
// conservative tile test and compacted in LDS (order kept), so pixels only
// loop over Gaussians that can contribute to this tile.
template <int TILE>
__global__ void __launch_bounds__(TILE* TILE) blend_kernel(RasterParams rp, Geo geo, const uint32_t* __restrict__ vals,
                                                          const uint2* __restrict__ ranges) {
    constexpr int NT = TILE * TILE;
    constexpr int NW = (NT + 63) / 64;
    __shared__ float2 s_xy[NT];
    __shared__ float4 s_co[NT];
    __shared__ float4 s_rgb[NT];
    __shared__ uint32_t s_id[NT];
    __shared__ int s_cnt[NW];
    const int ntl = rp.gx * rp.gy;
    const int view = blockIdx.x / ntl;
    const int tile = blockIdx.x - view * ntl;
    float* __restrict__ out = rp.views[view].out;
    int* __restrict__ surface = rp.views[view].surface;
    const uint32_t vbase = rp.seg ? (uint32_t)rp.seg[view] : (uint32_t)view * (uint32_t)rp.P;
    const int tx = tile % rp.gx, ty = tile / rp.gx;
    const int px = tx * TILE + (threadIdx.x % TILE), py = ty * TILE + (threadIdx.x / TILE);
    const bool inside = px < rp.W && py < rp.H;
    const float fxp = (float)px, fyp = (float)py;
    const float bx0 = (float)(tx * TILE), by0 = (float)(ty * TILE);
    const float bx1 = (float)min(tx * TILE + TILE - 1, rp.W - 1), by1 = (float)min(ty * TILE + TILE - 1, rp.H - 1);
    const uint2 rg = ranges[blockIdx.x];
    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    bool done = !inside;
    const int n = (int)(rg.y - rg.x);
    const int wid = threadIdx.x >> 6, wl = threadIdx.x & 63;
    for (int b0 = 0; b0 < n; b0 += NT) {
        if (__syncthreads_count(done) == NT) break;
        const int i = b0 + (int)threadIdx.x;
        bool keep = false;
        uint32_t g = 0;
        float2 xy;
        float4 co;
        if (i < n) {
            g = vals[rg.x + i];
            xy = geo.xy[g];
            co = geo.conic[g];
            keep = tile_may_contribute(xy, co, bx0, by0, bx1, by1);
        }
        const uint64_t bal = __ballot(keep);
        if (wl == 0) s_cnt[wid] = __popcll(bal);
        __syncthreads();
        int base = 0, m = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const int cw = s_cnt[w];
            base += (w < wid) ? cw : 0;
            m += cw;
        }
        if (keep) {
            const int pos = base + __popcll(bal & ((1ull << wl) - 1ull));
            s_id[pos] = g;
            s_xy[pos] = xy;
            s_co[pos] = co;
            s_rgb[pos] = geo.rgb[g];
        }
        __syncthreads();
        for (int j = 0; j < m && !done; j++) {
            const float2 xy = s_xy[j];
            const float4 co = s_co[j];
            const float dx = xy.x - fxp, dy = xy.y - fyp;
            const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
            if (power > 0.0f) continue;
            const float alpha = fminf(0.99f, co.w * blend_exp(power));
            if (alpha < 1.0f / 255.0f) continue;
            const float test_T = T * (1.0f - alpha);
            if (test_T < 0.0001f) { done = true; break; }
            const float4 c = s_rgb[j];
            C0 += c.x * alpha * T;
            C1 += c.y * alpha * T;
            C2 += c.z * alpha * T;
            if (surface) surface[s_id[j] - vbase] = 1;
            T = test_T;
        }
    }
    if (inside) {
        const int64_t HW = (int64_t)rp.W * rp.H, o = (int64_t)py * rp.W + px;
        out[o] = C0 + T * rp.bg[0];
        out[HW + o] = C1 + T * rp.bg[1];
        out[2 * HW + o] = C2 + T * rp.bg[2];
    }
}

// 16x16 tiles: one workgroup per tile, one wave per 8x8 quadrant. Each batch
// of 256 of the tile's Gaussians is staged in LDS once; every loader thread
// tests its Gaussian against the four quadrant boxes and the per-quadrant
// survivors are compacted into index lists (depth order kept). Each wave then
// walks only its own quadrant's list, four Gaussians per step: the four
// power / exp / alpha evaluations are independent and overlap, and the
// front-to-back compositing that follows runs in order with exactly the
// per-pixel arithmetic of the one-pixel-per-thread form, so the image is the
// same bit for bit; only the work on Gaussians that cannot reach 1/255 in the
// quadrant is skipped.
// U (2, 4 or 8) is the number of Gaussians whose alphas are evaluated together
// before compositing them in order; it changes only the instruction schedule.
// Deep-tile split (SPLIT): the block is segment b of the global segment list
// (tile t holds cnt[t] = max(1, ceil(n_t / split)) segments starting at
// off[t]); blocks past the last segment exit. A tile of one segment is blended
// and written exactly as without the split. A segment of a split tile blends
// its slice of the tile's depth-ordered list from transmittance 1 and writes
// its pixels' (C, T) to part[b], with T negated when the local test stopped
// the pixel. split_composite_kernel then combines the segments front to back
// (C += T C_s, T *= T_s) up to the segment where the pixel's transmittance
// first falls below 1e-4 (the running product only decreases, so that is the
// segment where the sequential blend stops it) and blends that one segment
// again from the true incoming transmittance with the sequential test, so the
// contributors are those of the tile-by-tile blend; only the association of
// the products and sums differs.
struct SplitArgs {
    const uint32_t* cnt;
    const uint32_t* off;
    const uint32_t* total;
    float4* part;
    int nt;
};

template <int U, bool SPLIT>
__global__ void __launch_bounds__(256) blend_quad_kernel(RasterParams rp, Geo geo, const uint32_t* __restrict__ vals,
                                                         const uint2* __restrict__ ranges, SplitArgs sa) {
    constexpr int NT = 256;
    __shared__ float2 s_xy[NT];
    __shared__ float4 s_co[NT];
    __shared__ float4 s_rgb[NT];
    __shared__ uint32_t s_id[NT];
    __shared__ __align__(8) uint8_t s_list[4][NT + 8];
    __shared__ int s_cnt[4][4];  // [quadrant][wave]
    const int ntl = rp.gx * rp.gy;
    int tg = blockIdx.x;          // tile of all views
    uint2 rg;
    int64_t pslot = -1;           // this block's partial slot (segment of a split tile)
    if constexpr (SPLIT) {
        const uint32_t b = blockIdx.x;
        if (b >= *sa.total) return;  // uniform per block, before any barrier
        int lo = 0, hi = sa.nt - 1;  // the last tile whose first segment is <= b
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sa.off[mid] <= b) lo = mid;
            else hi = mid - 1;
        }
        tg = lo;
        const uint32_t sidx = b - sa.off[tg];
        const uint2 full = ranges[tg];
        rg.x = full.x + sidx * (uint32_t)rp.split;
        rg.y = min(full.y, rg.x + (uint32_t)rp.split);
        if (sa.cnt[tg] > 1) pslot = (int64_t)b;
    } else {
        rg = ranges[tg];
    }
    const int view = tg / ntl;
    const int tile = tg - view * ntl;
    float* __restrict__ out = rp.views[view].out;
    int* __restrict__ surface = rp.views[view].surface;
    const uint32_t vbase = rp.seg ? (uint32_t)rp.seg[view] : (uint32_t)view * (uint32_t)rp.P;
    const int tx = tile % rp.gx, ty = tile / rp.gx;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = tx * 16 + (wv & 1) * 8 + (lane & 7), py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
    const bool inside = px < rp.W && py < rp.H;
    const float fxp = (float)px, fyp = (float)py;
    // quadrant boxes (pixel centres inside the image)
    float qb[4][4];
    bool qok[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int x0 = tx * 16 + (q & 1) * 8, y0 = ty * 16 + (q >> 1) * 8;
        qok[q] = x0 < rp.W && y0 < rp.H;
        qb[q][0] = (float)x0;
        qb[q][1] = (float)y0;
        qb[q][2] = (float)min(x0 + 7, rp.W - 1);
        qb[q][3] = (float)min(y0 + 7, rp.H - 1);
    }
    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    bool done = !inside;
    const int n = (int)(rg.y - rg.x);
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    for (int b0 = 0; b0 < n; b0 += NT) {
        if (__syncthreads_count(done) == NT) break;
        const int i = b0 + (int)threadIdx.x;
        int qmask = 0;
        if (i < n) {
            const uint32_t g = vals[rg.x + i];
            const float2 xy = geo.xy[g];
            const float4 co = geo.conic[g];
            const float lthr = cull_log_threshold(co.w);
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (qok[q] && box_may_contribute(xy, co, lthr, qb[q][0], qb[q][1], qb[q][2], qb[q][3])) qmask |= 1 << q;
            if (qmask) {
                s_id[threadIdx.x] = g;
                s_xy[threadIdx.x] = xy;
                s_co[threadIdx.x] = co;
                s_rgb[threadIdx.x] = geo.rgb[g];
            }
        }
        uint64_t bal[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            bal[q] = __ballot((qmask >> q) & 1);
            if (lane == 0) s_cnt[q][wv] = __popcll(bal[q]);
        }
        __syncthreads();
        int m = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int base = 0;
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const int cw = s_cnt[q][w];
                base += (w < wv) ? cw : 0;
                if (q == wv) m += cw;
            }
            if ((qmask >> q) & 1) s_list[q][base + __popcll(bal[q] & lt_mask)] = (uint8_t)threadIdx.x;
        }
        __syncthreads();
        const uint8_t* lst = s_list[wv];
        for (int j = 0; j < m; j += U) {
            if (__all(done)) break;
            int k[U];
            float alpha[U];
            bool use[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                k[u] = (int)lst[j + u];  // entries past m are stale bytes: in range, and unused
                const float2 xy = s_xy[k[u]];
                const float4 co = s_co[k[u]];
                const float dx = xy.x - fxp, dy = xy.y - fyp;
                const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
                alpha[u] = fminf(0.99f, co.w * blend_exp(power));
                use[u] = (j + u < m) && !(power > 0.0f) && !(alpha[u] < 1.0f / 255.0f);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (use[u] && !done) {
                    const float test_T = T * (1.0f - alpha[u]);
                    if (test_T < 0.0001f) {
                        done = true;
                    } else {
                        const float4 c = s_rgb[k[u]];
                        C0 += c.x * alpha[u] * T;
                        C1 += c.y * alpha[u] * T;
                        C2 += c.z * alpha[u] * T;
                        if (surface) surface[s_id[k[u]] - vbase] = 1;
                        T = test_T;
                    }
                }
            }
        }
    }
    if (SPLIT && pslot >= 0) {
        sa.part[pslot * NT + threadIdx.x] = make_float4(C0, C1, C2, (done && inside) ? -T : T);
    } else if (inside) {
        const int64_t HW = (int64_t)rp.W * rp.H, o = (int64_t)py * rp.W + px;
        out[o] = C0 + T * rp.bg[0];
        out[HW + o] = C1 + T * rp.bg[1];
        out[2 * HW + o] = C2 + T * rp.bg[2];
    }
}

__global__ void split_count_kernel(int64_t nt, const uint2* __restrict__ ranges, int split, uint32_t* __restrict__ cnt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint32_t n = ranges[t].y - ranges[t].x;
    cnt[t] = n <= (uint32_t)split ? 1u : (n + (uint32_t)split - 1u) / (uint32_t)split;
}

// Front-to-back composite of a split tile's segments, per pixel (the pixel
// mapping of blend_quad_kernel), with the stopping segment blended again.
__global__ void __launch_bounds__(256) split_composite_kernel(RasterParams rp, Geo geo,
                                                              const uint32_t* __restrict__ vals,
                                                              const uint2* __restrict__ ranges, SplitArgs sa) {
    constexpr int NT = 256;
    __shared__ float2 s_xy[NT];
    __shared__ float4 s_co[NT];
    __shared__ float4 s_rgb[NT];
    const int tg = blockIdx.x;
    const uint32_t ns = sa.cnt[tg];
    if (ns <= 1) return;  // uniform per block
    const int ntl = rp.gx * rp.gy;
    const int view = tg / ntl;
    const int tile = tg - view * ntl;
    const int tx = tile % rp.gx, ty = tile / rp.gx;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = tx * 16 + (wv & 1) * 8 + (lane & 7), py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
    const bool inside = px < rp.W && py < rp.H;
    const float fxp = (float)px, fyp = (float)py;
    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    int stop = -1;  // the segment where this pixel's blend stops
    const float4* p = sa.part + (int64_t)sa.off[tg] * NT + threadIdx.x;
    if (inside) {
        for (uint32_t s = 0; s < ns; s++) {
            const float4 v = p[(int64_t)s * NT];
            if (v.w < 0.0f || T * v.w < 0.0001f) {
                stop = (int)s;
                break;
            }
            C0 += T * v.x;
            C1 += T * v.y;
            C2 += T * v.z;
            T *= v.w;
        }
    }
    const uint2 full = ranges[tg];
    for (uint32_t s = 0; s < ns; s++) {
        if (!__syncthreads_or(stop == (int)s)) continue;
        const uint32_t lo = full.x + s * (uint32_t)rp.split, hi = min(full.y, lo + (uint32_t)rp.split);
        bool done = stop != (int)s;
        for (uint32_t b0 = lo; b0 < hi; b0 += NT) {
            if (__syncthreads_count(done) == NT) break;
            const uint32_t i = b0 + threadIdx.x;
            if (i < hi) {
                const uint32_t g = vals[i];
                s_xy[threadIdx.x] = geo.xy[g];
                s_co[threadIdx.x] = geo.conic[g];
                s_rgb[threadIdx.x] = geo.rgb[g];
            }
            __syncthreads();
            const int m = (int)min((uint32_t)NT, hi - b0);
            for (int j = 0; j < m && !done; j++) {
                const float2 xy = s_xy[j];
                const float4 co = s_co[j];
                const float dx = xy.x - fxp, dy = xy.y - fyp;
                const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
                if (power > 0.0f) continue;
                const float alpha = fminf(0.99f, co.w * blend_exp(power));
                if (alpha < 1.0f / 255.0f) continue;
                const float test_T = T * (1.0f - alpha);
                if (test_T < 0.0001f) {
                    done = true;
                    break;
                }
                const float4 c = s_rgb[j];
                C0 += c.x * alpha * T;
                C1 += c.y * alpha * T;
                C2 += c.z * alpha * T;
                T = test_T;
            }
            __syncthreads();
        }
    }
    if (!inside) return;
    float* __restrict__ out = rp.views[view].out;
    const int64_t HW = (int64_t)rp.W * rp.H, o = (int64_t)py * rp.W + px;
    out[o] = C0 + T * rp.bg[0];
    out[HW + o] = C1 + T * rp.bg[1];
    out[2 * HW + o] = C2 + T * rp.bg[2];
}

// ----------------------------------------------------------------- backward
// Gradient of one pixel's composite C = sum_i c_i a_i T_i + T_final bg with
// respect to the Gaussians it composited, walking them back to front. With
// S_i = sum_{j>i} c_j a_j T_j + T_final bg (the colour behind i):
//   dL/dc_i = a_i T_i dL/dC,   dL/da_i = sum_ch dL/dC_ch (c_i,ch T_i - S_i,ch / (1 - a_i)),
// T_i = T_{i+1} / (1 - a_i). a_i = min(0.99, o_i G_i) with G_i = exp(power_i)
// and power_i = -(A dx^2 + C dy^2)/2 - B dx dy (d = mean - pixel): the clamp
// has zero slope. Opacities are not differentiated.
//
// One workgroup per tile, one thread per pixel. The tile first re-runs the
// forward pass (same float operations, so the same contributors, final
// transmittance and stopping point per pixel) to find each pixel's last
// contributor; the reverse pass then restages the tile's batches in LDS from
// the deepest one up. For every Gaussian each wave sums its 64 pixels'
// gradients with lane shuffles into a per-wave LDS slot; after the batch one
// thread per Gaussian adds the wave slots and issues one set of 8 global
// atomic adds per (tile, Gaussian).
template <int TILE>
__global__ void __launch_bounds__(TILE* TILE) blend_backward_kernel(RasterParams rp, Geo geo,
                                                                   const uint32_t* __restrict__ vals,
                                                                   const uint2* __restrict__ ranges,
                                                                   float* __restrict__ grec) {
    constexpr int NT = TILE * TILE;
    constexpr int NW = NT / 64;
    static_assert(NT % 64 == 0, "whole waves per tile");
    __shared__ float2 s_xy[NT];
    __shared__ float4 s_co[NT];
    __shared__ float4 s_rgb[NT];
    __shared__ uint32_t s_id[NT];
    __shared__ float s_g[NW][NT][8];
    __shared__ int s_last;
    const int ntl = rp.gx * rp.gy;
    const int view = blockIdx.x / ntl;
    const int tile = blockIdx.x - view * ntl;
    const float* __restrict__ dout = rp.views[view].dout;
    const int tx = tile % rp.gx, ty = tile / rp.gx;
    const int px = tx * TILE + (threadIdx.x % TILE), py = ty * TILE + (threadIdx.x / TILE);
    const bool inside = px < rp.W && py < rp.H;
    const float fxp = (float)px, fyp = (float)py;
    const uint2 rg = ranges[blockIdx.x];
    const int n = (int)(rg.y - rg.x);
    const int wid = threadIdx.x >> 6, wl = threadIdx.x & 63;
    if (threadIdx.x == 0) s_last = -1;

    // forward pass: final transmittance and last contributor of this pixel
    float T = 1.0f;
    int last = -1;
    bool done = !inside;
    for (int b0 = 0; b0 < n; b0 += NT) {
        if (__syncthreads_count(done) == NT) break;
        const int i = b0 + (int)threadIdx.x;
        if (i < n) {
            const uint32_t g = vals[rg.x + i];
            s_xy[threadIdx.x] = geo.xy[g];
            s_co[threadIdx.x] = geo.conic[g];
        }
        __syncthreads();
        const int m = min(NT, n - b0);
        for (int j = 0; j < m && !done; j++) {
            const float2 xy = s_xy[j];
            const float4 co = s_co[j];
            const float dx = xy.x - fxp, dy = xy.y - fyp;
            const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
            if (power > 0.0f) continue;
            const float alpha = fminf(0.99f, co.w * blend_exp(power));
            if (alpha < 1.0f / 255.0f) continue;
            const float test_T = T * (1.0f - alpha);
            if (test_T < 0.0001f) { done = true; break; }
            T = test_T;
            last = b0 + j;
        }
    }
    __syncthreads();
    if (last >= 0) atomicMax(&s_last, last);
    __syncthreads();
    const int tile_last = s_last;
    if (tile_last < 0) return;  // uniform: nothing in this tile contributed

    float dC[3] = {0.0f, 0.0f, 0.0f};
    if (inside && dout) {
        const int64_t HW = (int64_t)rp.W * rp.H, o = (int64_t)py * rp.W + px;
        dC[0] = dout[o]; dC[1] = dout[HW + o]; dC[2] = dout[2 * HW + o];
    }
    float S0 = T * rp.bg[0], S1 = T * rp.bg[1], S2 = T * rp.bg[2];
    float Tc = T;
    for (int b0 = (tile_last / NT) * NT; b0 >= 0; b0 -= NT) {
        __syncthreads();
        const int i = b0 + (int)threadIdx.x;
        if (i < n) {
            const uint32_t g = vals[rg.x + i];
            s_id[threadIdx.x] = g;
            s_xy[threadIdx.x] = geo.xy[g];
            s_co[threadIdx.x] = geo.conic[g];
            s_rgb[threadIdx.x] = geo.rgb[g];
        }
        __syncthreads();
        const int jtop = min(NT - 1, tile_last - b0);
        for (int j = jtop; j >= 0; j--) {
            float gv[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            bool act = false;
            if (b0 + j <= last) {
                const float2 xy = s_xy[j];
                const float4 co = s_co[j];
                const float dx = xy.x - fxp, dy = xy.y - fyp;
                const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
                if (power <= 0.0f) {
                    const float G = blend_exp(power);
                    const float alpha = fminf(0.99f, co.w * G);
                    if (alpha >= 1.0f / 255.0f) {
                        act = true;
                        const float4 c = s_rgb[j];
                        const float om = 1.0f - alpha;
                        const float Ti = Tc / om;
                        const float wgt = alpha * Ti;
                        gv[5] = wgt * dC[0];
                        gv[6] = wgt * dC[1];
                        gv[7] = wgt * dC[2];
                        const float da = dC[0] * (c.x * Ti - S0 / om) + dC[1] * (c.y * Ti - S1 / om) +
                                         dC[2] * (c.z * Ti - S2 / om);
                        S0 += c.x * wgt;
                        S1 += c.y * wgt;
                        S2 += c.z * wgt;
                        Tc = Ti;
                        if (co.w * G < 0.99f) {
                            const float dpow = da * co.w * G;
                            gv[0] = dpow * (-co.x * dx - co.y * dy);
                            gv[1] = dpow * (-co.z * dy - co.y * dx);
                            gv[2] = -0.5f * dx * dx * dpow;
                            gv[3] = -dx * dy * dpow;
                            gv[4] = -0.5f * dy * dy * dpow;
                        }
                    }
                }
            }
            if (__ballot(act)) {
#pragma unroll
                for (int k = 0; k < 8; k++) gv[k] = wave_sum(gv[k]);
            }
            if (wl == 0) {
#pragma unroll
                for (int k = 0; k < 8; k++) s_g[wid][j][k] = gv[k];
            }
        }
        __syncthreads();
        const int t = (int)threadIdx.x;
        if (t <= jtop) {
            float acc[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                float a = 0.0f;
#pragma unroll
                for (int w = 0; w < NW; w++) a += s_g[w][t][k];
                acc[k] = a;
            }
            // component-major: component k of record gv at k * V * P + gv
            const int64_t PV = (int64_t)rp.V * rp.P;
            float* dst = grec + s_id[t];
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (acc[k] != 0.0f) unsafeAtomicAdd(dst + k * PV, acc[k]);
        }
    }
}

// Backward of project_one for one (view, Gaussian) record whose blend
// gradient is gr = d(x, y, conic a, b, c, r, g, b). The forward quantities
// are recomputed with the forward's float operations; the clamps of the
// forward (view-space tangent limits, colour >= 0) pass no gradient where
// they are active.
__device__ __forceinline__ void project_backward(const RasterParams& rp, const RasterView& vw, int g,
                                                 const float* __restrict__ sh, const float gr[8], float dp[3],
                                                 float dq[4], float ds[3], float dsh[48]) {
    const float px = rp.means[3 * g], py = rp.means[3 * g + 1], pz = rp.means[3 * g + 2];
    const float* V = vw.view;
    const float* Pm = vw.proj;
    const float vx = V[0] * px + V[4] * py + V[8] * pz + V[12];
    const float vy = V[1] * px + V[5] * py + V[9] * pz + V[13];
    const float vz = V[2] * px + V[6] * py + V[10] * pz + V[14];
    const float hx = Pm[0] * px + Pm[4] * py + Pm[8] * pz + Pm[12];
    const float hy = Pm[1] * px + Pm[5] * py + Pm[9] * pz + Pm[13];
    const float hw = Pm[3] * px + Pm[7] * py + Pm[11] * pz + Pm[15];
    const float pw = 1.0f / (hw + 0.0000001f);
    const float q[4] = {rp.rots[4 * g], rp.rots[4 * g + 1], rp.rots[4 * g + 2], rp.rots[4 * g + 3]};
    const float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    float R[3][3];
    R[0][0] = 1.0f - 2.0f * (qy * qy + qz * qz);
    R[0][1] = 2.0f * (qx * qy - qw * qz);
    R[0][2] = 2.0f * (qx * qz + qw * qy);
    R[1][0] = 2.0f * (qx * qy + qw * qz);
    R[1][1] = 1.0f - 2.0f * (qx * qx + qz * qz);
    R[1][2] = 2.0f * (qy * qz - qw * qx);
    R[2][0] = 2.0f * (qx * qz - qw * qy);
    R[2][1] = 2.0f * (qy * qz + qw * qx);
    R[2][2] = 1.0f - 2.0f * (qx * qx + qy * qy);
    float sc[3], s2[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        sc[k] = rp.scale_mod * rp.scales[3 * g + k];
        s2[k] = sc[k] * sc[k];
    }
    float S[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            S[i][j] = R[i][0] * s2[0] * R[j][0] + R[i][1] * s2[1] * R[j][1] + R[i][2] * s2[2] * R[j][2];
    const float limx = 1.3f * vw.tanfovx, limy = 1.3f * vw.tanfovy;
    const float txtz = vx / vz, tytz = vy / vz;
    const float tx = fminf(limx, fmaxf(-limx, txtz)) * vz;
    const float ty = fminf(limy, fmaxf(-limy, tytz)) * vz;
    const float fx = vw.focal_x, fy = vw.focal_y;
    const float j00 = fx / vz, j02 = -(fx * tx) / (vz * vz);
    const float j11 = fy / vz, j12 = -(fy * ty) / (vz * vz);
    float T[2][3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        T[0][c] = j00 * V[c * 4 + 0] + j02 * V[c * 4 + 2];
        T[1][c] = j11 * V[c * 4 + 1] + j12 * V[c * 4 + 2];
    }
    float ST0[3], ST1[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        ST0[i] = S[i][0] * T[0][0] + S[i][1] * T[0][1] + S[i][2] * T[0][2];
        ST1[i] = S[i][0] * T[1][0] + S[i][1] * T[1][1] + S[i][2] * T[1][2];
    }
    const float ca = T[0][0] * ST0[0] + T[0][1] * ST0[1] + T[0][2] * ST0[2] + 0.3f;
    const float cb = T[0][0] * ST1[0] + T[0][1] * ST1[1] + T[0][2] * ST1[2];
    const float cc = T[1][0] * ST1[0] + T[1][1] * ST1[1] + T[1][2] * ST1[2] + 0.3f;
    const float det = ca * cc - cb * cb;

    // colour
    const float ux0 = px - vw.campos[0], uy0 = py - vw.campos[1], uz0 = pz - vw.campos[2];
    const float dn = __builtin_sqrtf(ux0 * ux0 + uy0 * uy0 + uz0 * uz0);
    const float ux = ux0 / dn, uy = uy0 / dn, uz = uz0 / dn;
    float b[16], db[16][3];
    sh_basis_and_grad(ux, uy, uz, b, db);
    const int nb = (rp.D + 1) * (rp.D + 1);
    float dcol[3];
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        float r = 0.0f;
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (k < nb) r += b[k] * sh[3 * k + ch];
        dcol[ch] = (r + 0.5f > 0.0f) ? gr[5 + ch] : 0.0f;
    }
    float ddir[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k < nb) {
            float w = 0.0f;
#pragma unroll
            for (int ch = 0; ch < 3; ch++) {
                dsh[3 * k + ch] += b[k] * dcol[ch];
                w += sh[3 * k + ch] * dcol[ch];
            }
            ddir[0] += w * db[k][0];
            ddir[1] += w * db[k][1];
            ddir[2] += w * db[k][2];
        }
    }
    const float dd = ux * ddir[0] + uy * ddir[1] + uz * ddir[2];
    dp[0] += (ddir[0] - ux * dd) / dn;
    dp[1] += (ddir[1] - uy * dd) / dn;
    dp[2] += (ddir[2] - uz * dd) / dn;

    // conic (a, b, c) = (cc, -cb, ca) / det
    const float gA = gr[2], gB = gr[3], gC = gr[4];
    const float id = 1.0f / det, id2 = id * id;
    const float dca = -gA * cc * cc * id2 + gB * cb * cc * id2 + gC * (id - ca * cc * id2);
    const float dcb = 2.0f * gA * cb * cc * id2 - gB * (id + 2.0f * cb * cb * id2) + 2.0f * gC * cb * ca * id2;
    const float dcc = gA * (id - cc * ca * id2) + gB * cb * ca * id2 - gC * ca * ca * id2;
    // cov2D = T Sigma T^T
    float dS[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            dS[i][j] = dca * T[0][i] * T[0][j] + dcb * T[0][i] * T[1][j] + dcc * T[1][i] * T[1][j];
    float dj00 = 0.0f, dj02 = 0.0f, dj11 = 0.0f, dj12 = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float dT0 = 2.0f * dca * ST0[c] + dcb * ST1[c];
        const float dT1 = 2.0f * dcc * ST1[c] + dcb * ST0[c];
        dj00 += dT0 * V[c * 4 + 0];
        dj02 += dT0 * V[c * 4 + 2];
        dj11 += dT1 * V[c * 4 + 1];
        dj12 += dT1 * V[c * 4 + 2];
    }
    const float iz = 1.0f / vz, iz2 = iz * iz, iz3 = iz2 * iz;
    float dvx = 0.0f, dvy = 0.0f;
    float dvz = -dj00 * fx * iz2 - dj11 * fy * iz2 + 2.0f * dj02 * fx * tx * iz3 + 2.0f * dj12 * fy * ty * iz3;
    const float dtx = -dj02 * fx * iz2, dty = -dj12 * fy * iz2;
    if (txtz < -limx) dvz -= dtx * limx;
    else if (txtz > limx) dvz += dtx * limx;
    else dvx += dtx;
    if (tytz < -limy) dvz -= dty * limy;
    else if (tytz > limy) dvz += dty * limy;
    else dvy += dty;
    // Sigma = R diag(s^2) R^T
    float dR[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int bb = 0; bb < 3; bb++) {
            float acc = 0.0f;
#pragma unroll
            for (int j = 0; j < 3; j++) acc += (dS[a][j] + dS[j][a]) * R[j][bb];
            dR[a][bb] = s2[bb] * acc;
        }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float dD = 0.0f;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dD += dS[i][j] * R[i][k] * R[j][k];
        ds[k] += dD * 2.0f * sc[k] * rp.scale_mod;
    }
    quat_to_mat_backward(q, dR, dq);
    // image position ((ndc + 1) * size - 1) / 2 through the projection
    const float dndx = gr[0] * 0.5f * (float)rp.W, dndy = gr[1] * 0.5f * (float)rp.H;
    const float dhx = dndx * pw, dhy = dndy * pw;
    const float dpw = dndx * hx + dndy * hy;
    const float dhw = -dpw * pw * pw;
    dp[0] += dhx * Pm[0] + dhy * Pm[1] + dhw * Pm[3] + dvx * V[0] + dvy * V[1] + dvz * V[2];
    dp[1] += dhx * Pm[4] + dhy * Pm[5] + dhw * Pm[7] + dvx * V[4] + dvy * V[5] + dvz * V[6];
    dp[2] += dhx * Pm[8] + dhy * Pm[9] + dhw * Pm[11] + dvx * V[8] + dvy * V[9] + dvz * V[10];
}

// One thread per Gaussian over all views (its SH read once).
__global__ void __launch_bounds__(256) preprocess_backward_kernel(RasterParams rp, Geo geo,
                                                                  const float* __restrict__ grec,
                                                                  float* __restrict__ gout) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= rp.P) return;
    float dp[3] = {0.0f, 0.0f, 0.0f}, dq[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ds[3] = {0.0f, 0.0f, 0.0f};
    float dsh[48];
#pragma unroll
    for (int k = 0; k < 48; k++) dsh[k] = 0.0f;
    float sh[48];
    const float4* s4 = reinterpret_cast<const float4*>(rp.shs + 48 * (int64_t)g);
#pragma unroll
    for (int i = 0; i < 12; i++) {
        const float4 v = s4[i];
        sh[4 * i] = v.x; sh[4 * i + 1] = v.y; sh[4 * i + 2] = v.z; sh[4 * i + 3] = v.w;
    }
    for (int v = 0; v < rp.V; v++) {
        const int64_t gv = (int64_t)v * rp.P + g;
        if (geo.touched[gv] == 0) continue;
        float gr[8];
        bool any = false;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            gr[k] = grec[k * (int64_t)rp.V * rp.P + gv];
            any = any || gr[k] != 0.0f;
        }
        if (!any) continue;
        const RasterView vw = rp.views[v];
        project_backward(rp, vw, g, sh, gr, dp, dq, ds, dsh);
    }
    float* o = gout + 58 * (int64_t)g;
#pragma unroll
    for (int k = 0; k < 3; k++) o[k] = dp[k];
#pragma unroll
    for (int k = 0; k < 4; k++) o[3 + k] = dq[k];
#pragma unroll
    for (int k = 0; k < 3; k++) o[7 + k] = ds[k];
#pragma unroll
    for (int k = 0; k < 48; k++) o[10 + k] = dsh[k];
}

__global__ void copy_count_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) { *dst = *src; }

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

size_t raster_geo_bytes(int64_t PV) {
    // depth, xy, conic, rgb, rect, touched, offsets (+ alignment slack)
    return (size_t)PV * (4 + 8 + 16 + 16 + 16 + 4 + 4) + 7 * 256;
}

static Geo carve_geo(void* base, int64_t PV) {
    char* p = (char*)base;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += (bytes + 255) & ~(size_t)255;
        return r;
    };
    Geo g;
    g.depth = (float*)take(4 * (size_t)PV);
    g.xy = (float2*)take(8 * (size_t)PV);
    g.conic = (float4*)take(16 * (size_t)PV);
    g.rgb = (float4*)take(16 * (size_t)PV);
    g.rect = (int4*)take(16 * (size_t)PV);
    g.touched = (uint32_t*)take(4 * (size_t)PV);
    return g;
}

static uint32_t* geo_offsets(void* base, int64_t PV) {
    char* p = (char*)base;
    size_t off = 0;
    const size_t sz[6] = {4, 8, 16, 16, 16, 4};
    for (int i = 0; i < 6; i++) off += (sz[i] * (size_t)PV + 255) & ~(size_t)255;
    return (uint32_t*)(p + off);
}

int raster_preprocess(const RasterParams& rp, void* geo_ws, uint32_t* scan_ws, uint32_t* d_counts, hipStream_t st) {
    const int64_t PV = raster_records(rp);
    Geo geo = carve_geo(geo_ws, PV);
    hipLaunchKernelGGL(depth_range_init_kernel, dim3(1), dim3(1), 0, st, d_counts + 1);
    if (rp.P > 0) {
        hipLaunchKernelGGL(preprocess_kernel, dim3(nblk(rp.P, 256)), dim3(256), 0, st, rp, geo, d_counts + 1);
        HIP_RET(hipGetLastError());
    }
    return launch_exclusive_scan(geo.touched, PV, geo_offsets(geo_ws, PV), d_counts, scan_ws, st);
}

// split workspace: cnt u32[nt], off u32[nt], total u32[64], scan workspace,
// then the partials (float4 per pixel of every segment), each 256-byte aligned
static size_t split_align(size_t b) { return (b + 255) & ~(size_t)255; }

size_t raster_split_bytes(int64_t nt, uint32_t n_rendered, int split) {
    if (split <= 0) return 0;
    const size_t segs = (size_t)nt + n_rendered / (uint32_t)split + 1;
    return 2 * split_align(4 * (size_t)nt) + 256 + split_align(4 * scan_workspace_words(nt)) + segs * 256 * 16;
}

static SplitArgs split_carve(void* ws, int64_t nt, uint32_t n_rendered, int split) {
    (void)n_rendered; (void)split;
    char* p = (char*)ws;
    SplitArgs a;
    a.cnt = (const uint32_t*)p;
    p += split_align(4 * (size_t)nt);
    a.off = (const uint32_t*)p;
    p += split_align(4 * (size_t)nt);
    a.total = (const uint32_t*)p;
    p += 256 + split_align(4 * scan_workspace_words(nt));
    a.part = (float4*)p;
    a.nt = (int)nt;
    return a;
}

static uint32_t* split_scan_ws(void* ws, int64_t nt) {
    return (uint32_t*)((char*)ws + 2 * split_align(4 * (size_t)nt) + 256);
}

size_t raster_record_bytes(int64_t PV) {
    return (size_t)PV * (4 + 4 + 4 + 4 + 4) + sizeof(uint32_t) * radix_workspace_words(PV) + 6 * 256;
}

int raster_bin_and_blend(const RasterParams& rp, void* geo_ws, uint32_t n_rendered, const uint32_t* d_total,
                         uint32_t dmin, int dbits, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                         uint32_t* vals_alt, uint32_t* sort_ws, uint2* ranges, void* rec_ws, uint32_t* scan_ws,
                         void* split_ws, hipStream_t st) {
    const int64_t PV = raster_records(rp);
    Geo geo = carve_geo(geo_ws, PV);
    const int ntiles = rp.gx * rp.gy;
    const int64_t nt = (int64_t)ntiles * rp.V;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(nblk(2 * nt, 256)), dim3(256), 0, st, (uint32_t*)ranges, 2 * nt);
    if (n_rendered > 0) {
        // records by depth
        char* p = (char*)rec_ws;
        auto take = [&](size_t bytes) {
            char* r = p;
            p += (bytes + 255) & ~(size_t)255;
            return r;
        };
        uint32_t* rkeys = (uint32_t*)take(4 * (size_t)PV);
        uint32_t* rkeys_alt = (uint32_t*)take(4 * (size_t)PV);
        uint32_t* rvals = (uint32_t*)take(4 * (size_t)PV);
        uint32_t* rvals_alt = (uint32_t*)take(4 * (size_t)PV);
        uint32_t* rtouched = (uint32_t*)take(4 * (size_t)PV);
        uint32_t* rsort_ws = (uint32_t*)take(sizeof(uint32_t) * radix_workspace_words(PV));
        hipLaunchKernelGGL(record_keys_kernel, dim3(nblk(PV, 256)), dim3(256), 0, st, PV, geo, dmin, rkeys, rvals);
        int rc = launch_radix_sort32(rkeys, rvals, rkeys_alt, rvals_alt, PV, dbits, rsort_ws, st);
        if (rc) return rc;
        // instance offsets in depth order, then the tiles of every record
        hipLaunchKernelGGL(gather_touched_kernel, dim3(nblk(PV, 256)), dim3(256), 0, st, PV, rvals, geo.touched,
                           rtouched);
        uint32_t* offs = geo_offsets(geo_ws, PV);
        rc = launch_exclusive_scan(rtouched, PV, offs, (uint32_t*)d_total, scan_ws, st);
        if (rc) return rc;
        if (rp.cull)
            hipLaunchKernelGGL(emit_tiles_cull_kernel, dim3(nblk(PV, 256)), dim3(256), 0, st, rp, PV, ntiles, geo, rvals,
                               offs, keys, vals);
        else
            hipLaunchKernelGGL(emit_tiles_kernel, dim3(nblk(PV, 256)), dim3(256), 0, st, PV, rp.P, rp.seg, rp.V, rp.gx,
                               ntiles, geo, rvals, offs, keys, vals);
        int tile_bits = 0;
        while ((int64_t(1) << tile_bits) < nt) tile_bits++;
        rc = launch_radix_sort32(keys, vals, keys_alt, vals_alt, n_rendered, tile_bits, sort_ws, st);
        if (rc) return rc;
        hipLaunchKernelGGL(ranges_kernel, dim3(std::min<unsigned>(nblk(n_rendered, 256), 4096)), dim3(256), 0, st,
                           keys, d_total, ranges);
    }
    if (rp.tile == 16 && rp.split > 0 && n_rendered > 0 && split_ws) {
        SplitArgs sa = split_carve(split_ws, nt, n_rendered, rp.split);
        hipLaunchKernelGGL(split_count_kernel, dim3(nblk(nt, 256)), dim3(256), 0, st, nt, ranges, rp.split,
                           const_cast<uint32_t*>(sa.cnt));
        int rc = launch_exclusive_scan(sa.cnt, nt, const_cast<uint32_t*>(sa.off), const_cast<uint32_t*>(sa.total),
                                       split_scan_ws(split_ws, nt), st);
        if (rc) return rc;
        const unsigned nb = (unsigned)(nt + n_rendered / (uint32_t)rp.split + 1);  // at least the segment count
        if (rp.unroll == 2)
            hipLaunchKernelGGL((blend_quad_kernel<2, true>), dim3(nb), dim3(256), 0, st, rp, geo, vals, ranges, sa);
        else if (rp.unroll == 8)
            hipLaunchKernelGGL((blend_quad_kernel<8, true>), dim3(nb), dim3(256), 0, st, rp, geo, vals, ranges, sa);
        else
            hipLaunchKernelGGL((blend_quad_kernel<4, true>), dim3(nb), dim3(256), 0, st, rp, geo, vals, ranges, sa);
        hipLaunchKernelGGL(split_composite_kernel, dim3((unsigned)nt), dim3(256), 0, st, rp, geo, vals, ranges, sa);
    } else if (rp.tile == 16) {
        const SplitArgs none{nullptr, nullptr, nullptr, nullptr, 0};
        if (rp.unroll == 2)
            hipLaunchKernelGGL((blend_quad_kernel<2, false>), dim3((unsigned)nt), dim3(256), 0, st, rp, geo, vals,
                               ranges, none);
        else if (rp.unroll == 8)
            hipLaunchKernelGGL((blend_quad_kernel<8, false>), dim3((unsigned)nt), dim3(256), 0, st, rp, geo, vals,
                               ranges, none);
        else
            hipLaunchKernelGGL((blend_quad_kernel<4, false>), dim3((unsigned)nt), dim3(256), 0, st, rp, geo, vals,
                               ranges, none);
    } else if (rp.tile == 8)
        hipLaunchKernelGGL(blend_kernel<8>, dim3((unsigned)nt), dim3(64), 0, st, rp, geo, vals, ranges);
    else
        return -1;
    HIP_RET(hipGetLastError());
    return 0;
}

int raster_blend_backward(const RasterParams& rp, void* geo_ws, const uint32_t* vals, const uint2* ranges,
                          float* grec, hipStream_t st) {
    if (rp.seg) return -1;  // segment-mode renders have no backward
    const int64_t PV = (int64_t)rp.P * rp.V;
    Geo geo = carve_geo(geo_ws, PV);
    const int64_t nt = (int64_t)rp.gx * rp.gy * rp.V;
    if (nt == 0) return 0;
    if (rp.tile == 16)
        hipLaunchKernelGGL(blend_backward_kernel<16>, dim3((unsigned)nt), dim3(256), 0, st, rp, geo, vals, ranges,
                           grec);
    else if (rp.tile == 8)
        hipLaunchKernelGGL(blend_backward_kernel<8>, dim3((unsigned)nt), dim3(64), 0, st, rp, geo, vals, ranges,
                           grec);
    else
        return -1;
    HIP_RET(hipGetLastError());
    return 0;
}

int raster_preprocess_backward(const RasterParams& rp, void* geo_ws, const float* grec, float* gout,
                               hipStream_t st) {
    if (rp.seg) return -1;
    if (rp.P <= 0) return 0;
    const int64_t PV = (int64_t)rp.P * rp.V;
    Geo geo = carve_geo(geo_ws, PV);
    hipLaunchKernelGGL(preprocess_backward_kernel, dim3(nblk(rp.P, 256)), dim3(256), 0, st, rp, geo, grec, gout);
    HIP_RET(hipGetLastError());
    return 0;
}

}  // namespace arapgs
