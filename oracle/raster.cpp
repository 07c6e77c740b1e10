// TEST INFRASTRUCTURE — the oracle (see common.hpp). A CPU 3D Gaussian
// Splatting forward rasteriser written from the 3DGS paper (Kerbl et al.,
// SIGGRAPH 2023): EWA projection, tile binning, depth sort and front-to-back
// alpha blending. The reference calls an un-vendored CUDA fork of it
// (CudaRasterizer::Rasterizer::forward, call at GaussianView.cpp:1099-1128);
// only its arguments are known, so parity with the reference is unpinned and
// this restatement fixes the arithmetic the HIP rasteriser must reproduce.
//
// Every preprocess quantity that feeds a sort key or a tile rectangle is
// evaluated with single float operations in a fixed order (no contraction), so
// the HIP preprocess can match it bit for bit.
#include "common.hpp"
#include "raster.hpp"

#include <algorithm>
#include <vector>

namespace {

const float SH_C0 = 0.28209479177387814f;
const float SH_C1 = 0.4886025119029199f;
const float SH_C2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f, -1.0925484305920792f,
                        0.5462742152960396f};
const float SH_C3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f, 0.3731763325901154f,
                        -0.4570457994644658f, 1.445305721320277f, -0.5900435899266435f};

}  // namespace

extern "C" {


// Per-Gaussian preprocess. Outputs (P each): depth, radius, xy (2), conic+opacity
// (4), rgb (3), rect (4: minx,miny,maxx,maxy), tiles touched.
void orc_raster_preprocess(const orc_raster_args* a, float* depth, int* radius, float* xy, float* conic, float* rgb,
                           int* rect, int* touched) {
    const int gx = (a->W + a->tile - 1) / a->tile, gy = (a->H + a->tile - 1) / a->tile;
    const float* V = a->view;
    const float* Pm = a->proj;
    // perspective: focal lengths in pixels; orthographic: pixels per view unit
    const float fx = a->ortho ? (float)a->W / (2.0f * a->ortho_scale) : (float)a->W / (2.0f * a->tanfovx);
    const float fy = a->ortho ? (float)a->H / (2.0f * a->ortho_scale) : (float)a->H / (2.0f * a->tanfovy);
#pragma omp parallel for schedule(static)
    for (int g = 0; g < a->P; g++) {
        radius[g] = 0;
        touched[g] = 0;
        const float px = a->means[3 * g], py = a->means[3 * g + 1], pz = a->means[3 * g + 2];
        if (a->boxmin && (px < a->boxmin[0] || py < a->boxmin[1] || pz < a->boxmin[2] || px > a->boxmax[0] ||
                          py > a->boxmax[1] || pz > a->boxmax[2]))
            continue;
        const float vx = V[0] * px + V[4] * py + V[8] * pz + V[12];
        const float vy = V[1] * px + V[5] * py + V[9] * pz + V[13];
        const float vz = V[2] * px + V[6] * py + V[10] * pz + V[14];
        if (vz <= 0.2f) continue;
        const float hx = Pm[0] * px + Pm[4] * py + Pm[8] * pz + Pm[12];
        const float hy = Pm[1] * px + Pm[5] * py + Pm[9] * pz + Pm[13];
        const float hw = Pm[3] * px + Pm[7] * py + Pm[11] * pz + Pm[15];
        const float pw = 1.0f / (hw + 0.0000001f);
        const float ndx = hx * pw, ndy = hy * pw;
        // 3D covariance R diag(s^2) R^T
        const float qw = a->rots[4 * g], qx = a->rots[4 * g + 1], qy = a->rots[4 * g + 2], qz = a->rots[4 * g + 3];
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
        for (int k = 0; k < 3; k++) {
            const float s = a->scale_mod * a->scales[3 * g + k];
            s2[k] = s * s;
        }
        float S[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                S[i][j] = R[i][0] * s2[0] * R[j][0] + R[i][1] * s2[1] * R[j][1] + R[i][2] * s2[2] * R[j][2];
        // EWA: J W Sigma W^T J^T; J is the projection's Jacobian at the mean:
        // perspective (x/z, y/z with the 1.3 tan(fov) clamp) or orthographic
        // (constant scale, no dependence on depth)
        float j00, j02, j11, j12;
        if (a->ortho) {
            j00 = fx; j02 = 0.0f;
            j11 = fy; j12 = 0.0f;
        } else {
            const float limx = 1.3f * a->tanfovx, limy = 1.3f * a->tanfovy;
            const float txtz = vx / vz, tytz = vy / vz;
            const float tx = std::min(limx, std::max(-limx, txtz)) * vz;
            const float ty = std::min(limy, std::max(-limy, tytz)) * vz;
            j00 = fx / vz; j02 = -(fx * tx) / (vz * vz);
            j11 = fy / vz; j12 = -(fy * ty) / (vz * vz);
        }
        float T[2][3];
        for (int c = 0; c < 3; c++) {
            // W(r,c) = V[c*4 + r]
            T[0][c] = j00 * V[c * 4 + 0] + j02 * V[c * 4 + 2];
            T[1][c] = j11 * V[c * 4 + 1] + j12 * V[c * 4 + 2];
        }
        float TS[2][3];
        for (int r = 0; r < 2; r++)
            for (int c = 0; c < 3; c++) TS[r][c] = T[r][0] * S[0][c] + T[r][1] * S[1][c] + T[r][2] * S[2][c];
        const float ca = TS[0][0] * T[0][0] + TS[0][1] * T[0][1] + TS[0][2] * T[0][2] + 0.3f;
        const float cb = TS[0][0] * T[1][0] + TS[0][1] * T[1][1] + TS[0][2] * T[1][2];
        const float cc = TS[1][0] * T[1][0] + TS[1][1] * T[1][1] + TS[1][2] * T[1][2] + 0.3f;
        const float det = ca * cc - cb * cb;
        if (det == 0.0f) continue;
        const float dinv = 1.0f / det;
        const float mid = 0.5f * (ca + cc);
        const float disc = std::sqrt(std::max(0.1f, mid * mid - det));
        const float l1 = mid + disc, l2 = mid - disc;
        const int rad = (int)std::ceil(3.0f * std::sqrt(std::max(l1, l2)));
        const float ix = ((ndx + 1.0f) * (float)a->W - 1.0f) * 0.5f;
        const float iy = ((ndy + 1.0f) * (float)a->H - 1.0f) * 0.5f;
        int ex = rad, ey = rad;
        if (a->use_rects) {
            ex = (int)std::ceil(3.0f * std::sqrt(ca));
            ey = (int)std::ceil(3.0f * std::sqrt(cc));
        }
        const float tf = (float)a->tile;
        const int r0x = std::min(gx, std::max(0, (int)((ix - (float)ex) / tf)));
        const int r0y = std::min(gy, std::max(0, (int)((iy - (float)ey) / tf)));
        const int r1x = std::min(gx, std::max(0, (int)((ix + (float)ex + tf - 1.0f) / tf)));
        const int r1y = std::min(gy, std::max(0, (int)((iy + (float)ey + tf - 1.0f) / tf)));
        const int area = (r1x - r0x) * (r1y - r0y);
        if (area == 0) continue;
        // view-dependent colour, SH degree D
        float dx = px - a->campos[0], dy = py - a->campos[1], dz = pz - a->campos[2];
        const float dn = std::sqrt(dx * dx + dy * dy + dz * dz);
        dx /= dn; dy /= dn; dz /= dn;
        const float* sh = a->shs + 48 * (size_t)g;
        float col[3];
        for (int ch = 0; ch < 3; ch++) {
            auto c = [&](int k) { return sh[k * 3 + ch]; };
            float r = SH_C0 * c(0);
            if (a->D > 0) {
                r = r - SH_C1 * dy * c(1) + SH_C1 * dz * c(2) - SH_C1 * dx * c(3);
                if (a->D > 1) {
                    const float xx = dx * dx, yy = dy * dy, zz = dz * dz, xy2 = dx * dy, yz = dy * dz, xz = dx * dz;
                    r = r + SH_C2[0] * xy2 * c(4) + SH_C2[1] * yz * c(5) + SH_C2[2] * (2.0f * zz - xx - yy) * c(6) +
                        SH_C2[3] * xz * c(7) + SH_C2[4] * (xx - yy) * c(8);
                    if (a->D > 2) {
                        r = r + SH_C3[0] * dy * (3.0f * xx - yy) * c(9) + SH_C3[1] * xy2 * dz * c(10) +
                            SH_C3[2] * dy * (4.0f * zz - xx - yy) * c(11) +
                            SH_C3[3] * dz * (2.0f * zz - 3.0f * xx - 3.0f * yy) * c(12) +
                            SH_C3[4] * dx * (4.0f * zz - xx - yy) * c(13) + SH_C3[5] * dz * (xx - yy) * c(14) +
                            SH_C3[6] * dx * (xx - 3.0f * yy) * c(15);
                    }
                }
            }
            col[ch] = std::max(r + 0.5f, 0.0f);
        }
        depth[g] = vz;
        radius[g] = rad;
        xy[2 * g] = ix; xy[2 * g + 1] = iy;
        conic[4 * g] = cc * dinv; conic[4 * g + 1] = -cb * dinv; conic[4 * g + 2] = ca * dinv;
        conic[4 * g + 3] = a->opac[g];
        for (int ch = 0; ch < 3; ch++) rgb[3 * g + ch] = col[ch];
        rect[4 * g] = r0x; rect[4 * g + 1] = r0y; rect[4 * g + 2] = r1x; rect[4 * g + 3] = r1y;
        touched[g] = area;
    }
}

// Key emission + stable sort + tile ranges. keys/vals sized to the total
// instance count (sum of touched); ranges is (gx*gy)*2 [start, end).
long orc_raster_bin(const orc_raster_args* a, const float* depth, const int* rect, const int* touched,
                    uint64_t* keys, uint32_t* vals, int* ranges) {
    const int gx = (a->W + a->tile - 1) / a->tile, gy = (a->H + a->tile - 1) / a->tile;
    long n = 0;
    for (int g = 0; g < a->P; g++) {
        if (!touched[g]) continue;
        uint32_t db;
        std::memcpy(&db, &depth[g], 4);
        for (int y = rect[4 * g + 1]; y < rect[4 * g + 3]; y++)
            for (int x = rect[4 * g]; x < rect[4 * g + 2]; x++) {
                keys[n] = ((uint64_t)(uint32_t)(y * gx + x) << 32) | db;
                vals[n] = (uint32_t)g;
                n++;
            }
    }
    std::vector<long> ord((size_t)n);
    for (long i = 0; i < n; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](long p, long q) { return keys[p] < keys[q]; });
    std::vector<uint64_t> k2((size_t)n);
    std::vector<uint32_t> v2((size_t)n);
    for (long i = 0; i < n; i++) { k2[i] = keys[ord[i]]; v2[i] = vals[ord[i]]; }
    std::copy(k2.begin(), k2.end(), keys);
    std::copy(v2.begin(), v2.end(), vals);
    for (int t = 0; t < gx * gy; t++) { ranges[2 * t] = 0; ranges[2 * t + 1] = 0; }
    for (long i = 0; i < n; i++) {
        const int t = (int)(keys[i] >> 32);
        if (i == 0 || (int)(keys[i - 1] >> 32) != t) ranges[2 * t] = (int)i;
        if (i == n - 1 || (int)(keys[i + 1] >> 32) != t) ranges[2 * t + 1] = (int)i + 1;
    }
    return n;
}

// Front-to-back blending per pixel over the tile's sorted list; out is planar
// CHW float RGB (shaders/copy.frag:38-40). surface (P ints, may be NULL) gets 1
// for every Gaussian that contributed to a pixel.
void orc_raster_blend(const orc_raster_args* a, const uint32_t* vals, const int* ranges, const float* xy,
                      const float* conic, const float* rgb, float* out, int* surface) {
    const int gx = (a->W + a->tile - 1) / a->tile;
    std::vector<char> contrib(surface ? (size_t)a->P : 0, 0);
#pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < a->H; py++) {
        for (int px = 0; px < a->W; px++) {
            const int t = (py / a->tile) * gx + px / a->tile;
            float T = 1.0f, C[3] = {0.0f, 0.0f, 0.0f};
            const float fxp = (float)px, fyp = (float)py;
            for (int i = ranges[2 * t]; i < ranges[2 * t + 1]; i++) {
                const uint32_t g = vals[i];
                const float dx = xy[2 * g] - fxp, dy = xy[2 * g + 1] - fyp;
                const float* co = conic + 4 * (size_t)g;
                const float power = -0.5f * (co[0] * dx * dx + co[2] * dy * dy) - co[1] * dx * dy;
                if (power > 0.0f) continue;
                const float alpha = std::min(0.99f, co[3] * std::exp(power));
                if (alpha < 1.0f / 255.0f) continue;
                const float test_T = T * (1.0f - alpha);
                if (test_T < 0.0001f) break;
                for (int ch = 0; ch < 3; ch++) C[ch] += rgb[3 * g + ch] * alpha * T;
                if (surface) contrib[g] = 1;
                T = test_T;
            }
            for (int ch = 0; ch < 3; ch++) out[(size_t)ch * a->W * a->H + (size_t)py * a->W + px] = C[ch] + T * a->bg[ch];
        }
    }
    if (surface)
        for (int g = 0; g < a->P; g++)
            if (contrib[g]) surface[g] = 1;
}

}  // extern "C"
