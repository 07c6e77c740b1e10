// TEST INFRASTRUCTURE — the oracle (see common.hpp). Scene AABB, the G^3 cell
// grid with its Gaussian lists and 4^3 samples per valid cell, and the density
// sample.
#include "common.hpp"

#include <algorithm>
#include <vector>

using namespace orc;

extern "C" {

// getOverallAABB (GaussianView.cpp:3601-3631) in float, then the grid step of
// getGridSamples (GaussianView.cpp:3873-3877): out = {min.xyz, max.xyz, step}.
void orc_scene_aabb(const float* pos, long N, int G, float* out7) {
    float mn[3] = {pos[0], pos[1], pos[2]}, mx[3] = {pos[0], pos[1], pos[2]};
    for (long i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) {
            mn[c] = std::min(mn[c], pos[3 * i + c]);
            mx[c] = std::max(mx[c], pos[3 * i + c]);
        }
    float step = 0.0f;
    for (int c = 0; c < 3; c++) {
        const float mean = (mx[c] + mn[c]) / 2.0f;
        const float lo = (mn[c] - mean) * 1.1f + mean;
        const float hi = ((mx[c] - mean) * 1.1f + mean - lo) * 1.0078125f + lo;
        out7[c] = std::min(lo, -0.75f);
        out7[3 + c] = std::max(hi, 0.75f);
        step = std::max(step, (out7[3 + c] - out7[c]) / (float)G);
    }
    out7[6] = step;
}

// Cell of a point (GaussianView.cpp:3909-3912, cudakdtree.cu:408-412).
static inline int cell_of(const float* p, const float* mn, float step, int G) {
    const int x = (int)std::floor((p[0] - mn[0]) / step);
    const int y = (int)std::floor((p[1] - mn[1]) / step);
    const int z = (int)std::floor((p[2] - mn[2]) / step);
    return x * G * G + y * G + z;
}

// First half of getGridSamples (GaussianView.cpp:3889-3953): order[new] = old,
// a stable sort of the Gaussians by the cell of their centre.
void orc_grid_reorder(const float* pos, long N, const float* mn, float step, int G, long* order) {
    std::vector<std::pair<int, long>> key((size_t)N);
    for (long i = 0; i < N; i++) key[i] = {cell_of(pos + 3 * i, mn, step, G), i};
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<int, long>& a, const std::pair<int, long>& b) { return a.first < b.first; });
    for (long i = 0; i < N; i++) order[i] = key[i].second;
}

// Per-Gaussian alpha-cutoff AABB (GaussianView.cpp:3960-4019): half-axes
// sqrt(-2 ln(alpha_c/opacity)) s_i (0 if opacity <= alpha_c) along the rotated
// axes. out: N*6 (min.xyz, max.xyz).
void orc_gaussian_aabbs(const float* pos, const float* rot, const float* scale, const float* opacity, long N,
                        float* out) {
    const float cutoff = 1.0f / 255.0f;
#pragma omp parallel for schedule(static)
    for (long g = 0; g < N; g++) {
        float R[3][3];
        quat_to_mat(quat_normalized(Qf{rot[4 * g], rot[4 * g + 1], rot[4 * g + 2], rot[4 * g + 3]}), R);
        float s3[3] = {0.0f, 0.0f, 0.0f};
        if (opacity[g] > cutoff) {
            // ln rounded from double: the correctly rounded float logarithm
            // (the reference's glibc logf agrees to within its <1 ulp error).
            const float lg = (float)std::log((double)(cutoff / opacity[g]));
            const float f = std::sqrt(-2.0f * lg);
            for (int i = 0; i < 3; i++) s3[i] = f * (scale[3 * g + i] + 0.0f);
        }
        float lo[3], hi[3];
        for (int c = 0; c < 3; c++) lo[c] = hi[c] = pos[3 * g + c];
        for (int d = 0; d < 3; d++) {
            for (int sgn = 0; sgn < 2; sgn++) {
                for (int c = 0; c < 3; c++) {
                    const float off = R[c][d] * s3[d];
                    const float v = sgn == 0 ? pos[3 * g + c] + off : pos[3 * g + c] + (-off);
                    lo[c] = std::min(lo[c], v);
                    hi[c] = std::max(hi[c], v);
                }
            }
        }
        for (int c = 0; c < 3; c++) { out[6 * g + c] = lo[c]; out[6 * g + 3 + c] = hi[c]; }
    }
}

// Second half of getGridSamples (GaussianView.cpp:4021-4133) plus the
// cell-range rule of GetBoxesGsGrid (cudakdtree.cu:425-451):
//   prefix (G^3, inclusive) of per-cell box counts, the per-cell Gaussian lists
//   in ascending Gaussian order, the valid (non-empty) cells ascending, and the
//   4^3 samples of every valid cell. Returns the list length; pass list=NULL to
//   size it first. valid/samples may be NULL.
long orc_grid_lists(const float* aabbs, long N, const float* mn, float step, int G, int padding, int* prefix,
                    int* list, int* valid, long* n_valid, float* samples) {
    const long GC = (long)G * G * G;
    std::vector<int> cnt((size_t)GC, 0);
    auto range = [&](long g, int lo[3], int hi[3]) {
        for (int c = 0; c < 3; c++) {
            lo[c] = std::max((int)(std::floor((aabbs[6 * g + c] - mn[c]) / step) - padding), 0);
            hi[c] = std::min((int)(std::floor((aabbs[6 * g + 3 + c] - mn[c]) / step) + padding), G - 1);
        }
    };
    for (long g = 0; g < N; g++) {
        int lo[3], hi[3];
        range(g, lo, hi);
        for (int x = lo[0]; x <= hi[0]; x++)
            for (int y = lo[1]; y <= hi[1]; y++)
                for (int z = lo[2]; z <= hi[2]; z++) cnt[(size_t)x * G * G + y * G + z]++;
    }
    long run = 0, nv = 0;
    for (long c = 0; c < GC; c++) {
        run += cnt[c];
        prefix[c] = (int)run;
        if (cnt[c]) {
            if (valid) valid[nv] = (int)c;
            nv++;
        }
    }
    if (n_valid) *n_valid = nv;
    if (list) {
        std::vector<int> used((size_t)GC, 0);
        for (long g = 0; g < N; g++) {
            int lo[3], hi[3];
            range(g, lo, hi);
            for (int x = lo[0]; x <= hi[0]; x++)
                for (int y = lo[1]; y <= hi[1]; y++)
                    for (int z = lo[2]; z <= hi[2]; z++) {
                        const long c = (long)x * G * G + y * G + z;
                        const long base = c == 0 ? 0 : prefix[c - 1];
                        list[base + used[c]++] = (int)g;
                    }
        }
    }
    if (samples && valid) {
        const float interval = step / 4;
        for (long v = 0; v < nv; v++) {
            const int c = valid[v];
            const int xi = c / (G * G), yi = (c - xi * G * G) / G, zi = c % G;
            const float x0 = (float)((double)(mn[0] + xi * step) + 0.5 * (double)interval);
            const float y0 = (float)((double)(mn[1] + yi * step) + 0.5 * (double)interval);
            const float z0 = (float)((double)(mn[2] + zi * step) + 0.5 * (double)interval);
            for (int s = 0; s < 64; s++) {
                const int sx = s / 16, sy = (s - sx * 16) / 4, sz = s % 4;
                float* o = samples + (v * 64 + s) * 3;
                o[0] = x0 + sx * interval;
                o[1] = y0 + sy * interval;
                o[2] = z0 + sz * interval;
            }
        }
    }
    return run;
}

// Density sample (the un-vendored forward3d_grid, GaussianView.cpp:4159-4186;
// formula defined here, parity with the reference unpinned): for every sample x
// of valid cell v and Gaussian g in v's list,
//   a_g(x) = o_g exp(-1/2 d^T (Sigma_g + L)^-1 d),  d = x - mu_g,
//   L = lpf I, or ada[6v ..] (xx, yy, zz, xy, xz, yz) when ada is given,
//   feature(x) = sum_g a_g(x) sh_g (48 floats),  opacity(x) = sum_g a_g(x).
// Computed in double, stored as float.
void orc_density_sample(long n_valid, const int* valid, const int* prefix, const int* list, const float* samples,
                        const float* pos, const float* rot, const float* scale, const float* opacity,
                        const float* sh, float lpf, float* feat, float* opac, const float* ada) {
#pragma omp parallel for schedule(dynamic, 16)
    for (long v = 0; v < n_valid; v++) {
        const int c = valid[v];
        const long b = c == 0 ? 0 : prefix[c - 1], e = prefix[c];
        for (int s = 0; s < 64; s++) {
            const float* x = samples + (v * 64 + s) * 3;
            double acc[48] = {0}, aop = 0.0;
            for (long li = b; li < e; li++) {
                const int g = list[li];
                float Rf[3][3];
                quat_to_mat(quat_normalized(Qf{rot[4 * g], rot[4 * g + 1], rot[4 * g + 2], rot[4 * g + 3]}), Rf);
                double S[3][3];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        double a = 0.0;
                        for (int k = 0; k < 3; k++) {
                            const double sk = scale[3 * g + k];
                            a += (double)Rf[i][k] * sk * sk * (double)Rf[j][k];
                        }
                        S[i][j] = a + (ada ? 0.0 : (i == j ? (double)lpf : 0.0));
                    }
                if (ada) {
                    const float* L = ada + 6 * v;
                    S[0][0] += L[0]; S[1][1] += L[1]; S[2][2] += L[2];
                    S[0][1] += L[3]; S[1][0] += L[3];
                    S[0][2] += L[4]; S[2][0] += L[4];
                    S[1][2] += L[5]; S[2][1] += L[5];
                }
                // inverse of the symmetric 3x3 by cofactors
                const double c00 = S[1][1] * S[2][2] - S[1][2] * S[2][1];
                const double c01 = S[1][2] * S[2][0] - S[1][0] * S[2][2];
                const double c02 = S[1][0] * S[2][1] - S[1][1] * S[2][0];
                const double det = S[0][0] * c00 + S[0][1] * c01 + S[0][2] * c02;
                const double i00 = c00 / det, i01 = c01 / det, i02 = c02 / det;
                const double i11 = (S[0][0] * S[2][2] - S[0][2] * S[2][0]) / det;
                const double i12 = (S[0][2] * S[1][0] - S[0][0] * S[1][2]) / det;
                const double i22 = (S[0][0] * S[1][1] - S[0][1] * S[1][0]) / det;
                const double d0 = (double)x[0] - pos[3 * g], d1 = (double)x[1] - pos[3 * g + 1],
                             d2 = (double)x[2] - pos[3 * g + 2];
                const double q = i00 * d0 * d0 + i11 * d1 * d1 + i22 * d2 * d2 +
                                 2.0 * (i01 * d0 * d1 + i02 * d0 * d2 + i12 * d1 * d2);
                const double a = (double)opacity[g] * std::exp(-0.5 * q);
                aop += a;
                for (int k = 0; k < 48; k++) acc[k] += a * (double)sh[48 * (size_t)g + k];
            }
            float* fo = feat + (v * 64 + s) * 48;
            for (int k = 0; k < 48; k++) fo[k] = (float)acc[k];
            opac[v * 64 + s] = (float)aop;
        }
    }
}

}  // extern "C"
