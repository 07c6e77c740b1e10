// Host construction of the SH-rotation resampling tables (ShRotTables): for
// bands 2 and 3, a set of 2l+1 unit directions whose basis matrix
// Yl[k][m] = Y_m(d_k) is well conditioned, and its inverse. The directions come
// from a fixed-seed search, so every context builds identical tables.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "common.hpp"

namespace arapgs {

namespace {

void basis(int l, const double d[3], double* b) {
    const double x = d[0], y = d[1], z = d[2], xx = x * x, yy = y * y, zz = z * z;
    if (l == 2) {
        b[0] = 1.0925484305920792 * x * y;
        b[1] = -1.0925484305920792 * y * z;
        b[2] = 0.31539156525252005 * (2.0 * zz - xx - yy);
        b[3] = -1.0925484305920792 * x * z;
        b[4] = 0.5462742152960396 * (xx - yy);
    } else {
        b[0] = -0.5900435899266435 * y * (3.0 * xx - yy);
        b[1] = 2.890611442640554 * x * y * z;
        b[2] = -0.4570457994644658 * y * (4.0 * zz - xx - yy);
        b[3] = 0.3731763325901154 * z * (2.0 * zz - 3.0 * xx - 3.0 * yy);
        b[4] = -0.4570457994644658 * x * (4.0 * zz - xx - yy);
        b[5] = 1.445305721320277 * z * (xx - yy);
        b[6] = -0.5900435899266435 * x * (xx - 3.0 * yy);
    }
}

// Gauss-Jordan with partial pivoting; returns false if singular.
bool invert(int n, const double* A, double* inv) {
    double a[7][14];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 2 * n; j++) a[i][j] = j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0);
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (std::fabs(a[p][c]) < 1e-12) return false;
        if (p != c)
            for (int j = 0; j < 2 * n; j++) std::swap(a[p][j], a[c][j]);
        const double d = a[c][c];
        for (int j = 0; j < 2 * n; j++) a[c][j] /= d;
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = a[r][c];
            for (int j = 0; j < 2 * n; j++) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) inv[i * n + j] = a[i][n + j];
    return true;
}

struct Lcg {
    uint64_t s;
    double uniform() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    }
};

void search(int l, float dir[][3], float* inv_out) {
    const int n = 2 * l + 1;
    Lcg rng{0x5eed0000ULL + (uint64_t)l};
    double best = 1e300;
    double bestD[7][3], bestInv[49];
    for (int trial = 0; trial < 4000; trial++) {
        double D[7][3], Y[49], I[49];
        for (int k = 0; k < n; k++) {
            // uniform direction: z in [-1,1], phi in [0, 2pi)
            const double z = 2.0 * rng.uniform() - 1.0, phi = 6.283185307179586 * rng.uniform();
            const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
            // round to float first: the kernels see the float directions
            D[k][0] = (double)(float)(r * std::cos(phi));
            D[k][1] = (double)(float)(r * std::sin(phi));
            D[k][2] = (double)(float)z;
            basis(l, D[k], Y + k * n);
        }
        if (!invert(n, Y, I)) continue;
        double fy = 0.0, fi = 0.0;
        for (int e = 0; e < n * n; e++) { fy += Y[e] * Y[e]; fi += I[e] * I[e]; }
        const double cond = std::sqrt(fy * fi);
        if (cond < best) {
            best = cond;
            std::memcpy(bestD, D, sizeof(D));
            std::memcpy(bestInv, I, sizeof(I));
        }
    }
    for (int k = 0; k < n; k++)
        for (int c = 0; c < 3; c++) dir[k][c] = (float)bestD[k][c];
    // inv_out[m][k] = (Y^-1)[m][k]
    for (int e = 0; e < n * n; e++) inv_out[e] = (float)bestInv[e];
}

}  // namespace

void build_shrot_tables(ShRotTables* T) {
    search(2, T->dir2, &T->inv2[0][0]);
    search(3, T->dir3, &T->inv3[0][0]);
}

}  // namespace arapgs
