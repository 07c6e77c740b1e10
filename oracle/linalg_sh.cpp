// TEST INFRASTRUCTURE — the oracle (see common.hpp). Polar decomposition and
// real spherical-harmonic rotation.
#include "common.hpp"

#include <algorithm>

namespace orc {

void polar_fp64(const double M[3][3], double R[3][3], double S[3][3]) {
    // One-sided Jacobi SVD: rotate column pairs of A = M V until they are
    // mutually orthogonal; then A = U diag(sig).
    double A[3][3], V[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { A[i][j] = M[i][j]; V[i][j] = (i == j) ? 1.0 : 0.0; }
    static const int P[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (int pr = 0; pr < 3; pr++) {
            const int p = P[pr][0], q = P[pr][1];
            double a = 0.0, b = 0.0, g = 0.0;
            for (int i = 0; i < 3; i++) {
                a += A[i][p] * A[i][p];
                b += A[i][q] * A[i][q];
                g += A[i][p] * A[i][q];
            }
            if (g == 0.0 || a == 0.0 || b == 0.0) continue;
            const double rel = std::fabs(g) / std::sqrt(a * b);
            off = std::max(off, rel);
            if (rel < 1e-16) continue;
            const double zeta = (b - a) / (2.0 * g);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
            const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
            for (int i = 0; i < 3; i++) {
                const double ap = A[i][p], aq = A[i][q];
                A[i][p] = c * ap - s * aq;
                A[i][q] = s * ap + c * aq;
                const double vp = V[i][p], vq = V[i][q];
                V[i][p] = c * vp - s * vq;
                V[i][q] = s * vp + c * vq;
            }
        }
        if (off < 1e-15) break;
    }
    double sig[3], U[3][3];
    for (int j = 0; j < 3; j++) {
        sig[j] = std::sqrt(A[0][j] * A[0][j] + A[1][j] * A[1][j] + A[2][j] * A[2][j]);
        for (int i = 0; i < 3; i++) U[i][j] = sig[j] > 0.0 ? A[i][j] / sig[j] : 0.0;
    }
    // Rank-deficient input: complete U with a unit vector orthogonal to the
    // other columns (the polar factor is not unique there).
    for (int j = 0; j < 3; j++) {
        if (sig[j] > 1e-300) continue;
        const int a = (j + 1) % 3, b = (j + 2) % 3;
        double c0 = U[1][a] * U[2][b] - U[2][a] * U[1][b];
        double c1 = U[2][a] * U[0][b] - U[0][a] * U[2][b];
        double c2 = U[0][a] * U[1][b] - U[1][a] * U[0][b];
        double n = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        if (n == 0.0) { c0 = (j == 0); c1 = (j == 1); c2 = (j == 2); n = 1.0; }
        U[0][j] = c0 / n; U[1][j] = c1 / n; U[2][j] = c2 / n;
    }
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) {
            double r = 0.0, s = 0.0;
            for (int j = 0; j < 3; j++) {
                r += U[i][j] * V[k][j];
                s += V[i][j] * sig[j] * V[k][j];
            }
            R[i][k] = r;
            S[i][k] = s;
        }
}

// ---------------------------------------------------------------------------
// SH rotation by the Ivanic-Ruedenberg recurrence (J. Ivanic, K. Ruedenberg,
// J. Phys. Chem. 100 (1996) 6342; erratum 102 (1998) 9099). This is the
// recurrence the reference unrolls for degrees 2 and 3 in
// Construct_SH_Rotation_Matrix (helper.cpp:979-1075). The band-1 block is the
// (y, z, x) permutation of R (helper.cpp:979-987).
// Index convention: band l matrix B_l[m+l][n+l], m,n in [-l, l].
namespace {

struct Bands { double b[4][7][7]; };

inline double& at(Bands& B, int l, int m, int n) { return B.b[l][m + l][n + l]; }
inline double get(const Bands& B, int l, int m, int n) { return B.b[l][m + l][n + l]; }

double Pfun(const Bands& B, int i, int l, int a, int b) {
    const double ri1 = get(B, 1, i, 1), rim1 = get(B, 1, i, -1), ri0 = get(B, 1, i, 0);
    if (b == l) return ri1 * get(B, l - 1, a, l - 1) - rim1 * get(B, l - 1, a, -l + 1);
    if (b == -l) return ri1 * get(B, l - 1, a, -l + 1) + rim1 * get(B, l - 1, a, l - 1);
    return ri0 * get(B, l - 1, a, b);
}

void build_bands(const float R[3][3], Bands& B) {
    std::memset(&B, 0, sizeof(B));
    B.b[0][0][0] = 1.0;
    static const int ax[3] = {1, 2, 0};  // m = -1, 0, 1  ->  y, z, x
    for (int m = -1; m <= 1; m++)
        for (int n = -1; n <= 1; n++) at(B, 1, m, n) = (double)R[ax[m + 1]][ax[n + 1]];
    for (int l = 2; l <= 3; l++) {
        for (int m = -l; m <= l; m++) {
            for (int n = -l; n <= l; n++) {
                const int d = (m == 0) ? 1 : 0;
                const int am = m < 0 ? -m : m;
                const double denom = (n == l || n == -l) ? (double)(2 * l) * (2 * l - 1)
                                                         : (double)(l + n) * (l - n);
                const double u = std::sqrt((double)(l + m) * (l - m) / denom);
                const double v = 0.5 * std::sqrt((1.0 + d) * (l + am - 1) * (l + am) / denom) * (1 - 2 * d);
                const double w = -0.5 * std::sqrt((double)(l - am - 1) * (l - am) / denom) * (1 - d);
                double val = 0.0;
                if (u != 0.0) val += u * Pfun(B, 0, l, m, n);
                if (v != 0.0) {
                    double V;
                    if (m == 0) {
                        V = Pfun(B, 1, l, 1, n) + Pfun(B, -1, l, -1, n);
                    } else if (m > 0) {
                        const double d1 = (m == 1) ? 1.0 : 0.0;
                        V = Pfun(B, 1, l, m - 1, n) * std::sqrt(1.0 + d1) - Pfun(B, -1, l, -m + 1, n) * (1.0 - d1);
                    } else {
                        const double d1 = (m == -1) ? 1.0 : 0.0;
                        V = Pfun(B, 1, l, m + 1, n) * (1.0 - d1) + Pfun(B, -1, l, -m - 1, n) * std::sqrt(1.0 + d1);
                    }
                    val += v * V;
                }
                if (w != 0.0) {
                    double W;
                    if (m > 0) W = Pfun(B, 1, l, m + 1, n) + Pfun(B, -1, l, -m - 1, n);
                    else W = Pfun(B, 1, l, m - 1, n) - Pfun(B, -1, l, -m + 1, n);
                    val += w * W;
                }
                at(B, l, m, n) = val;
            }
        }
    }
}

}  // namespace

// SH_Rotation (helper.cpp:955-977) with the odd-index sign wrapper of
// GaussianView.cpp:3138-3156 (and the folded device form, cudakdtree.cu:160-195).
// The band matrices are rounded to float, as the reference stores them in
// Eigen::Matrix<float,...>, and applied per colour channel.
void sh_rotate(const float R[3][3], float* sh, bool flip_odd) {
    Bands B;
    build_bands(R, B);
    float in[16][3];
    for (int c = 0; c < 16; c++)
        for (int ch = 0; ch < 3; ch++) {
            const float v = sh[c * 3 + ch];
            in[c][ch] = (flip_odd && (c & 1)) ? -v : v;
        }
    float out[16][3];
    for (int ch = 0; ch < 3; ch++) out[0][ch] = in[0][ch];
    for (int l = 1; l <= 3; l++) {
        const int base = l * l;
        for (int m = 0; m < 2 * l + 1; m++)
            for (int ch = 0; ch < 3; ch++) {
                float acc = 0.0f;
                for (int n = 0; n < 2 * l + 1; n++) acc += (float)B.b[l][m][n] * in[base + n][ch];
                out[base + m][ch] = acc;
            }
    }
    for (int c = 0; c < 16; c++)
        for (int ch = 0; ch < 3; ch++) {
            const float v = out[c][ch];
            sh[c * 3 + ch] = (flip_odd && (c & 1)) ? -v : v;
        }
}

}  // namespace orc

extern "C" {

// Polar decomposition (test hook): M, R, S row-major 3x3 FP64.
void orc_polar(const double* M, double* R, double* S) {
    double m[3][3], r[3][3], s[3][3];
    for (int i = 0; i < 9; i++) m[i / 3][i % 3] = M[i];
    orc::polar_fp64(m, r, s);
    for (int i = 0; i < 9; i++) { R[i] = r[i / 3][i % 3]; S[i] = s[i / 3][i % 3]; }
}

// SH rotation (test hook): R row-major float 3x3; sh: n x 48 in place.
void orc_sh_rotate(const float* R, float* sh, long n, int flip_odd) {
    float r[3][3];
    for (int i = 0; i < 9; i++) r[i / 3][i % 3] = R[i];
    for (long i = 0; i < n; i++) orc::sh_rotate(r, sh + i * 48, flip_odd != 0);
}

// Quaternion <-> matrix (test hooks), quaternion (w,x,y,z), matrix row-major.
void orc_mat_to_quat(const float* m, float* q) {
    float a[3][3];
    for (int i = 0; i < 9; i++) a[i / 3][i % 3] = m[i];
    orc::Qf r = orc::mat_to_quat(a);
    q[0] = r.w; q[1] = r.x; q[2] = r.y; q[3] = r.z;
}
void orc_quat_to_mat(const float* q, float* m) {
    float a[3][3];
    orc::quat_to_mat(orc::Qf{q[0], q[1], q[2], q[3]}, a);
    for (int i = 0; i < 9; i++) m[i] = a[i / 3][i % 3];
}

}  // extern "C"
