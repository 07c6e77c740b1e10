// TEST INFRASTRUCTURE — the oracle (see common.hpp). FP64 forward of the
// differentiable deform -> density-sample -> render chain, written once over a
// scalar type T and instantiated for double (values) and for a forward-mode
// dual number (value + directional derivative). The dual instantiation gives
// the exact FP64 directional derivative J v of every output along a direction
// v of the inputs (node R and t, rest rotations q0, rest scales s0); the HIP
// backward is checked through the adjoint identity <g, J v> = <J^T g, v>.
//
// Chain (reference call sites): GetEndPoints (GaussianView.cpp:4643-4668),
// predict_mesh / predict_samples (Deform.hpp:230-268),
// UpdateAsSixPointsWithdrawBad (GaussianView.cpp:3081-3166) with q_old = q0 and
// the backup scales = s0, Rasterizer::forward (GaussianView.cpp:1099-1128,
// un-vendored: the 3DGS forward of raster.cpp), and the density sample of the
// deformed Gaussians at the blended samples (the formula of grid.cpp's
// density restatement). Where the HIP library evaluates in float, this
// restatement evaluates in FP64: the polar factor by Higham's Newton iteration
// (X <- (X + X^-T) / 2) instead of a Jacobi SVD, the SH rotation by the
// Ivanic-Ruedenberg recurrence instead of resampling. The discrete structure of
// the render (culling, tile rectangles, depth order) is taken from the float
// restatement on the float-rounded deformed Gaussians, as the GPU sees them.
#include "common.hpp"

#include <algorithm>
#include <vector>

#include "raster.hpp"

namespace {

// ------------------------------------------------------------------ dual numbers
struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double x) : v(x), d(0.0) {}  // NOLINT: implicit from constants
    Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
inline Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
inline Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
inline Dual& operator/=(Dual& a, Dual b) { return a = a / b; }
inline Dual sqrt(Dual a) {
    const double s = std::sqrt(a.v);
    return {s, s > 0.0 ? a.d / (2.0 * s) : 0.0};
}
inline Dual exp(Dual a) {
    const double e = std::exp(a.v);
    return {e, e * a.d};
}
inline double val(double x) { return x; }
inline double val(Dual x) { return x.v; }
template <class T> inline T tmin(T a, T b) { return val(b) < val(a) ? b : a; }
template <class T> inline T tmax(T a, T b) { return val(b) > val(a) ? b : a; }
using std::exp;
using std::sqrt;

// ------------------------------------------------------------------ rotations
template <class T>
void quat_to_mat_t(const T q[4], T R[3][3]) {
    const T w = q[0], x = q[1], y = q[2], z = q[3];
    R[0][0] = 1.0 - 2.0 * (y * y + z * z); R[0][1] = 2.0 * (x * y - w * z); R[0][2] = 2.0 * (x * z + w * y);
    R[1][0] = 2.0 * (x * y + w * z); R[1][1] = 1.0 - 2.0 * (x * x + z * z); R[1][2] = 2.0 * (y * z - w * x);
    R[2][0] = 2.0 * (x * z - w * y); R[2][1] = 2.0 * (y * z + w * x); R[2][2] = 1.0 - 2.0 * (x * x + y * y);
}

template <class T>
void normalize4(T q[4]) {
    const T n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (val(n) > 0.0)
        for (int i = 0; i < 4; i++) q[i] = q[i] / n;
}

// Shoemake's conversion (the branch is chosen on values).
template <class T>
void mat_to_quat_t(const T m[3][3], T q[4]) {
    const T t = m[0][0] + m[1][1] + m[2][2];
    if (val(t) > 0.0) {
        T s = sqrt(t + 1.0);
        q[0] = 0.5 * s;
        s = 0.5 / s;
        q[1] = (m[2][1] - m[1][2]) * s;
        q[2] = (m[0][2] - m[2][0]) * s;
        q[3] = (m[1][0] - m[0][1]) * s;
        return;
    }
    int i = 0;
    if (val(m[1][1]) > val(m[0][0])) i = 1;
    if (val(m[2][2]) > val(m[i][i])) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    T s = sqrt(m[i][i] - m[j][j] - m[k][k] + 1.0);
    T v[3];
    v[i] = 0.5 * s;
    s = 0.5 / s;
    q[0] = (m[k][j] - m[j][k]) * s;
    v[j] = (m[j][i] + m[i][j]) * s;
    v[k] = (m[k][i] + m[i][k]) * s;
    q[1] = v[0]; q[2] = v[1]; q[3] = v[2];
}

template <class T>
void quat_mul_t(const T a[4], const T b[4], T o[4]) {
    o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    o[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    o[2] = a[0] * b[2] + a[2] * b[0] + a[3] * b[1] - a[1] * b[3];
    o[3] = a[0] * b[3] + a[3] * b[0] + a[1] * b[2] - a[2] * b[1];
}

template <class T>
bool inverse3(const T A[3][3], T I[3][3]) {
    const T c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const T c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const T c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    const T det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    if (val(det) == 0.0) return false;
    const T id = 1.0 / det;
    I[0][0] = c00 * id;
    I[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
    I[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
    I[1][0] = c01 * id;
    I[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
    I[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
    I[2][0] = c02 * id;
    I[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
    I[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
    return true;
}

// Polar factor of a nonsingular M by Newton's iteration X <- (X + X^-T)/2
// (N. J. Higham, SIAM J. Sci. Stat. Comput. 7 (1986) 1160), run to a fixed
// count past convergence so the dual part converges with it; diag(R^T M).
template <class T>
void polar_t(const T M[3][3], T R[3][3], T Sd[3]) {
    T X[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) X[i][j] = M[i][j];
    for (int it = 0; it < 40; it++) {
        T Xi[3][3];
        if (!inverse3(X, Xi)) break;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) X[i][j] = 0.5 * (X[i][j] + Xi[j][i]);
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) R[i][j] = X[i][j];
    for (int i = 0; i < 3; i++) Sd[i] = R[0][i] * M[0][i] + R[1][i] * M[1][i] + R[2][i] * M[2][i];
}

// SH rotation by the Ivanic-Ruedenberg recurrence with the odd-index sign
// wrapper (the oracle's sh_rotate in linalg_sh.cpp, over T and unrounded).
template <class T>
struct BandsT { T b[4][7][7]; };

template <class T>
T bget(const BandsT<T>& B, int l, int m, int n) { return B.b[l][m + l][n + l]; }

template <class T>
T Pfun_t(const BandsT<T>& B, int i, int l, int a, int b) {
    const T ri1 = bget(B, 1, i, 1), rim1 = bget(B, 1, i, -1), ri0 = bget(B, 1, i, 0);
    if (b == l) return ri1 * bget(B, l - 1, a, l - 1) - rim1 * bget(B, l - 1, a, -l + 1);
    if (b == -l) return ri1 * bget(B, l - 1, a, -l + 1) + rim1 * bget(B, l - 1, a, l - 1);
    return ri0 * bget(B, l - 1, a, b);
}

template <class T>
void build_bands_t(const T R[3][3], BandsT<T>& B) {
    for (int l = 0; l < 4; l++)
        for (int i = 0; i < 7; i++)
            for (int j = 0; j < 7; j++) B.b[l][i][j] = 0.0;
    B.b[0][0][0] = 1.0;
    static const int ax[3] = {1, 2, 0};
    for (int m = -1; m <= 1; m++)
        for (int n = -1; n <= 1; n++) B.b[1][m + 1][n + 1] = R[ax[m + 1]][ax[n + 1]];
    for (int l = 2; l <= 3; l++)
        for (int m = -l; m <= l; m++)
            for (int n = -l; n <= l; n++) {
                const int d = (m == 0) ? 1 : 0;
                const int am = m < 0 ? -m : m;
                const double denom = (n == l || n == -l) ? (double)(2 * l) * (2 * l - 1) : (double)(l + n) * (l - n);
                const double u = std::sqrt((double)(l + m) * (l - m) / denom);
                const double v = 0.5 * std::sqrt((1.0 + d) * (l + am - 1) * (l + am) / denom) * (1 - 2 * d);
                const double w = -0.5 * std::sqrt((double)(l - am - 1) * (l - am) / denom) * (1 - d);
                T acc = 0.0;
                if (u != 0.0) acc += u * Pfun_t(B, 0, l, m, n);
                if (v != 0.0) {
                    T Vv;
                    if (m == 0) {
                        Vv = Pfun_t(B, 1, l, 1, n) + Pfun_t(B, -1, l, -1, n);
                    } else if (m > 0) {
                        const double d1 = (m == 1) ? 1.0 : 0.0;
                        Vv = Pfun_t(B, 1, l, m - 1, n) * std::sqrt(1.0 + d1) - Pfun_t(B, -1, l, -m + 1, n) * (1.0 - d1);
                    } else {
                        const double d1 = (m == -1) ? 1.0 : 0.0;
                        Vv = Pfun_t(B, 1, l, m + 1, n) * (1.0 - d1) + Pfun_t(B, -1, l, -m - 1, n) * std::sqrt(1.0 + d1);
                    }
                    acc += v * Vv;
                }
                if (w != 0.0) {
                    T Ww;
                    if (m > 0) Ww = Pfun_t(B, 1, l, m + 1, n) + Pfun_t(B, -1, l, -m - 1, n);
                    else Ww = Pfun_t(B, 1, l, m - 1, n) - Pfun_t(B, -1, l, -m + 1, n);
                    acc += w * Ww;
                }
                B.b[l][m + l][n + l] = acc;
            }
}

template <class T>
void sh_rotate_t(const T R[3][3], const T in[48], T out[48]) {
    BandsT<T> B;
    build_bands_t(R, B);
    T a[16][3];
    for (int c = 0; c < 16; c++)
        for (int ch = 0; ch < 3; ch++) a[c][ch] = (c & 1) ? -in[c * 3 + ch] : in[c * 3 + ch];
    T o[16][3];
    for (int ch = 0; ch < 3; ch++) o[0][ch] = a[0][ch];
    for (int l = 1; l <= 3; l++) {
        const int base = l * l;
        for (int m = 0; m < 2 * l + 1; m++)
            for (int ch = 0; ch < 3; ch++) {
                T acc = 0.0;
                for (int n = 0; n < 2 * l + 1; n++) acc += B.b[l][m][n] * a[base + n][ch];
                o[base + m][ch] = acc;
            }
    }
    for (int c = 0; c < 16; c++)
        for (int ch = 0; ch < 3; ch++) out[c * 3 + ch] = (c & 1) ? -o[c][ch] : o[c][ch];
}

}  // namespace

extern "C" {

struct orc_chain_args {
    long N;
    int K, M;
    const float* pos0;      // N*3
    const float* sh0;       // N*48
    const float* opac;      // N
    const uint32_t* ends_idx; const float* ends_w;  // N*6*K
    const float* node_g;    // M*3 (the table's g)
    long nv;                // valid cells (0: no density term)
    const int* valid; const int* prefix; const int* list;
    const float* samples0;  // S*3, S = 64 nv
    const uint32_t* sample_idx; const float* sample_w;  // S*K
    float lpf;
    int V, W, H, D;
    const float* bg;        // 3
    const float* view;      // V*16
    const float* proj;      // V*16
    const float* campos;    // V*3
    const float* tanfov;    // V*2
    float pad, coeff;
    const float* ada;       // nv*6 adaptive filters (xx, yy, zz, xy, xz, yz) or NULL: lpf I
};

}  // extern "C"

namespace {

template <class T>
void blend_t(const orc_chain_args* a, const T* Rn, const T* tn, const uint32_t* idx, const float* w, const T p[3],
             T o[3]) {
    o[0] = o[1] = o[2] = 0.0;
    for (int j = 0; j < a->K; j++) {
        const uint32_t n = idx[j];
        const double g0 = a->node_g[3 * n], g1 = a->node_g[3 * n + 1], g2 = a->node_g[3 * n + 2];
        const T d0 = p[0] - g0, d1 = p[1] - g1, d2 = p[2] - g2;
        const T* R = Rn + 9 * (size_t)n;
        const T* t = tn + 3 * (size_t)n;
        const double wj = w[j];
        o[0] += wj * (R[0] * d0 + R[1] * d1 + R[2] * d2 + t[0] + g0);
        o[1] += wj * (R[3] * d0 + R[4] * d1 + R[5] * d2 + t[1] + g1);
        o[2] += wj * (R[6] * d0 + R[7] * d1 + R[8] * d2 + t[2] + g2);
    }
}

template <class T>
struct Deformed {
    std::vector<T> pos, rot, scale, sh, samples, opac;
};

template <class T>
void deform_t(const orc_chain_args* a, const T* Rn, const T* tn, const T* q0, const T* s0, Deformed<T>& D) {
    const long N = a->N;
    D.pos.assign(3 * N, T(0.0)); D.rot.assign(4 * N, T(0.0)); D.scale.assign(3 * N, T(0.0)); D.sh.assign(48 * N, T(0.0));
    D.opac.resize(N);
    for (long g = 0; g < N; g++) D.opac[g] = (double)a->opac[g];
#pragma omp parallel for schedule(static)
    for (long g = 0; g < N; g++) {
        T qn[4] = {q0[4 * g], q0[4 * g + 1], q0[4 * g + 2], q0[4 * g + 3]};
        normalize4(qn);
        T R0[3][3];
        quat_to_mat_t(qn, R0);
        T e[6][3];
        for (int i = 0; i < 3; i++) {
            const T len = (s0[3 * g + i] + (double)a->pad) * (double)a->coeff;
            for (int c = 0; c < 3; c++) {
                const T off = R0[c][i] * len;
                e[2 * i][c] = (double)a->pos0[3 * g + c] + off;
                e[2 * i + 1][c] = (double)a->pos0[3 * g + c] - off;
            }
        }
        T eb[6][3];
        for (int p = 0; p < 6; p++) {
            const size_t pt = 6 * (size_t)g + p;
            blend_t(a, Rn, tn, a->ends_idx + pt * a->K, a->ends_w + pt * a->K, e[p], eb[p]);
        }
        for (int c = 0; c < 3; c++) {
            T sum = 0.0;
            for (int p = 0; p < 6; p++) sum += eb[p][c];
            D.pos[3 * g + c] = sum / 6.0;
        }
        T M[3][3];
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++) M[c][i] = 0.5 * (eb[2 * i][c] - eb[2 * i + 1][c]);
        T R[3][3], Sd[3];
        polar_t(M, R, Sd);
        T q[4];
        mat_to_quat_t(R, q);
        normalize4(q);
        for (int k = 0; k < 4; k++) D.rot[4 * g + k] = q[k];
        for (int i = 0; i < 3; i++) {
            const T s = s0[3 * g + i];
            D.scale[3 * g + i] = Sd[i] / ((s + (double)a->pad) * (double)a->coeff) * s;
        }
        const T n2 = q0[4 * g] * q0[4 * g] + q0[4 * g + 1] * q0[4 * g + 1] + q0[4 * g + 2] * q0[4 * g + 2] +
                     q0[4 * g + 3] * q0[4 * g + 3];
        const T qi[4] = {q0[4 * g] / n2, -q0[4 * g + 1] / n2, -q0[4 * g + 2] / n2, -q0[4 * g + 3] / n2};
        T qr[4];
        quat_mul_t(q, qi, qr);
        normalize4(qr);
        T Rr[3][3];
        quat_to_mat_t(qr, Rr);
        T in[48], out[48];
        for (int k = 0; k < 48; k++) in[k] = (double)a->sh0[48 * g + k];
        sh_rotate_t(Rr, in, out);
        for (int k = 0; k < 48; k++) D.sh[48 * g + k] = out[k];
    }
    const long S = 64 * a->nv;
    D.samples.assign(3 * S, T(0.0));
#pragma omp parallel for schedule(static)
    for (long s = 0; s < S; s++) {
        const T p[3] = {(double)a->samples0[3 * s], (double)a->samples0[3 * s + 1], (double)a->samples0[3 * s + 2]};
        T o[3];
        blend_t(a, Rn, tn, a->sample_idx + s * a->K, a->sample_w + s * a->K, p, o);
        for (int c = 0; c < 3; c++) D.samples[3 * s + c] = o[c];
    }
}

// One view: the 3DGS forward of raster.cpp over T for the continuous values;
// culling, tile rectangles and depth order from the float restatement.
template <class T>
void render_t(const orc_chain_args* a, int v, const Deformed<T>& D, T* img) {
    const long N = a->N;
    const int W = a->W, H = a->H, tile = 16;
    std::vector<float> fpos(3 * N), frot(4 * N), fscale(3 * N), fsh(48 * N);
    for (long i = 0; i < 3 * N; i++) fpos[i] = (float)val(D.pos[i]);
    for (long i = 0; i < 4 * N; i++) frot[i] = (float)val(D.rot[i]);
    for (long i = 0; i < 3 * N; i++) fscale[i] = (float)val(D.scale[i]);
    for (long i = 0; i < 48 * N; i++) fsh[i] = (float)val(D.sh[i]);
    const float* Vm = a->view + 16 * v;
    const float* Pm = a->proj + 16 * v;
    const float* cp = a->campos + 3 * v;
    const float tfx = a->tanfov[2 * v], tfy = a->tanfov[2 * v + 1];
    orc_raster_args ra{(int)N, a->D, W, H, tile, a->bg, fpos.data(), fsh.data(), a->opac, fscale.data(), 1.0f,
                       frot.data(), Vm, Pm, cp, tfx, tfy, 0, nullptr, nullptr, 0, 1.0f};
    std::vector<float> depth(N), xyf(2 * N), conf(4 * N), rgbf(3 * N);
    std::vector<int> radius(N), rect(4 * N), touched(N);
    orc_raster_preprocess(&ra, depth.data(), radius.data(), xyf.data(), conf.data(), rgbf.data(), rect.data(),
                          touched.data());
    long total = 0;
    for (long g = 0; g < N; g++) total += touched[g];
    const int gx = (W + tile - 1) / tile, gy = (H + tile - 1) / tile;
    std::vector<uint64_t> keys(std::max<long>(total, 1));
    std::vector<uint32_t> vals(std::max<long>(total, 1));
    std::vector<int> ranges(2 * gx * gy);
    orc_raster_bin(&ra, depth.data(), rect.data(), touched.data(), keys.data(), vals.data(), ranges.data());
    // continuous per-Gaussian quantities in T
    std::vector<T> xy(2 * N), co(3 * N), rgb(3 * N);
    const double fx = (double)((float)W / (2.0f * tfx)), fy = (double)((float)H / (2.0f * tfy));
    const double limx = 1.3 * (double)tfx, limy = 1.3 * (double)tfy;
#pragma omp parallel for schedule(static)
    for (long g = 0; g < N; g++) {
        if (!touched[g]) continue;
        const T px = D.pos[3 * g], py = D.pos[3 * g + 1], pz = D.pos[3 * g + 2];
        const T vx = (double)Vm[0] * px + (double)Vm[4] * py + (double)Vm[8] * pz + (double)Vm[12];
        const T vy = (double)Vm[1] * px + (double)Vm[5] * py + (double)Vm[9] * pz + (double)Vm[13];
        const T vz = (double)Vm[2] * px + (double)Vm[6] * py + (double)Vm[10] * pz + (double)Vm[14];
        const T hx = (double)Pm[0] * px + (double)Pm[4] * py + (double)Pm[8] * pz + (double)Pm[12];
        const T hy = (double)Pm[1] * px + (double)Pm[5] * py + (double)Pm[9] * pz + (double)Pm[13];
        const T hw = (double)Pm[3] * px + (double)Pm[7] * py + (double)Pm[11] * pz + (double)Pm[15];
        const T pw = 1.0 / (hw + 0.0000001);
        const T q[4] = {D.rot[4 * g], D.rot[4 * g + 1], D.rot[4 * g + 2], D.rot[4 * g + 3]};
        T R[3][3];
        quat_to_mat_t(q, R);
        T s2[3];
        for (int k = 0; k < 3; k++) s2[k] = D.scale[3 * g + k] * D.scale[3 * g + k];
        T S[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) S[i][j] = R[i][0] * s2[0] * R[j][0] + R[i][1] * s2[1] * R[j][1] + R[i][2] * s2[2] * R[j][2];
        const T txtz = vx / vz, tytz = vy / vz;
        const T tx = tmin(T(limx), tmax(T(-limx), txtz)) * vz;
        const T ty = tmin(T(limy), tmax(T(-limy), tytz)) * vz;
        const T j00 = fx / vz, j02 = -(fx * tx) / (vz * vz);
        const T j11 = fy / vz, j12 = -(fy * ty) / (vz * vz);
        T Tm[2][3];
        for (int c = 0; c < 3; c++) {
            Tm[0][c] = j00 * (double)Vm[c * 4 + 0] + j02 * (double)Vm[c * 4 + 2];
            Tm[1][c] = j11 * (double)Vm[c * 4 + 1] + j12 * (double)Vm[c * 4 + 2];
        }
        T TS[2][3];
        for (int r = 0; r < 2; r++)
            for (int c = 0; c < 3; c++) TS[r][c] = Tm[r][0] * S[0][c] + Tm[r][1] * S[1][c] + Tm[r][2] * S[2][c];
        const T ca = TS[0][0] * Tm[0][0] + TS[0][1] * Tm[0][1] + TS[0][2] * Tm[0][2] + 0.3;
        const T cb = TS[0][0] * Tm[1][0] + TS[0][1] * Tm[1][1] + TS[0][2] * Tm[1][2];
        const T cc = TS[1][0] * Tm[1][0] + TS[1][1] * Tm[1][1] + TS[1][2] * Tm[1][2] + 0.3;
        const T det = ca * cc - cb * cb;
        xy[2 * g] = ((hx * pw + 1.0) * (double)W - 1.0) * 0.5;
        xy[2 * g + 1] = ((hy * pw + 1.0) * (double)H - 1.0) * 0.5;
        co[3 * g] = cc / det; co[3 * g + 1] = -cb / det; co[3 * g + 2] = ca / det;
        T dx = px - (double)cp[0], dy = py - (double)cp[1], dz = pz - (double)cp[2];
        const T dn = sqrt(dx * dx + dy * dy + dz * dz);
        dx = dx / dn; dy = dy / dn; dz = dz / dn;
        const T xx = dx * dx, yy = dy * dy, zz = dz * dz;
        const double C0 = 0.28209479177387814, C1 = 0.4886025119029199;
        const T b[16] = {T(C0), -C1 * dy, C1 * dz, -C1 * dx,
                         1.0925484305920792 * dx * dy, -1.0925484305920792 * dy * dz,
                         0.31539156525252005 * (2.0 * zz - xx - yy), -1.0925484305920792 * dx * dz,
                         0.5462742152960396 * (xx - yy),
                         -0.5900435899266435 * dy * (3.0 * xx - yy), 2.890611442640554 * dx * dy * dz,
                         -0.4570457994644658 * dy * (4.0 * zz - xx - yy),
                         0.3731763325901154 * dz * (2.0 * zz - 3.0 * xx - 3.0 * yy),
                         -0.4570457994644658 * dx * (4.0 * zz - xx - yy), 1.445305721320277 * dz * (xx - yy),
                         -0.5900435899266435 * dx * (xx - 3.0 * yy)};
        const int nb = (a->D + 1) * (a->D + 1);
        for (int ch = 0; ch < 3; ch++) {
            T r = 0.0;
            for (int k = 0; k < nb; k++) r += b[k] * D.sh[48 * g + 3 * k + ch];
            rgb[3 * g + ch] = tmax(r + 0.5, T(0.0));
        }
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < H; py++)
        for (int px = 0; px < W; px++) {
            const int t = (py / tile) * gx + px / tile;
            T Tt = 1.0, C[3] = {0.0, 0.0, 0.0};
            for (int i = ranges[2 * t]; i < ranges[2 * t + 1]; i++) {
                const uint32_t g = vals[i];
                const T dx = xy[2 * g] - (double)px, dy = xy[2 * g + 1] - (double)py;
                const T power = -0.5 * (co[3 * g] * dx * dx + co[3 * g + 2] * dy * dy) - co[3 * g + 1] * dx * dy;
                if (val(power) > 0.0) continue;
                const T alpha = tmin(T(0.99), (double)a->opac[g] * exp(power));
                if (val(alpha) < 1.0 / 255.0) continue;
                const T test_T = Tt * (1.0 - alpha);
                if (val(test_T) < 0.0001) break;
                for (int ch = 0; ch < 3; ch++) C[ch] += rgb[3 * g + ch] * alpha * Tt;
                Tt = test_T;
            }
            for (int ch = 0; ch < 3; ch++) img[(size_t)ch * W * H + (size_t)py * W + px] = C[ch] + Tt * (double)a->bg[ch];
        }
}

template <class T>
void density_t(const orc_chain_args* a, const Deformed<T>& D, T* feat, T* dens) {
    // Sigma_g + L (L = lpf I, or the cell's adaptive filter when a->ada is set)
    const long N = a->N;
    std::vector<T> Sg(6 * N), mu(3 * N);
    std::vector<char> need(N, 0);
    for (long v = 0; v < a->nv; v++) {
        const int c = a->valid[v];
        const int beg = c == 0 ? 0 : a->prefix[c - 1], end = a->prefix[c];
        for (int li = beg; li < end; li++) need[a->list[li]] = 1;
    }
#pragma omp parallel for schedule(static)
    for (long g = 0; g < N; g++) {
        if (!need[g]) continue;
        T q[4] = {D.rot[4 * g], D.rot[4 * g + 1], D.rot[4 * g + 2], D.rot[4 * g + 3]};
        normalize4(q);
        T R[3][3];
        quat_to_mat_t(q, R);
        T S[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                T acc = 0.0;
                for (int k = 0; k < 3; k++) acc += R[i][k] * D.scale[3 * g + k] * D.scale[3 * g + k] * R[j][k];
                S[i][j] = acc;
            }
        Sg[6 * g] = S[0][0]; Sg[6 * g + 1] = S[1][1]; Sg[6 * g + 2] = S[2][2];
        Sg[6 * g + 3] = S[0][1]; Sg[6 * g + 4] = S[0][2]; Sg[6 * g + 5] = S[1][2];
        for (int k = 0; k < 3; k++) mu[3 * g + k] = D.pos[3 * g + k];
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (long v = 0; v < a->nv; v++) {
        const int c = a->valid[v];
        const int beg = c == 0 ? 0 : a->prefix[c - 1], end = a->prefix[c];
        double L[6] = {(double)a->lpf, (double)a->lpf, (double)a->lpf, 0.0, 0.0, 0.0};
        if (a->ada)
            for (int k = 0; k < 6; k++) L[k] = (double)a->ada[6 * v + k];
        std::vector<T> Qc(6 * (size_t)(end - beg));
        for (int li = beg; li < end; li++) {
            const T* sg = &Sg[6 * (size_t)a->list[li]];
            T S[3][3] = {{sg[0] + L[0], sg[3] + L[3], sg[4] + L[4]},
                         {sg[3] + L[3], sg[1] + L[1], sg[5] + L[5]},
                         {sg[4] + L[4], sg[5] + L[5], sg[2] + L[2]}};
            T I[3][3];
            inverse3(S, I);
            T* q = &Qc[6 * (size_t)(li - beg)];
            q[0] = I[0][0]; q[1] = I[1][1]; q[2] = I[2][2]; q[3] = I[0][1]; q[4] = I[0][2]; q[5] = I[1][2];
        }
        for (int s = 0; s < 64; s++) {
            const long si = v * 64 + s;
            T acc[48], ao = 0.0;
            for (int k = 0; k < 48; k++) acc[k] = 0.0;
            for (int li = beg; li < end; li++) {
                const int g = a->list[li];
                const T d0 = D.samples[3 * si] - mu[3 * g], d1 = D.samples[3 * si + 1] - mu[3 * g + 1],
                        d2 = D.samples[3 * si + 2] - mu[3 * g + 2];
                const T* q = &Qc[6 * (size_t)(li - beg)];
                const T qf = q[0] * d0 * d0 + q[1] * d1 * d1 + q[2] * d2 * d2 +
                             2.0 * (q[3] * d0 * d1 + q[4] * d0 * d2 + q[5] * d1 * d2);
                const T al = D.opac[g] * exp(-0.5 * qf);
                ao += al;
                for (int k = 0; k < 48; k++) acc[k] += al * D.sh[48 * g + k];
            }
            for (int k = 0; k < 48; k++) feat[48 * si + k] = acc[k];
            dens[si] = ao;
        }
    }
}

template <class T>
void chain_t(const orc_chain_args* a, const T* Rn, const T* tn, const T* q0, const T* s0, T* images, T* feat,
             T* dens, Deformed<T>* keep) {
    Deformed<T> D;
    deform_t(a, Rn, tn, q0, s0, D);
    const size_t img = 3 * (size_t)a->W * a->H;
    for (int v = 0; v < a->V; v++) render_t(a, v, D, images + v * img);
    if (a->nv > 0) density_t(a, D, feat, dens);
    if (keep) *keep = D;
}

}  // namespace

extern "C" {

// Forward of the chain in FP64 at (Rn (M*9, row-major), tn (M*3), q0 (N*4),
// s0 (N*3)) and, when the direction (dR, dt, dq0, ds0) is given (non-NULL), the
// exact directional derivatives of every output (d_images, d_feat, d_dens).
// Outputs: images V*3*H*W, feat S*48, dens S (S = 64 nv); optional deformed
// state pos (N*3), rot (N*4), scale (N*3), sh (N*48), samples (S*3).
void orc_chain_jvp(const orc_chain_args* a, const double* Rn, const double* tn, const double* q0, const double* s0,
                   const double* dR, const double* dt, const double* dq0, const double* ds0, double* images,
                   double* feat, double* dens, double* d_images, double* d_feat, double* d_dens, double* st_pos,
                   double* st_rot, double* st_scale, double* st_sh, double* st_samples) {
    const long N = a->N, M = a->M, S = 64 * a->nv;
    const size_t nimg = (size_t)a->V * 3 * a->W * a->H;
    std::vector<Dual> R(9 * M), t(3 * M), q(4 * N), s(3 * N);
    for (long i = 0; i < 9 * M; i++) R[i] = Dual(Rn[i], dR ? dR[i] : 0.0);
    for (long i = 0; i < 3 * M; i++) t[i] = Dual(tn[i], dt ? dt[i] : 0.0);
    for (long i = 0; i < 4 * N; i++) q[i] = Dual(q0[i], dq0 ? dq0[i] : 0.0);
    for (long i = 0; i < 3 * N; i++) s[i] = Dual(s0[i], ds0 ? ds0[i] : 0.0);
    std::vector<Dual> im(nimg), fe(48 * S), de(S);
    Deformed<Dual> D;
    chain_t<Dual>(a, R.data(), t.data(), q.data(), s.data(), im.data(), fe.data(), de.data(), &D);
    for (size_t i = 0; i < nimg; i++) {
        images[i] = im[i].v;
        if (d_images) d_images[i] = im[i].d;
    }
    for (long i = 0; i < 48 * S; i++) {
        if (feat) feat[i] = fe[i].v;
        if (d_feat) d_feat[i] = fe[i].d;
    }
    for (long i = 0; i < S; i++) {
        if (dens) dens[i] = de[i].v;
        if (d_dens) d_dens[i] = de[i].d;
    }
    if (st_pos) for (long i = 0; i < 3 * N; i++) st_pos[i] = D.pos[i].v;
    if (st_rot) for (long i = 0; i < 4 * N; i++) st_rot[i] = D.rot[i].v;
    if (st_scale) for (long i = 0; i < 3 * N; i++) st_scale[i] = D.scale[i].v;
    if (st_sh) for (long i = 0; i < 48 * N; i++) st_sh[i] = D.sh[i].v;
    if (st_samples) for (long i = 0; i < 3 * S; i++) st_samples[i] = D.samples[i].v;
}

// The same forward in plain FP64 (for central differences).
void orc_chain_forward(const orc_chain_args* a, const double* Rn, const double* tn, const double* q0,
                       const double* s0, double* images, double* feat, double* dens) {
    chain_t<double>(a, Rn, tn, q0, s0, images, feat, dens, nullptr);
}

}  // extern "C"

extern "C" {

// The density sample alone (the formula of arapgs_density_sample) over its own
// Gaussian attributes and sample positions, with forward-mode derivatives
// along (dpos, drot, dscale, dopac, dsh, dsamples) (each may be NULL = 0);
// ada: per-cell adaptive filters (nv*6) or NULL for lpf I.
void orc_density_jvp(long nv, const int* valid, const int* prefix, const int* list, long N, const float* samples,
                     const float* pos, const float* rot, const float* scale, const float* opac, const float* sh,
                     float lpf, const double* dpos, const double* drot, const double* dscale, const double* dopac,
                     const double* dsh, const double* dsamples, double* feat, double* dens, double* d_feat,
                     double* d_dens, const float* ada) {
    orc_chain_args a{};
    a.ada = ada;
    a.N = N;
    a.nv = nv;
    a.valid = valid;
    a.prefix = prefix;
    a.list = list;
    a.lpf = lpf;
    a.opac = opac;
    const long S = 64 * nv;
    Deformed<Dual> D;
    auto fill = [](std::vector<Dual>& out, const float* v, const double* dv, long n) {
        out.resize(n);
        for (long i = 0; i < n; i++) out[i] = Dual((double)v[i], dv ? dv[i] : 0.0);
    };
    fill(D.pos, pos, dpos, 3 * N);
    fill(D.rot, rot, drot, 4 * N);
    fill(D.scale, scale, dscale, 3 * N);
    fill(D.opac, opac, dopac, N);
    fill(D.sh, sh, dsh, 48 * N);
    fill(D.samples, samples, dsamples, 3 * S);
    std::vector<Dual> fe(48 * S), de(S);
    density_t(&a, D, fe.data(), de.data());
    for (long i = 0; i < 48 * S; i++) {
        feat[i] = fe[i].v;
        d_feat[i] = fe[i].d;
    }
    for (long i = 0; i < S; i++) {
        dens[i] = de[i].v;
        d_dens[i] = de[i].d;
    }
}

}  // extern "C"
