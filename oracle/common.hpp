// TEST INFRASTRUCTURE — the oracle.
//
// A CPU restatement (C++/OpenMP, FP64 where the reference is FP64) of the
// reference's per-frame "ARAP solve -> deform Gaussians -> render" path. It is
// the checker for the HIP library and the CPU baseline timed by bench.py.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it; the product path never calls into it.
//
// Build: `make -C oracle` (g++, no -march, so no FMA contraction — the same
// arithmetic the reference's x86-64 host build performs).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

struct V3f { float x, y, z; };
struct V3d { double x, y, z; };

static inline V3f ld3(const float* p) { return V3f{p[0], p[1], p[2]}; }
static inline void st3(float* p, V3f v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// Row-major 3x3 in double.
struct M3d { double m[3][3]; };
static inline M3d mul(const M3d& a, const M3d& b) {
    M3d c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += a.m[i][k] * b.m[k][j];
            c.m[i][j] = s;
        }
    return c;
}
static inline M3d transpose(const M3d& a) {
    M3d t;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t.m[i][j] = a.m[j][i];
    return t;
}

// Quaternion with float components, stored (w,x,y,z) like the reference's Rot.
struct Qf { float w, x, y, z; };

// Rotation matrix (row-major, float) of a unit quaternion: the textbook
// formula, matching Eigen::Quaternion::toRotationMatrix used at
// GaussianView.cpp:3150 and cudakdtree.cu:220.
static inline void quat_to_mat(const Qf& q, float R[3][3]) {
    const float tx = 2.0f * q.x, ty = 2.0f * q.y, tz = 2.0f * q.z;
    const float twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const float txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const float tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    R[0][0] = 1.0f - (tyy + tzz); R[0][1] = txy - twz;          R[0][2] = txz + twy;
    R[1][0] = txy + twz;          R[1][1] = 1.0f - (txx + tzz); R[1][2] = tyz - twx;
    R[2][0] = txz - twy;          R[2][1] = tyz + twx;          R[2][2] = 1.0f - (txx + tyy);
}

static inline Qf quat_normalized(Qf q) {
    float n = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    if (n > 0.0f) { q.w /= n; q.x /= n; q.y /= n; q.z /= n; }
    return q;
}

// Shoemake's matrix -> quaternion conversion (K. Shoemake, "Animating rotation
// with quaternion curves", SIGGRAPH 1985), in float; the method behind
// Eigen::Quaternionf(Matrix3f) used at GaussianView.cpp:3119 and helper.cpp:3175.
static inline Qf mat_to_quat(const float m[3][3]) {
    Qf q;
    const float t = m[0][0] + m[1][1] + m[2][2];
    if (t > 0.0f) {
        float s = std::sqrt(t + 1.0f);
        q.w = 0.5f * s;
        s = 0.5f / s;
        q.x = (m[2][1] - m[1][2]) * s;
        q.y = (m[0][2] - m[2][0]) * s;
        q.z = (m[1][0] - m[0][1]) * s;
    } else {
        int i = 0;
        if (m[1][1] > m[0][0]) i = 1;
        if (m[2][2] > m[i][i]) i = 2;
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        float s = std::sqrt(m[i][i] - m[j][j] - m[k][k] + 1.0f);
        float v[3];
        v[i] = 0.5f * s;
        s = 0.5f / s;
        q.w = (m[k][j] - m[j][k]) * s;
        v[j] = (m[j][i] + m[i][j]) * s;
        v[k] = (m[k][i] + m[i][k]) * s;
        q.x = v[0]; q.y = v[1]; q.z = v[2];
    }
    return q;
}

static inline Qf quat_mul(const Qf& a, const Qf& b) {
    return Qf{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
              a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
              a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
              a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x};
}
static inline Qf quat_inverse(const Qf& q) {
    const float n2 = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    return Qf{q.w / n2, -q.x / n2, -q.y / n2, -q.z / n2};
}

// Polar decomposition M = R S of a real 3x3 matrix in FP64, from a one-sided
// (Hestenes) Jacobi SVD M = U diag(sig) V^T: R = U V^T, S = V diag(sig) V^T.
// The orthogonal polar factor and S are unique for non-singular M, so this
// equals the JacobiSVD-based getOthogonalMatrixWithK (helper.cpp:429-440).
void polar_fp64(const double M[3][3], double R[3][3], double S[3][3]);

// Real SH rotation (degree <= 3) of 48 interleaved coefficients (coef*3+ch).
// `flip_odd` applies the odd-index sign convention of GaussianView.cpp:3138-3156.
void sh_rotate(const float R[3][3], float* sh48, bool flip_odd);

}  // namespace orc
