// Device-side math shared by the gfx950 kernels: quaternions, the FP64 polar
// decomposition of a 3x3, the real SH basis (degree <= 3, the 3DGS sign
// convention) and SH rotation by resampling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace arapgs {

struct f3 { float x, y, z; };

__device__ __forceinline__ f3 ld_f3(const float* p) { return f3{p[0], p[1], p[2]}; }

// ---------------------------------------------------------------- quaternions
// (w,x,y,z) unit quaternion -> row-major rotation matrix (textbook formula, as
// Eigen::Quaternion::toRotationMatrix, used at GaussianView.cpp:3150,4655).
__device__ __forceinline__ void quat_to_mat(float w, float x, float y, float z, float R[3][3]) {
#pragma clang fp contract(off)
    const float tx = 2.0f * x, ty = 2.0f * y, tz = 2.0f * z;
    const float twx = tx * w, twy = ty * w, twz = tz * w;
    const float txx = tx * x, txy = ty * x, txz = tz * x;
    const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
    R[0][0] = 1.0f - (tyy + tzz); R[0][1] = txy - twz;          R[0][2] = txz + twy;
    R[1][0] = txy + twz;          R[1][1] = 1.0f - (txx + tzz); R[1][2] = tyz - twx;
    R[2][0] = txz - twy;          R[2][1] = tyz + twx;          R[2][2] = 1.0f - (txx + tyy);
}

__device__ __forceinline__ void quat_normalize(float q[4]) {
#pragma clang fp contract(off)
    const float n = __builtin_sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (n > 0.0f) { q[0] /= n; q[1] /= n; q[2] /= n; q[3] /= n; }
}

// Shoemake's matrix -> quaternion (SIGGRAPH 1985), float, out (w,x,y,z).
// Largest-diagonal branch of mat_to_quat for a fixed pivot I (static indices,
// so the matrix stays in registers).
template <int I>
__device__ __forceinline__ void mat_to_quat_pivot(const float m[3][3], float q[4]) {
#pragma clang fp contract(off)
    constexpr int J = (I + 1) % 3, K = (J + 1) % 3;
    float s = __builtin_sqrtf(m[I][I] - m[J][J] - m[K][K] + 1.0f);
    float v[3];
    v[I] = 0.5f * s;
    s = 0.5f / s;
    q[0] = (m[K][J] - m[J][K]) * s;
    v[J] = (m[J][I] + m[I][J]) * s;
    v[K] = (m[K][I] + m[I][K]) * s;
    q[1] = v[0]; q[2] = v[1]; q[3] = v[2];
}

__device__ __forceinline__ void mat_to_quat(const float m[3][3], float q[4]) {
#pragma clang fp contract(off)
    const float t = m[0][0] + m[1][1] + m[2][2];
    if (t > 0.0f) {
        float s = __builtin_sqrtf(t + 1.0f);
        q[0] = 0.5f * s;
        s = 0.5f / s;
        q[1] = (m[2][1] - m[1][2]) * s;
        q[2] = (m[0][2] - m[2][0]) * s;
        q[3] = (m[1][0] - m[0][1]) * s;
    } else {
        int i = 0;
        if (m[1][1] > m[0][0]) i = 1;
        if (m[2][2] > (i == 1 ? m[1][1] : m[0][0])) i = 2;
        if (i == 0) mat_to_quat_pivot<0>(m, q);
        else if (i == 1) mat_to_quat_pivot<1>(m, q);
        else mat_to_quat_pivot<2>(m, q);
    }
}

// a * b, (w,x,y,z)
__device__ __forceinline__ void quat_mul(const float a[4], const float b[4], float o[4]) {
    o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    o[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    o[2] = a[0] * b[2] + a[2] * b[0] + a[3] * b[1] - a[1] * b[3];
    o[3] = a[0] * b[3] + a[3] * b[0] + a[1] * b[2] - a[2] * b[1];
}

// ---------------------------------------------------------------- polar (FP64)
// M = R S via a one-sided Jacobi SVD in double: column pairs of A = M V are
// rotated until orthogonal, then A = U diag(sig); R = U V^T, S = V diag(sig) V^T.
// Only diag(S) is returned (the six-point update needs nothing else).
__device__ __forceinline__ void polar3_fp64(const double Min[3][3], double R[3][3], double Sdiag[3]) {
    double A[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) { A[i][j] = Min[i][j]; V[i][j] = (i == j) ? 1.0 : 0.0; }
    for (int sweep = 0; sweep < 12; sweep++) {
        double off = 0.0;
#pragma unroll
        for (int pr = 0; pr < 3; pr++) {
            const int p = pr == 2 ? 1 : 0;
            const int q = pr == 0 ? 1 : 2;
            double a = 0.0, b = 0.0, g = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                a = fma(A[i][p], A[i][p], a);
                b = fma(A[i][q], A[i][q], b);
                g = fma(A[i][p], A[i][q], g);
            }
            const double ab = a * b;
            if (!(ab > 0.0) || g == 0.0) continue;
            const double rel2 = (g * g) / ab;
            off = fmax(off, rel2);
            if (rel2 < 1e-32) continue;
            const double zeta = (b - a) / (2.0 * g);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(fma(zeta, zeta, 1.0)));
            const double c = rsqrt(fma(t, t, 1.0)), s = c * t;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const double ap = A[i][p], aq = A[i][q];
                A[i][p] = c * ap - s * aq;
                A[i][q] = s * ap + c * aq;
                const double vp = V[i][p], vq = V[i][q];
                V[i][p] = c * vp - s * vq;
                V[i][q] = s * vp + c * vq;
            }
        }
        if (off < 1e-30) break;
    }
    double sig[3], U[3][3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        sig[j] = sqrt(A[0][j] * A[0][j] + A[1][j] * A[1][j] + A[2][j] * A[2][j]);
        const double inv = sig[j] > 0.0 ? 1.0 / sig[j] : 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) U[i][j] = A[i][j] * inv;
    }
    // rank-deficient input: complete U by a cross product (polar not unique)
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (sig[j] > 1e-300) continue;
        const int a = (j + 1) % 3, b = (j + 2) % 3;
        double c0 = U[1][a] * U[2][b] - U[2][a] * U[1][b];
        double c1 = U[2][a] * U[0][b] - U[0][a] * U[2][b];
        double c2 = U[0][a] * U[1][b] - U[1][a] * U[0][b];
        double n = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        if (n == 0.0) { c0 = (j == 0); c1 = (j == 1); c2 = (j == 2); n = 1.0; }
        U[0][j] = c0 / n; U[1][j] = c1 / n; U[2][j] = c2 / n;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int k = 0; k < 3; k++) R[i][k] = U[i][0] * V[k][0] + U[i][1] * V[k][1] + U[i][2] * V[k][2];
        Sdiag[i] = V[i][0] * sig[0] * V[i][0] + V[i][1] * sig[1] * V[i][1] + V[i][2] * sig[2] * V[i][2];
    }
}

// ---------------------------------------------------------------- SH basis
// Real SH basis values of band l at unit direction d, in the 3DGS convention
// (the signs of the standard 3DGS evaluator).
#define ARAPGS_SH_C1 0.4886025119029199f
__device__ __forceinline__ void sh_band2(float x, float y, float z, float b[5]) {
    const float xx = x * x, yy = y * y, zz = z * z;
    b[0] = 1.0925484305920792f * x * y;
    b[1] = -1.0925484305920792f * y * z;
    b[2] = 0.31539156525252005f * (2.0f * zz - xx - yy);
    b[3] = -1.0925484305920792f * x * z;
    b[4] = 0.5462742152960396f * (xx - yy);
}
__device__ __forceinline__ void sh_band3(float x, float y, float z, float b[7]) {
    const float xx = x * x, yy = y * y, zz = z * z;
    b[0] = -0.5900435899266435f * y * (3.0f * xx - yy);
    b[1] = 2.890611442640554f * x * y * z;
    b[2] = -0.4570457994644658f * y * (4.0f * zz - xx - yy);
    b[3] = 0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    b[4] = -0.4570457994644658f * x * (4.0f * zz - xx - yy);
    b[5] = 1.445305721320277f * z * (xx - yy);
    b[6] = -0.5900435899266435f * x * (xx - 3.0f * yy);
}

// ---------------------------------------------------------------- SH rotation
// Rotation of a band-limited function f(d) = sum_m c_m Y_m(d) into
// f'(d) = f(R^T d) by resampling: for band l pick 2l+1 fixed directions d_k
// (chosen at library init so the basis matrix Yl[k][m] = Y_m(d_k) is well
// conditioned), evaluate F_k = f(R^T d_k) and solve Yl c' = F with the
// precomputed inverse. Band 1 is the (y,z,x)-permutation conjugation of R with
// the 3DGS signs folded in. The result is the same linear map as the
// reference's recurrence with its odd-index sign wrapper (helper.cpp:938-1075,
// GaussianView.cpp:3134-3156, cudakdtree.cu:150-199).
struct ShRotTables {
    float dir2[5][3];
    float inv2[5][5];   // inv2[m][k]
    float dir3[7][3];
    float inv3[7][7];
};

// SH rotation f'(d) = f(R^T d) band by band (R row-major, coefficients
// coef-major, channel-minor: band l holds coefficients l^2 .. (l+1)^2 - 1, so
// band 1 = floats 3..11, band 2 = 12..26, band 3 = 27..47; band 0 is invariant).
// Band 1: basis (-y, z, -x) * C1, so c' = M c with M[m][n] = s_m s_n R[a_m][a_n],
// a = (y, z, x), s = (-1, 1, -1). Bands 2 and 3 resample: evaluate the band at
// the 2l+1 rotated table directions R^T d_k and map the values back to
// coefficients with the table's inverse basis matrix.
__device__ __forceinline__ void sh_rotate_band1(const float R[3][3], float (&b)[9]) {
    const int ax[3] = {1, 2, 0};
    const float sg[3] = {-1.0f, 1.0f, -1.0f};
    float o[3][3];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            float acc = 0.0f;
#pragma unroll
            for (int n = 0; n < 3; n++) acc += sg[m] * sg[n] * R[ax[m]][ax[n]] * b[n * 3 + ch];
            o[m][ch] = acc;
        }
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int ch = 0; ch < 3; ch++) b[m * 3 + ch] = o[m][ch];
}

__device__ __forceinline__ void sh_rotate_band2(const ShRotTables& T, const float R[3][3], float (&b)[15]) {
    float F[5][3];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const float dx = T.dir2[k][0], dy = T.dir2[k][1], dz = T.dir2[k][2];
        const float rx = R[0][0] * dx + R[1][0] * dy + R[2][0] * dz;
        const float ry = R[0][1] * dx + R[1][1] * dy + R[2][1] * dz;
        const float rz = R[0][2] * dx + R[1][2] * dy + R[2][2] * dz;
        float y[5];
        sh_band2(rx, ry, rz, y);
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            float acc = 0.0f;
#pragma unroll
            for (int m = 0; m < 5; m++) acc += y[m] * b[m * 3 + ch];
            F[k][ch] = acc;
        }
    }
#pragma unroll
    for (int m = 0; m < 5; m++)
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 5; k++) acc += T.inv2[m][k] * F[k][ch];
            b[m * 3 + ch] = acc;
        }
}

__device__ __forceinline__ void sh_rotate_band3(const ShRotTables& T, const float R[3][3], float (&b)[21]) {
    float F[7][3];
#pragma unroll
    for (int k = 0; k < 7; k++) {
        const float dx = T.dir3[k][0], dy = T.dir3[k][1], dz = T.dir3[k][2];
        const float rx = R[0][0] * dx + R[1][0] * dy + R[2][0] * dz;
        const float ry = R[0][1] * dx + R[1][1] * dy + R[2][1] * dz;
        const float rz = R[0][2] * dx + R[1][2] * dy + R[2][2] * dz;
        float y[7];
        sh_band3(rx, ry, rz, y);
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            float acc = 0.0f;
#pragma unroll
            for (int m = 0; m < 7; m++) acc += y[m] * b[m * 3 + ch];
            F[k][ch] = acc;
        }
    }
#pragma unroll
    for (int m = 0; m < 7; m++)
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 7; k++) acc += T.inv3[m][k] * F[k][ch];
            b[m * 3 + ch] = acc;
        }
}

// Rotates one 48-float SH set in global memory in place: twelve 16-byte loads
// and stores (sh must be 16-byte aligned), bands 1-3 rotated in registers.
__device__ __forceinline__ void sh_rotate_f4(const ShRotTables& T, const float R[3][3], float* __restrict__ sh) {
    float c[48];
    const float4* s4 = reinterpret_cast<const float4*>(sh);
#pragma unroll
    for (int i = 0; i < 12; i++) {
        const float4 v = s4[i];
        c[4 * i] = v.x; c[4 * i + 1] = v.y; c[4 * i + 2] = v.z; c[4 * i + 3] = v.w;
    }
    float b1[9], b2[15], b3[21];
#pragma unroll
    for (int i = 0; i < 9; i++) b1[i] = c[3 + i];
#pragma unroll
    for (int i = 0; i < 15; i++) b2[i] = c[12 + i];
#pragma unroll
    for (int i = 0; i < 21; i++) b3[i] = c[27 + i];
    sh_rotate_band1(R, b1);
    sh_rotate_band2(T, R, b2);
    sh_rotate_band3(T, R, b3);
#pragma unroll
    for (int i = 0; i < 9; i++) c[3 + i] = b1[i];
#pragma unroll
    for (int i = 0; i < 15; i++) c[12 + i] = b2[i];
#pragma unroll
    for (int i = 0; i < 21; i++) c[27 + i] = b3[i];
    float4* o4 = reinterpret_cast<float4*>(sh);
#pragma unroll
    for (int i = 0; i < 12; i++) o4[i] = make_float4(c[4 * i], c[4 * i + 1], c[4 * i + 2], c[4 * i + 3]);
}

}  // namespace arapgs
