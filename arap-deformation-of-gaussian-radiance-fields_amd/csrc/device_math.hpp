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

// ---------------------------------------------------------------- adjoints
// Reverse-mode companions of the quaternion helpers above: each adds the
// gradient of its inputs given the gradient of its output.

// R = quat_to_mat(q) for any (w,x,y,z), not necessarily unit: dq += (dR/dq)^T dR.
__device__ __forceinline__ void quat_to_mat_backward(const float q[4], const float dR[3][3], float dq[4]) {
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    dq[0] += 2.0f * (-z * dR[0][1] + y * dR[0][2] + z * dR[1][0] - x * dR[1][2] - y * dR[2][0] + x * dR[2][1]);
    dq[1] += 2.0f * (y * dR[0][1] + z * dR[0][2] + y * dR[1][0] - 2.0f * x * dR[1][1] - w * dR[1][2] +
                     z * dR[2][0] + w * dR[2][1] - 2.0f * x * dR[2][2]);
    dq[2] += 2.0f * (-2.0f * y * dR[0][0] + x * dR[0][1] + w * dR[0][2] + x * dR[1][0] + z * dR[1][2] -
                     w * dR[2][0] + z * dR[2][1] - 2.0f * y * dR[2][2]);
    dq[3] += 2.0f * (-2.0f * z * dR[0][0] - w * dR[0][1] + x * dR[0][2] + w * dR[1][0] - 2.0f * z * dR[1][1] +
                     y * dR[1][2] + x * dR[2][0] + y * dR[2][1]);
}

// u = q / |q|: dq = (du - u <u, du>) / |q| (q is the unnormalised input).
__device__ __forceinline__ void quat_normalize_backward(const float q[4], const float du[4], float dq[4]) {
    const float n = __builtin_sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (!(n > 0.0f)) return;
    const float u[4] = {q[0] / n, q[1] / n, q[2] / n, q[3] / n};
    const float d = u[0] * du[0] + u[1] * du[1] + u[2] * du[2] + u[3] * du[3];
#pragma unroll
    for (int i = 0; i < 4; i++) dq[i] += (du[i] - u[i] * d) / n;
}

// o = a * b (Hamilton product, (w,x,y,z)): gradients of both factors.
__device__ __forceinline__ void quat_mul_backward(const float a[4], const float b[4], const float d[4], float da[4],
                                                  float db[4]) {
    da[0] += d[0] * b[0] + d[1] * b[1] + d[2] * b[2] + d[3] * b[3];
    da[1] += -d[0] * b[1] + d[1] * b[0] - d[2] * b[3] + d[3] * b[2];
    da[2] += -d[0] * b[2] + d[1] * b[3] + d[2] * b[0] - d[3] * b[1];
    da[3] += -d[0] * b[3] - d[1] * b[2] + d[2] * b[1] + d[3] * b[0];
    db[0] += d[0] * a[0] + d[1] * a[1] + d[2] * a[2] + d[3] * a[3];
    db[1] += -d[0] * a[1] + d[1] * a[0] + d[2] * a[3] - d[3] * a[2];
    db[2] += -d[0] * a[2] - d[1] * a[3] + d[2] * a[0] + d[3] * a[1];
    db[3] += -d[0] * a[3] + d[1] * a[2] - d[2] * a[1] + d[3] * a[0];
}

// q = mat_to_quat(m) (the branch the forward takes): dm += (dq/dm)^T dq.
template <int I>
__device__ __forceinline__ void mat_to_quat_pivot_backward(const float m[3][3], const float dq[4], float dm[3][3]) {
    constexpr int J = (I + 1) % 3, K = (J + 1) % 3;
    const float s = __builtin_sqrtf(m[I][I] - m[J][J] - m[K][K] + 1.0f);
    const float inv = 0.5f / s;
    const float dvI = dq[1 + I], dvJ = dq[1 + J], dvK = dq[1 + K];
    dm[K][J] += dq[0] * inv; dm[J][K] -= dq[0] * inv;
    dm[J][I] += dvJ * inv; dm[I][J] += dvJ * inv;
    dm[K][I] += dvK * inv; dm[I][K] += dvK * inv;
    const float dinv = dq[0] * (m[K][J] - m[J][K]) + dvJ * (m[J][I] + m[I][J]) + dvK * (m[K][I] + m[I][K]);
    const float ds = 0.5f * dvI - dinv * 0.5f / (s * s);
    const float dt = ds * 0.5f / s;
    dm[I][I] += dt; dm[J][J] -= dt; dm[K][K] -= dt;
}

__device__ __forceinline__ void mat_to_quat_backward(const float m[3][3], const float dq[4], float dm[3][3]) {
    const float t = m[0][0] + m[1][1] + m[2][2];
    if (t > 0.0f) {
        const float s = __builtin_sqrtf(t + 1.0f);
        const float inv = 0.5f / s;
        dm[2][1] += dq[1] * inv; dm[1][2] -= dq[1] * inv;
        dm[0][2] += dq[2] * inv; dm[2][0] -= dq[2] * inv;
        dm[1][0] += dq[3] * inv; dm[0][1] -= dq[3] * inv;
        const float dinv = dq[1] * (m[2][1] - m[1][2]) + dq[2] * (m[0][2] - m[2][0]) + dq[3] * (m[1][0] - m[0][1]);
        const float ds = 0.5f * dq[0] - dinv * 0.5f / (s * s);
        const float dt = ds * 0.5f / s;
        dm[0][0] += dt; dm[1][1] += dt; dm[2][2] += dt;
    } else {
        int i = 0;
        if (m[1][1] > m[0][0]) i = 1;
        if (m[2][2] > (i == 1 ? m[1][1] : m[0][0])) i = 2;
        if (i == 0) mat_to_quat_pivot_backward<0>(m, dq, dm);
        else if (i == 1) mat_to_quat_pivot_backward<1>(m, dq, dm);
        else mat_to_quat_pivot_backward<2>(m, dq, dm);
    }
}

// Sum of v over the 64 lanes of a wave (every lane gets the total).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ---------------------------------------------------------------- polar (FP64)
// M = R S via a one-sided Jacobi SVD in double: column pairs of A = M V are
// rotated until orthogonal, then A = U diag(sig); R = U V^T, S = V diag(sig) V^T.
// Only diag(S) is returned (the six-point update needs nothing else).
// The same factorisation with V and the singular values sig exposed
// (M = U diag(sig) V^T, R = U V^T, S = V diag(sig) V^T).
__device__ __forceinline__ void polar3_fp64_svd(const double Min[3][3], double R[3][3], double V[3][3],
                                                double sig[3]) {
    double A[3][3];
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
    double U[3][3];
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
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) R[i][k] = U[i][0] * V[k][0] + U[i][1] * V[k][1] + U[i][2] * V[k][2];
}

__device__ __forceinline__ void polar3_fp64(const double Min[3][3], double R[3][3], double Sdiag[3]) {
    double V[3][3], sig[3];
    polar3_fp64_svd(Min, R, V, sig);
#pragma unroll
    for (int i = 0; i < 3; i++)
        Sdiag[i] = V[i][0] * sig[0] * V[i][0] + V[i][1] * sig[1] * V[i][1] + V[i][2] * sig[2] * V[i][2];
}

// Reverse mode of M -> (R, diag S) for M = R S (S symmetric positive
// semi-definite): with dR = R Omega and S Omega + Omega S = A - A^T for
// A = R^T dM, the adjoint is dM = R (diag(dSd) + V Ht V^T), where
// G = R^T dR_bar - diag(dSd) S, Gt = V^T G V and
// Ht_ij = (Gt_ij - Gt_ji) / (sig_i + sig_j).
__device__ __forceinline__ void polar3_fp64_backward(const double R[3][3], const double V[3][3], const double sig[3],
                                                     const double dR[3][3], const double dSd[3], double dM[3][3]) {
    double S[3][3], G[3][3], Gt[3][3], H[3][3], tmp[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            S[i][j] = V[i][0] * sig[0] * V[j][0] + V[i][1] * sig[1] * V[j][1] + V[i][2] * sig[2] * V[j][2];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            G[i][j] = R[0][i] * dR[0][j] + R[1][i] * dR[1][j] + R[2][i] * dR[2][j] - dSd[i] * S[i][j];
    // Gt = V^T G V
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) tmp[i][j] = G[i][0] * V[0][j] + G[i][1] * V[1][j] + G[i][2] * V[2][j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Gt[i][j] = V[0][i] * tmp[0][j] + V[1][i] * tmp[1][j] + V[2][i] * tmp[2][j];
    // Ht (in place in tmp), then H = V Ht V^T
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double den = sig[i] + sig[j];
            tmp[i][j] = den > 0.0 ? (Gt[i][j] - Gt[j][i]) / den : 0.0;
        }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) G[i][j] = tmp[i][0] * V[j][0] + tmp[i][1] * V[j][1] + tmp[i][2] * V[j][2];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) H[i][j] = V[i][0] * G[0][j] + V[i][1] * G[1][j] + V[i][2] * G[2][j];
#pragma unroll
    for (int i = 0; i < 3; i++) H[i][i] += dSd[i];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) dM[i][j] = R[i][0] * H[0][j] + R[i][1] * H[1][j] + R[i][2] * H[2][j];
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

// Real SH basis of degree <= 3 at unit direction (x, y, z), in the 3DGS sign
// convention (the renderer's colour evaluator, sh_band2 / sh_band3 below), and its gradient with respect to (x, y, z).
__device__ __forceinline__ void sh_basis_and_grad(float x, float y, float z, float b[16], float db[16][3]) {
    const float C0 = 0.28209479177387814f, C1 = 0.4886025119029199f;
    const float B1 = 1.0925484305920792f, B2 = 0.31539156525252005f, B3 = 0.5462742152960396f;
    const float E0 = 0.5900435899266435f, E1 = 2.890611442640554f, E2 = 0.4570457994644658f,
                E3 = 0.3731763325901154f, E4 = 1.445305721320277f;
    const float xx = x * x, yy = y * y, zz = z * z;
    b[0] = C0; db[0][0] = 0.0f; db[0][1] = 0.0f; db[0][2] = 0.0f;
    b[1] = -C1 * y; db[1][0] = 0.0f; db[1][1] = -C1; db[1][2] = 0.0f;
    b[2] = C1 * z; db[2][0] = 0.0f; db[2][1] = 0.0f; db[2][2] = C1;
    b[3] = -C1 * x; db[3][0] = -C1; db[3][1] = 0.0f; db[3][2] = 0.0f;
    b[4] = B1 * x * y; db[4][0] = B1 * y; db[4][1] = B1 * x; db[4][2] = 0.0f;
    b[5] = -B1 * y * z; db[5][0] = 0.0f; db[5][1] = -B1 * z; db[5][2] = -B1 * y;
    b[6] = B2 * (2.0f * zz - xx - yy); db[6][0] = -2.0f * B2 * x; db[6][1] = -2.0f * B2 * y; db[6][2] = 4.0f * B2 * z;
    b[7] = -B1 * x * z; db[7][0] = -B1 * z; db[7][1] = 0.0f; db[7][2] = -B1 * x;
    b[8] = B3 * (xx - yy); db[8][0] = 2.0f * B3 * x; db[8][1] = -2.0f * B3 * y; db[8][2] = 0.0f;
    b[9] = -E0 * y * (3.0f * xx - yy);
    db[9][0] = -E0 * 6.0f * x * y; db[9][1] = -E0 * (3.0f * xx - 3.0f * yy); db[9][2] = 0.0f;
    b[10] = E1 * x * y * z; db[10][0] = E1 * y * z; db[10][1] = E1 * x * z; db[10][2] = E1 * x * y;
    b[11] = -E2 * y * (4.0f * zz - xx - yy);
    db[11][0] = E2 * 2.0f * x * y; db[11][1] = -E2 * (4.0f * zz - xx - 3.0f * yy); db[11][2] = -E2 * 8.0f * y * z;
    b[12] = E3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    db[12][0] = -E3 * 6.0f * x * z; db[12][1] = -E3 * 6.0f * y * z; db[12][2] = E3 * (6.0f * zz - 3.0f * xx - 3.0f * yy);
    b[13] = -E2 * x * (4.0f * zz - xx - yy);
    db[13][0] = -E2 * (4.0f * zz - 3.0f * xx - yy); db[13][1] = E2 * 2.0f * x * y; db[13][2] = -E2 * 8.0f * x * z;
    b[14] = E4 * z * (xx - yy); db[14][0] = E4 * 2.0f * x * z; db[14][1] = -E4 * 2.0f * y * z; db[14][2] = E4 * (xx - yy);
    b[15] = -E0 * x * (xx - 3.0f * yy);
    db[15][0] = -E0 * (3.0f * xx - 3.0f * yy); db[15][1] = E0 * 6.0f * x * y; db[15][2] = 0.0f;
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
__device__ __forceinline__ void sh_rotate_regs(const ShRotTables& T, const float R[3][3], float (&c)[48]) {
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
}

__device__ __forceinline__ void sh_rotate_f4(const ShRotTables& T, const float R[3][3], float* __restrict__ sh) {
    float c[48];
    const float4* s4 = reinterpret_cast<const float4*>(sh);
#pragma unroll
    for (int i = 0; i < 12; i++) {
        const float4 v = s4[i];
        c[4 * i] = v.x; c[4 * i + 1] = v.y; c[4 * i + 2] = v.z; c[4 * i + 3] = v.w;
    }
    sh_rotate_regs(T, R, c);
    float4* o4 = reinterpret_cast<float4*>(sh);
#pragma unroll
    for (int i = 0; i < 12; i++) o4[i] = make_float4(c[4 * i], c[4 * i + 1], c[4 * i + 2], c[4 * i + 3]);
}

// Reverse mode of sh_rotate_f4's map (bands 1-3, R row-major) with respect to
// R: given the input coefficients c (48) and dL/d(rotated coefficients) dc,
// adds dL/dR. Band 1 is linear in R; bands 2 and 3 pass dc back through the
// inverse basis matrix to the resampled values F_k, then through the basis
// at r_k = R^T d_k (r_k,j = sum_i R_ij d_k,i).
__device__ __forceinline__ void sh_rotate_backward(const ShRotTables& T, const float R[3][3], const float* c,
                                                   const float* dc, float dR[3][3]) {
    {
        const int ax[3] = {1, 2, 0};
        const float sg[3] = {-1.0f, 1.0f, -1.0f};
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
            for (int n = 0; n < 3; n++) {
                float a = 0.0f;
#pragma unroll
                for (int ch = 0; ch < 3; ch++) a += dc[3 + m * 3 + ch] * c[3 + n * 3 + ch];
                dR[ax[m]][ax[n]] += sg[m] * sg[n] * a;
            }
    }
    float bb[16], dbb[16][3];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const float dx = T.dir2[k][0], dy = T.dir2[k][1], dz = T.dir2[k][2];
        const float rx = R[0][0] * dx + R[1][0] * dy + R[2][0] * dz;
        const float ry = R[0][1] * dx + R[1][1] * dy + R[2][1] * dz;
        const float rz = R[0][2] * dx + R[1][2] * dy + R[2][2] * dz;
        sh_basis_and_grad(rx, ry, rz, bb, dbb);
        float dr[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int m = 0; m < 5; m++) {
            float dy_m = 0.0f;
#pragma unroll
            for (int ch = 0; ch < 3; ch++) {
                float dF = 0.0f;
#pragma unroll
                for (int mm = 0; mm < 5; mm++) dF += T.inv2[mm][k] * dc[12 + mm * 3 + ch];
                dy_m += dF * c[12 + m * 3 + ch];
            }
            dr[0] += dy_m * dbb[4 + m][0];
            dr[1] += dy_m * dbb[4 + m][1];
            dr[2] += dy_m * dbb[4 + m][2];
        }
        const float d[3] = {dx, dy, dz};
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dR[i][j] += d[i] * dr[j];
    }
#pragma unroll
    for (int k = 0; k < 7; k++) {
        const float dx = T.dir3[k][0], dy = T.dir3[k][1], dz = T.dir3[k][2];
        const float rx = R[0][0] * dx + R[1][0] * dy + R[2][0] * dz;
        const float ry = R[0][1] * dx + R[1][1] * dy + R[2][1] * dz;
        const float rz = R[0][2] * dx + R[1][2] * dy + R[2][2] * dz;
        sh_basis_and_grad(rx, ry, rz, bb, dbb);
        float dr[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int m = 0; m < 7; m++) {
            float dy_m = 0.0f;
#pragma unroll
            for (int ch = 0; ch < 3; ch++) {
                float dF = 0.0f;
#pragma unroll
                for (int mm = 0; mm < 7; mm++) dF += T.inv3[mm][k] * dc[27 + mm * 3 + ch];
                dy_m += dF * c[27 + m * 3 + ch];
            }
            dr[0] += dy_m * dbb[9 + m][0];
            dr[1] += dy_m * dbb[9 + m][1];
            dr[2] += dy_m * dbb[9 + m][2];
        }
        const float d[3] = {dx, dy, dz};
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dR[i][j] += d[i] * dr[j];
    }
}

}  // namespace arapgs
