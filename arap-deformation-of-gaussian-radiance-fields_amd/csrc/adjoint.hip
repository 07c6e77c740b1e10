// Reverse mode of the deform -> density-sample -> render chain (gfx950).
//
// Forward chain, per Gaussian (rest mean mu0, rest rotation q0, rest scale
// s0, SH c0): six end points e = mu0 +- R(q0/|q0|)[:,i] (s0_i + pad) coeff
// (GetEndPoints, GaussianView.cpp:4643-4668), each blended by its K nodes
// e' = sum_j w_j (R_j (e - g_j) + t_j + g_j) (predict_mesh, Deform.hpp:230-247),
// then the six-point update (UpdateAsSixPointsWithdrawBad,
// GaussianView.cpp:3081-3166): mu' = mean of e', R, S = polar of
// M = [(e'_{i+} - e'_{i-}) / 2], q' = quat(R) / |quat(R)|,
// s'_i = S_ii / ((s0_i + pad) coeff) s0_i, c' = SH rotation of c0 by
// R(normalise(q' q0^-1)). Samples x are blended the same way. The renderer
// and the density sampler read (mu', q', s', c') and the blended samples.
//
// The kernels here take the output-side gradients (from the renderer's
// backward and the density backward below) back to the node transforms
// (R_j, t_j), q0 and s0:
//   density_backward_kernel   one workgroup per valid cell: sample gradients
//                             (owned by the cell) and per-(cell, Gaussian)
//                             atomic adds of d mu', d Q (inverse covariance)
//                             and d c'
//   six_point_backward_kernel one thread per Gaussian: density covariance
//                             gradient to (q', s'), SH rotation, quaternion,
//                             polar and centre adjoints, blend adjoint to the
//                             rest end points and the end-point adjoint to
//                             (q0, s0); writes d e' per end point
//   node_gather_kernel        one workgroup per node over the node's entries
//                             of the (end point | sample) x slot inverse
//                             index: d R_j, d t_j in FP64, no atomics
#pragma clang fp contract(off)
#include "adjoint.hpp"
#include "common.hpp"
#include "sort.hpp"

namespace arapgs {

namespace {

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__global__ void zero_f32_kernel(float* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

// Inverse covariance of a density Gaussian as the density kernel forms it in
// valid cell v: Sigma = R(q/|q|) diag(s^2) R^T + L (L = lpf I or the cell's
// adaptive filter), Q = Sigma^-1 by cofactors. Returns Q (q00, q11, q22, q01,
// q02, q12), R and Sigma for the backward.
__device__ __forceinline__ void density_inverse_cov(const float* __restrict__ rot, const float* __restrict__ scale,
                                                    int64_t g, const DensityFilter& flt, int64_t v, float Q[6],
                                                    float R[3][3], float S[3][3]) {
#pragma clang fp contract(fast)
    float q[4] = {rot[4 * g], rot[4 * g + 1], rot[4 * g + 2], rot[4 * g + 3]};
    quat_normalize(q);
    quat_to_mat(q[0], q[1], q[2], q[3], R);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float a = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float sk = scale[3 * g + k];
                a += R[i][k] * sk * sk * R[j][k];
            }
            S[i][j] = a;
        }
    density_filter_add(flt, v, S);
    const float c00 = S[1][1] * S[2][2] - S[1][2] * S[2][1];
    const float c01 = S[1][2] * S[2][0] - S[1][0] * S[2][2];
    const float c02 = S[1][0] * S[2][1] - S[1][1] * S[2][0];
    const float det = S[0][0] * c00 + S[0][1] * c01 + S[0][2] * c02;
    const float id = 1.0f / det;
    Q[0] = c00 * id;
    Q[1] = (S[0][0] * S[2][2] - S[0][2] * S[2][0]) * id;
    Q[2] = (S[0][0] * S[1][1] - S[0][1] * S[1][0]) * id;
    Q[3] = c01 * id;
    Q[4] = c02 * id;
    Q[5] = (S[0][2] * S[1][0] - S[0][0] * S[1][2]) * id;
}

// dSigma = -Q dQ Q for Q = Sigma^-1, both symmetric and packed (00, 11, 22,
// 01, 02, 12); the off-diagonal entries of dq6 and of the result count both
// symmetric entries (dL/dX01 + dL/dX10).
__device__ __forceinline__ void dinv_to_dsigma(const float Qv[6], const float dq6[6], float ds6[6]) {
    const float Qm[3][3] = {{Qv[0], Qv[3], Qv[4]}, {Qv[3], Qv[1], Qv[5]}, {Qv[4], Qv[5], Qv[2]}};
    const float dQ[3][3] = {{dq6[0], 0.5f * dq6[3], 0.5f * dq6[4]},
                            {0.5f * dq6[3], dq6[1], 0.5f * dq6[5]},
                            {0.5f * dq6[4], 0.5f * dq6[5], dq6[2]}};
    float tmp[3][3], dS[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) tmp[i][j] = dQ[i][0] * Qm[0][j] + dQ[i][1] * Qm[1][j] + dQ[i][2] * Qm[2][j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) dS[i][j] = -(Qm[i][0] * tmp[0][j] + Qm[i][1] * tmp[1][j] + Qm[i][2] * tmp[2][j]);
    ds6[0] = dS[0][0]; ds6[1] = dS[1][1]; ds6[2] = dS[2][2];
    ds6[3] = dS[0][1] + dS[1][0]; ds6[4] = dS[0][2] + dS[2][0]; ds6[5] = dS[1][2] + dS[2][1];
}

// a_gs = o_g exp(-q/2), q = d^T Q d, d = x_s - mu_g; feat_s = sum_g a_gs c_g,
// dens_s = sum_g a_gs. With da = dL/dfeat_s . c_g + dL/ddens_s:
//   dL/dc_g += a dL/dfeat_s,  dL/dq = -a da / 2,  dL/dQ += dL/dq d d^T,
//   dL/dx_s += 2 dL/dq Q d,   dL/dmu_g -= 2 dL/dq Q d.
// Threads own Gaussians of the cell's list (64 at a time, parameters and SH
// in registers) and walk the cell's 64 samples, whose positions and output
// gradients sit in LDS (every lane reads the same sample: broadcast). The
// sample gradient of each step is a wave sum. gdens is component-major
// (component k of Gaussian g at k * gstride + g): d mu (3), d Q (q00, q11,
// q22, then q01, q02, q12 each counted for both symmetric entries; with an
// adaptive filter, whose Q differs per cell, d Sigma = -Q dQ Q in the same
// packing instead), d c (48), d o (1).
__global__ void __launch_bounds__(64) density_backward_kernel(int64_t nv, const int* __restrict__ valid,
                                                              const int64_t* __restrict__ prefix,
                                                              const int* __restrict__ list,
                                                              const float* __restrict__ samples,
                                                              const float* __restrict__ pos,
                                                              const float* __restrict__ rot,
                                                              const float* __restrict__ scale,
                                                              const float* __restrict__ opacity,
                                                              const float* __restrict__ sh, DensityFilter flt,
                                                              const float* __restrict__ dfeat,
                                                              const float* __restrict__ ddens,
                                                              float* __restrict__ gsamp, float* __restrict__ gdens,
                                                              int64_t gstride, const uint8_t* __restrict__ skip) {
#pragma clang fp contract(fast)
    __shared__ float s_x[64][3];
    __shared__ float s_df[64][49];
    __shared__ float s_dx[64][3];
    const int64_t v = blockIdx.x;
    if (v >= nv) return;
    if (skip && skip[v]) {  // a static or converged cell passes no gradient
        if (gsamp)
#pragma unroll
            for (int k = 0; k < 3; k++) gsamp[3 * (v * 64 + threadIdx.x) + k] = 0.0f;
        return;
    }
    const int t = threadIdx.x;
    const int c = valid[v];
    const int64_t beg = c == 0 ? 0 : prefix[c - 1], end = prefix[c];
    const int64_t si = v * 64 + t;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        s_x[t][k] = samples[3 * si + k];
        s_dx[t][k] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 48; k++) s_df[t][k] = dfeat ? dfeat[48 * si + k] : 0.0f;
    s_df[t][48] = ddens ? ddens[si] : 0.0f;
    __syncthreads();
    for (int64_t b0 = beg; b0 < end; b0 += 64) {
        const int64_t li = b0 + t;
        const bool has = li < end;
        const int g = has ? list[li] : 0;
        float mu[3] = {0.0f, 0.0f, 0.0f}, Q[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, o = 0.0f;
        float cg[48];
#pragma unroll
        for (int k = 0; k < 48; k++) cg[k] = 0.0f;
        if (has) {
            float R[3][3], S[3][3];
            density_inverse_cov(rot, scale, g, flt, v, Q, R, S);
            mu[0] = pos[3 * g]; mu[1] = pos[3 * g + 1]; mu[2] = pos[3 * g + 2];
            o = opacity[g];
            const float4* s4 = reinterpret_cast<const float4*>(sh + 48 * (int64_t)g);
#pragma unroll
            for (int k = 0; k < 12; k++) {
                const float4 q4 = s4[k];
                cg[4 * k] = q4.x; cg[4 * k + 1] = q4.y; cg[4 * k + 2] = q4.z; cg[4 * k + 3] = q4.w;
            }
        }
        float gmu[3] = {0.0f, 0.0f, 0.0f}, gQ[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, gc[48], go = 0.0f;
#pragma unroll
        for (int k = 0; k < 48; k++) gc[k] = 0.0f;
        for (int s = 0; s < 64; s++) {
            const float d0 = s_x[s][0] - mu[0], d1 = s_x[s][1] - mu[1], d2 = s_x[s][2] - mu[2];
            const float qf = Q[0] * d0 * d0 + Q[1] * d1 * d1 + Q[2] * d2 * d2 +
                             2.0f * (Q[3] * d0 * d1 + Q[4] * d0 * d2 + Q[5] * d1 * d2);
            const float ex = __expf(-0.5f * qf);
            const float a = o * ex;
            float da = s_df[s][48];
#pragma unroll
            for (int k = 0; k < 48; k++) da += s_df[s][k] * cg[k];
#pragma unroll
            for (int k = 0; k < 48; k++) gc[k] += a * s_df[s][k];
            go += ex * da;
            const float dq = -0.5f * a * da;
            gQ[0] += dq * d0 * d0;
            gQ[1] += dq * d1 * d1;
            gQ[2] += dq * d2 * d2;
            gQ[3] += 2.0f * dq * d0 * d1;
            gQ[4] += 2.0f * dq * d0 * d2;
            gQ[5] += 2.0f * dq * d1 * d2;
            const float e0 = 2.0f * dq * (Q[0] * d0 + Q[3] * d1 + Q[4] * d2);
            const float e1 = 2.0f * dq * (Q[3] * d0 + Q[1] * d1 + Q[5] * d2);
            const float e2 = 2.0f * dq * (Q[4] * d0 + Q[5] * d1 + Q[2] * d2);
            gmu[0] -= e0; gmu[1] -= e1; gmu[2] -= e2;
            const float w0 = wave_sum(e0), w1 = wave_sum(e1), w2 = wave_sum(e2);
            if (t == 0) {
                s_dx[s][0] += w0;
                s_dx[s][1] += w1;
                s_dx[s][2] += w2;
            }
        }
        if (has && flt.ada) {
            const float gq[6] = {gQ[0], gQ[1], gQ[2], gQ[3], gQ[4], gQ[5]};
            dinv_to_dsigma(Q, gq, gQ);
        }
        if (has) {
            // component-major: the wave's atomics of one component land on
            // neighbouring addresses (list entries are sorted by Gaussian)
            float* dst = gdens + g;
#pragma unroll
            for (int k = 0; k < 3; k++) if (gmu[k] != 0.0f) unsafeAtomicAdd(dst + k * gstride, gmu[k]);
#pragma unroll
            for (int k = 0; k < 6; k++) if (gQ[k] != 0.0f) unsafeAtomicAdd(dst + (3 + k) * gstride, gQ[k]);
#pragma unroll
            for (int k = 0; k < 48; k++) if (gc[k] != 0.0f) unsafeAtomicAdd(dst + (9 + k) * gstride, gc[k]);
            if (go != 0.0f) unsafeAtomicAdd(dst + 57 * gstride, go);
        }
    }
    __syncthreads();
    if (gsamp)
#pragma unroll
        for (int k = 0; k < 3; k++) gsamp[3 * si + k] = s_dx[t][k];
}

// dL/dSigma of a density Gaussian (packed as dinv_to_dsigma writes it) back to
// its rotation q (unnormalised, as stored) and scale s, Sigma = R(q/|q|)
// diag(s^2) R^T + L. Adds into dq and ds.
__device__ __forceinline__ void density_sigma_backward(const float* __restrict__ rot,
                                                       const float* __restrict__ scale, int64_t g, const float ds6[6],
                                                       float dq[4], float ds[3]) {
    const float s1[3] = {scale[3 * g], scale[3 * g + 1], scale[3 * g + 2]};
    const float q1[4] = {rot[4 * g], rot[4 * g + 1], rot[4 * g + 2], rot[4 * g + 3]};
    float qn[4] = {q1[0], q1[1], q1[2], q1[3]};
    quat_normalize(qn);
    float R[3][3];
    quat_to_mat(qn[0], qn[1], qn[2], qn[3], R);
    const float dS[3][3] = {{ds6[0], 0.5f * ds6[3], 0.5f * ds6[4]},
                            {0.5f * ds6[3], ds6[1], 0.5f * ds6[5]},
                            {0.5f * ds6[4], 0.5f * ds6[5], ds6[2]}};
    float dR[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            float acc = 0.0f;
#pragma unroll
            for (int j = 0; j < 3; j++) acc += (dS[a][j] + dS[j][a]) * R[j][b];
            dR[a][b] = s1[b] * s1[b] * acc;
        }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float dD = 0.0f;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dD += dS[i][j] * R[i][k] * R[j][k];
        ds[k] += dD * 2.0f * s1[k];
    }
    float dqn[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    quat_to_mat_backward(qn, dR, dqn);
    quat_normalize_backward(q1, dqn, dq);
}

// The accumulated covariance gradient d6 (d[3..8] of the density backward) of
// Gaussian g back to (q, s): dL/dQ with the scalar filter (Sigma = R diag(s^2)
// R^T + lpf I, dSigma = -Q dQ Q), dL/dSigma itself with an adaptive one.
__device__ __forceinline__ void density_cov_backward(const float* __restrict__ rot, const float* __restrict__ scale,
                                                     int64_t g, const DensityFilter& flt, const float d6[6],
                                                     float dq[4], float ds[3]) {
    float ds6[6];
    if (flt.ada) {
#pragma unroll
        for (int k = 0; k < 6; k++) ds6[k] = d6[k];
    } else {
        float Qv[6], R[3][3], S[3][3];
        density_inverse_cov(rot, scale, g, flt, 0, Qv, R, S);
        dinv_to_dsigma(Qv, d6, ds6);
    }
    density_sigma_backward(rot, scale, g, ds6, dq, ds);
}

// The density gradients of every Gaussian's own attributes (the VJP of the
// density sample, arapgs_density_backward): position, rotation (as stored),
// scale, opacity and SH from the component-major accumulators gd (58 x N).
__global__ void __launch_bounds__(256) density_attr_grad_kernel(int64_t N, const float* __restrict__ rot,
                                                                const float* __restrict__ scale,
                                                                const float* __restrict__ gd, DensityFilter flt,
                                                                float* __restrict__ dpos, float* __restrict__ drot,
                                                                float* __restrict__ dscale, float* __restrict__ dopac,
                                                                float* __restrict__ dsh) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    float dq6[6];
#pragma unroll
    for (int k = 0; k < 6; k++) dq6[k] = gd[(3 + k) * N + g];
    float dq[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ds[3] = {0.0f, 0.0f, 0.0f};
    density_cov_backward(rot, scale, g, flt, dq6, dq, ds);
#pragma unroll
    for (int k = 0; k < 3; k++) dpos[3 * g + k] = gd[k * N + g];
#pragma unroll
    for (int k = 0; k < 4; k++) drot[4 * g + k] = dq[k];
#pragma unroll
    for (int k = 0; k < 3; k++) dscale[3 * g + k] = ds[k];
    dopac[g] = gd[57 * N + g];
#pragma unroll
    for (int k = 0; k < 48; k++) dsh[48 * g + k] = gd[(9 + k) * N + g];
}

// One thread per Gaussian: the adjoint of everything between the rest
// attributes and (mu', q', s', c'). gr: the renderer's per-Gaussian gradient
// (d mu' 3, d q' 4, d s' 3, d c' 48); gd: the density's, component-major
// (58 x N: d mu' 3, d Q 6, d c' 48, d o) or null; opacities are constants here. Writes d e' (18) and d q0, d s0.
__global__ void __launch_bounds__(128) six_point_backward_kernel(
    int64_t N, int K, const NodeXf* __restrict__ table, const uint32_t* __restrict__ ends_idx,
    const float* __restrict__ ends_w, const float* __restrict__ ends, const float* __restrict__ rot0,
    const float* __restrict__ scale0, const float* __restrict__ sh0, const float* __restrict__ rot1,
    const float* __restrict__ scale1, const float* __restrict__ gr, const float* __restrict__ gd, DensityFilter flt,
    float pad, float coeff, const ShRotTables* __restrict__ Tsh, float* __restrict__ gends,
    float* __restrict__ drot0, float* __restrict__ dscale0) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    const float* r = gr + 58 * g;
    float dmu[3] = {r[0], r[1], r[2]};
    float dq1[4] = {r[3], r[4], r[5], r[6]};
    float ds1[3] = {r[7], r[8], r[9]};
    float dc1[48];
#pragma unroll
    for (int k = 0; k < 48; k++) dc1[k] = r[10 + k];
    if (gd) {
        float d[57];
#pragma unroll
        for (int k = 0; k < 57; k++) d[k] = gd[k * N + g];
#pragma unroll
        for (int k = 0; k < 3; k++) dmu[k] += d[k];
#pragma unroll
        for (int k = 0; k < 48; k++) dc1[k] += d[9 + k];
        density_cov_backward(rot1, scale1, g, flt, d + 3, dq1, ds1);
    }

    // recompute the six-point forward from the blended end points
    float e[6][3];
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
        for (int c = 0; c < 3; c++) e[p][c] = ends[18 * g + 3 * p + c];
    double M[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) M[c][i] = 0.5 * ((double)e[2 * i][c] - (double)e[2 * i + 1][c]);
    double Rd[3][3], Vd[3][3], sig[3];
    polar3_fp64_svd(M, Rd, Vd, sig);
    double Sd[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
        Sd[i] = Vd[i][0] * sig[0] * Vd[i][0] + Vd[i][1] * sig[1] * Vd[i][1] + Vd[i][2] * sig[2] * Vd[i][2];
    float Rf[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Rf[i][j] = (float)Rd[i][j];
    float qm[4];
    mat_to_quat(Rf, qm);
    float q[4] = {qm[0], qm[1], qm[2], qm[3]};
    quat_normalize(q);

    // SH: c' = rot(R(normalise(q qi))) c0, qi = conj(q0) / |q0|^2
    const float qo[4] = {rot0[4 * g], rot0[4 * g + 1], rot0[4 * g + 2], rot0[4 * g + 3]};
    const float n2 = qo[0] * qo[0] + qo[1] * qo[1] + qo[2] * qo[2] + qo[3] * qo[3];
    const float qi[4] = {qo[0] / n2, -qo[1] / n2, -qo[2] / n2, -qo[3] / n2};
    float qm2[4];
    quat_mul(q, qi, qm2);
    float qr[4] = {qm2[0], qm2[1], qm2[2], qm2[3]};
    quat_normalize(qr);
    float Rr[3][3];
    quat_to_mat(qr[0], qr[1], qr[2], qr[3], Rr);
    float c0[48];
    const float4* c4 = reinterpret_cast<const float4*>(sh0 + 48 * g);
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const float4 v = c4[k];
        c0[4 * k] = v.x; c0[4 * k + 1] = v.y; c0[4 * k + 2] = v.z; c0[4 * k + 3] = v.w;
    }
    float dRr[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    sh_rotate_backward(*Tsh, Rr, c0, dc1, dRr);
    float dqr[4] = {0.0f, 0.0f, 0.0f, 0.0f}, dqm2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    quat_to_mat_backward(qr, dRr, dqr);
    quat_normalize_backward(qm2, dqr, dqm2);
    float dqi[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    quat_mul_backward(q, qi, dqm2, dq1, dqi);
    // qi_k = c_k qo_k / n2 (c = 1, -1, -1, -1)
    const float cs[4] = {1.0f, -1.0f, -1.0f, -1.0f};
    float dq0[4];
    {
        float sdot = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) sdot += dqi[k] * cs[k] * qo[k];
#pragma unroll
        for (int k = 0; k < 4; k++) dq0[k] = cs[k] * dqi[k] / n2 - 2.0f * qo[k] * sdot / (n2 * n2);
    }

    // scales: s'_i = S_ii / ((s0_i + pad) coeff) s0_i
    double dSd[3];
    float ds0[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float s0 = scale0[3 * g + i];
        const float den = (s0 + pad) * coeff;
        dSd[i] = (double)(ds1[i] * s0 / den);
        ds0[i] = ds1[i] * (float)Sd[i] * pad / (coeff * (s0 + pad) * (s0 + pad));
    }
    // q' = normalise(mat_to_quat(R))
    float dqm[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    quat_normalize_backward(qm, dq1, dqm);
    float dRf[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    mat_to_quat_backward(Rf, dqm, dRf);
    double dRd[3][3], dM[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) dRd[i][j] = (double)dRf[i][j];
    polar3_fp64_backward(Rd, Vd, sig, dRd, dSd, dM);

    // d e' from M and from the centre
    float de[6][3];
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
        for (int c = 0; c < 3; c++) de[p][c] = dmu[c] / 6.0f;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float h = (float)(0.5 * dM[c][i]);
            de[2 * i][c] += h;
            de[2 * i + 1][c] -= h;
        }
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
        for (int c = 0; c < 3; c++) gends[18 * g + 3 * p + c] = de[p][c];

    // blend adjoint to the rest end points: d e = sum_j w_j R_j^T d e'
    float dE[6][3];
#pragma unroll
    for (int p = 0; p < 6; p++) {
        dE[p][0] = dE[p][1] = dE[p][2] = 0.0f;
        const int64_t pt = 6 * g + p;
        for (int j = 0; j < K; j++) {
            const NodeXf& nd = table[ends_idx[pt * K + j]];
            const float w = ends_w[pt * K + j];
#pragma unroll
            for (int c = 0; c < 3; c++)
                dE[p][c] += w * (nd.R[c] * de[p][0] + nd.R[3 + c] * de[p][1] + nd.R[6 + c] * de[p][2]);
        }
    }
    // end points: e_{i+-} = mu0 +- R0[:, i] (s0_i + pad) coeff, R0 = R(q0 / |q0|)
    float qn0[4] = {qo[0], qo[1], qo[2], qo[3]};
    quat_normalize(qn0);
    float R0[3][3];
    quat_to_mat(qn0[0], qn0[1], qn0[2], qn0[3], R0);
    float dR0[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float len = (scale0[3 * g + i] + pad) * coeff;
        float dlen = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float dd = dE[2 * i][c] - dE[2 * i + 1][c];
            dlen += R0[c][i] * dd;
            dR0[c][i] = len * dd;
        }
        ds0[i] += dlen * coeff;
    }
    float dqn0[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    quat_to_mat_backward(qn0, dR0, dqn0);
    quat_normalize_backward(qo, dqn0, dq0);
#pragma unroll
    for (int k = 0; k < 4; k++) drot0[4 * g + k] = dq0[k];
#pragma unroll
    for (int k = 0; k < 3; k++) dscale0[3 * g + k] = ds0[k];
}

// Entry e of the inverse index is (point, slot) = (e / K, e % K) over the
// 6N end points followed by the S samples; its node is the blend id.
__global__ void entry_keys_kernel(int64_t n_end_entries, int64_t n_entries, const uint32_t* __restrict__ ends_idx,
                                  const uint32_t* __restrict__ sample_idx, uint64_t* __restrict__ keys,
                                  uint32_t* __restrict__ vals) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_entries) return;
    keys[e] = e < n_end_entries ? ends_idx[e] : sample_idx[e - n_end_entries];
    vals[e] = (uint32_t)e;
}

__global__ void zero_u32_kernel(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void node_ranges_kernel(const uint64_t* __restrict__ keys, int64_t n, int M, uint2* __restrict__ rng) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = (uint32_t)keys[i];
    if (k >= (uint32_t)M) return;  // not a node id: dropped
    if (i == 0 || (uint32_t)keys[i - 1] != k) rng[k].x = (uint32_t)i;
    if (i == n - 1 || (uint32_t)keys[i + 1] != k) rng[k].y = (uint32_t)(i + 1);
}

// d R_j = sum w d e' (p - g_j)^T and d t_j = sum w d e' over the node's
// entries (p the blend input: rest end point or rest sample). Thread partials
// in FP64, fixed-order block reduction: deterministic.
__global__ void __launch_bounds__(256) node_gather_kernel(int M, int K, int64_t n_end_entries,
                                                          const uint2* __restrict__ rng,
                                                          const uint32_t* __restrict__ entries,
                                                          const NodeXf* __restrict__ table,
                                                          const float* __restrict__ ends_w,
                                                          const float* __restrict__ ends0,
                                                          const float* __restrict__ gends,
                                                          const float* __restrict__ sample_w,
                                                          const float* __restrict__ samples0,
                                                          const float* __restrict__ gsamp,
                                                          NodeXf* __restrict__ dtable) {
    __shared__ double s_red[4][12];
    const int n = blockIdx.x;
    if (n >= M) return;
    const uint2 r = rng[n];
    const float gx = table[n].g[0], gy = table[n].g[1], gz = table[n].g[2];
    double acc[12];
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = 0.0;
    for (uint32_t i = r.x + threadIdx.x; i < r.y; i += blockDim.x) {
        const uint32_t e = entries[i];
        float w, px, py, pz, d0, d1, d2;
        if ((int64_t)e < n_end_entries) {
            const uint32_t p = e / (uint32_t)K;
            w = ends_w[e];
            px = ends0[3 * (int64_t)p]; py = ends0[3 * (int64_t)p + 1]; pz = ends0[3 * (int64_t)p + 2];
            d0 = gends[3 * (int64_t)p]; d1 = gends[3 * (int64_t)p + 1]; d2 = gends[3 * (int64_t)p + 2];
        } else {
            const int64_t es = (int64_t)e - n_end_entries;
            const int64_t s = es / K;
            w = sample_w[es];
            px = samples0[3 * s]; py = samples0[3 * s + 1]; pz = samples0[3 * s + 2];
            d0 = gsamp[3 * s]; d1 = gsamp[3 * s + 1]; d2 = gsamp[3 * s + 2];
        }
        const double tx = (double)(px - gx), ty = (double)(py - gy), tz = (double)(pz - gz);
        const double wd0 = (double)w * d0, wd1 = (double)w * d1, wd2 = (double)w * d2;
        acc[0] += wd0 * tx; acc[1] += wd0 * ty; acc[2] += wd0 * tz;
        acc[3] += wd1 * tx; acc[4] += wd1 * ty; acc[5] += wd1 * tz;
        acc[6] += wd2 * tx; acc[7] += wd2 * ty; acc[8] += wd2 * tz;
        acc[9] += wd0; acc[10] += wd1; acc[11] += wd2;
    }
#pragma unroll
    for (int k = 0; k < 12; k++) {
        double v = acc[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        acc[k] = v;
    }
    const int wid = threadIdx.x >> 6, wl = threadIdx.x & 63;
    if (wl == 0)
#pragma unroll
        for (int k = 0; k < 12; k++) s_red[wid][k] = acc[k];
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        const float v = (float)(s_red[0][k] + s_red[1][k] + s_red[2][k] + s_red[3][k]);
        float* o = reinterpret_cast<float*>(dtable + n);
        if (k < 9) o[3 + k] = v;   // R, row-major
        else o[12 + (k - 9)] = v;  // t
    }
    if (threadIdx.x >= 12 && threadIdx.x < 16) {
        float* o = reinterpret_cast<float*>(dtable + n);
        const int k = threadIdx.x - 12;  // g (3) and pad: not differentiated
        o[k < 3 ? k : 15] = 0.0f;
    }
}

// ---------------------------------------------------------------- realignment
// raw parameters -> activated: s = exp(log s), o = sigmoid(logit o)
__global__ void __launch_bounds__(256) activate_kernel(int64_t N, const float* __restrict__ log_scale,
                                                       const float* __restrict__ logit_opac, float* __restrict__ scale,
                                                       float* __restrict__ opac) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
#pragma unroll
    for (int k = 0; k < 3; k++) scale[3 * g + k] = expf(log_scale[3 * g + k]);
    opac[g] = 1.0f / (1.0f + expf(-logit_opac[g]));
}

// L1 loss sum |feat - aim| + w sum |dens - aim_dens| and its subgradient
// (sign, 0 at equality); the loss is reduced in FP64 (block sum, one atomic).
__global__ void __launch_bounds__(256) l1_grad_kernel(int64_t S, const float* __restrict__ feat,
                                                      const float* __restrict__ dens, const float* __restrict__ aim_feat,
                                                      const float* __restrict__ aim_dens, float w,
                                                      float* __restrict__ dfeat, float* __restrict__ ddens,
                                                      double* __restrict__ loss, float* __restrict__ cell_loss) {
    __shared__ double s_red[4];
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (s < S) {
#pragma unroll 4
        for (int k = 0; k < 48; k++) {
            const float d = feat[48 * s + k] - aim_feat[48 * s + k];
            dfeat[48 * s + k] = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
            acc += fabs((double)d);
        }
        const float d = dens[s] - aim_dens[s];
        ddens[s] = d > 0.0f ? w : (d < 0.0f ? -w : 0.0f);
        acc += (double)w * fabs((double)d);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    // one wave = the 64 samples of one valid cell
    if ((threadIdx.x & 63) == 0 && cell_loss && s < S) cell_loss[s >> 6] = (float)acc;
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0 && loss) unsafeAtomicAdd(loss, s_red[0] + s_red[1] + s_red[2] + s_red[3]);
}

// Cells whose lists hold no moved Gaussian (the reference's current static
// grids, GaussianView.cpp:3850-3852): one wave per valid cell.
__global__ void __launch_bounds__(64) static_cells_kernel(int64_t nv, const int* __restrict__ valid,
                                                          const int64_t* __restrict__ prefix,
                                                          const int* __restrict__ list,
                                                          const uint8_t* __restrict__ moved,
                                                          uint8_t* __restrict__ stat) {
    const int64_t v = blockIdx.x;
    if (v >= nv) return;
    const int c = valid[v];
    const int64_t beg = c == 0 ? 0 : prefix[c - 1], end = prefix[c];
    bool any = false;
    for (int64_t i = beg + threadIdx.x; i < end && !any; i += 64) any = moved[list[i]] != 0;
    any = __ballot(any) != 0ull;
    if (threadIdx.x == 0) stat[v] = any ? 0 : 1;
}

// Per-cell convergence of the realignment: the first step of a run (a new
// realignment or new options) records every cell's loss and clears the
// converged flags; a later step marks a cell converged once its loss is at
// most tol times that record (sticky, also across re-binnings, as the
// reference's grid_is_converged). skip = static or converged.
__global__ void __launch_bounds__(256) converge_kernel(int64_t nv, const float* __restrict__ cell_loss,
                                                       float* __restrict__ cell_ref,
                                                       const uint8_t* __restrict__ stat, float tol, int first,
                                                       uint8_t* __restrict__ conv, uint8_t* __restrict__ skip) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const bool st = stat && stat[v];
    if (first) {
        cell_ref[v] = cell_loss[v];
        conv[v] = 0;
        skip[v] = st ? 1 : 0;
        return;
    }
    const bool c = conv[v] || (tol > 0.0f && cell_loss[v] <= tol * cell_ref[v]);
    conv[v] = c ? 1 : 0;
    skip[v] = (st || c) ? 1 : 0;
}

// ExpandOpRange (GaussianView.cpp:4320-4360): every cell that is, or is a face
// neighbour (clamped to the grid) of, a non-converged valid cell is marked in
// map (G^3 bytes, cleared by the caller) ...
__global__ void __launch_bounds__(256) expand_mark_kernel(int64_t nv, const int* __restrict__ valid,
                                                          const uint8_t* __restrict__ conv, int G,
                                                          uint8_t* __restrict__ map) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv || conv[v]) return;
    const int c = valid[v];
    const int z = c % G, y = (c / G) % G, x = c / (G * G);
    const int64_t GG = (int64_t)G * G;
    map[c] = 1;
    map[(int64_t)max(x - 1, 0) * GG + (int64_t)y * G + z] = 1;
    map[(int64_t)min(x + 1, G - 1) * GG + (int64_t)y * G + z] = 1;
    map[(int64_t)x * GG + (int64_t)max(y - 1, 0) * G + z] = 1;
    map[(int64_t)x * GG + (int64_t)min(y + 1, G - 1) * G + z] = 1;
    map[(int64_t)x * GG + (int64_t)y * G + max(z - 1, 0)] = 1;
    map[(int64_t)x * GG + (int64_t)y * G + min(z + 1, G - 1)] = 1;
}

// ... and every marked valid cell is no longer converged.
__global__ void __launch_bounds__(256) expand_apply_kernel(int64_t nv, const int* __restrict__ valid,
                                                           const uint8_t* __restrict__ map,
                                                           const uint8_t* __restrict__ stat,
                                                           uint8_t* __restrict__ conv, uint8_t* __restrict__ skip) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv || !map[valid[v]]) return;
    conv[v] = 0;
    skip[v] = (stat && stat[v]) ? 1 : 0;
}

// One Adam step per Gaussian over its 59 raw parameters (pos 3, rot 4,
// log-scale 3, logit-opacity 1, SH 48); the activated gradients are chained to
// the raw ones (d log s = d s * s, d logit o = d o * o (1 - o)). m and v are
// component-major (k * N + g); bc1 / bc2 are the bias corrections 1 - beta^t.
__global__ void __launch_bounds__(256) adam_kernel(int64_t N, const float* __restrict__ dpos,
                                                   const float* __restrict__ drot, const float* __restrict__ dscale,
                                                   const float* __restrict__ dopac, const float* __restrict__ dsh,
                                                   const float* __restrict__ scale, const float* __restrict__ opac,
                                                   float* __restrict__ pos, float* __restrict__ rot,
                                                   float* __restrict__ log_scale, float* __restrict__ logit_opac,
                                                   float* __restrict__ sh, float* __restrict__ m, float* __restrict__ v,
                                                   float lr_pos, float lr_rot, float lr_scale, float lr_opac,
                                                   float lr_sh, float b1, float b2, float eps, float bc1, float bc2,
                                                   const uint8_t* __restrict__ moved,
                                                   const float* __restrict__ max_scale) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    if (moved && !moved[g]) return;  // "Only Optimize Deformed Region": the rest stays as it is
    auto step = [&](int k, float grad, float lr, float& p) {
        float& mk = m[k * N + g];
        float& vk = v[k * N + g];
        mk = b1 * mk + (1.0f - b1) * grad;
        vk = b2 * vk + (1.0f - b2) * grad * grad;
        p -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
    };
#pragma unroll
    for (int c = 0; c < 3; c++) step(c, dpos[3 * g + c], lr_pos, pos[3 * g + c]);
#pragma unroll
    for (int c = 0; c < 4; c++) step(3 + c, drot[4 * g + c], lr_rot, rot[4 * g + c]);
#pragma unroll
    for (int c = 0; c < 3; c++) step(7 + c, dscale[3 * g + c] * scale[3 * g + c], lr_scale, log_scale[3 * g + c]);
    const float o = opac[g];
    step(10, dopac[g] * o * (1.0f - o), lr_opac, logit_opac[g]);
#pragma unroll 4
    for (int c = 0; c < 48; c++) step(11 + c, dsh[48 * g + c], lr_sh, sh[48 * g + c]);
    if (max_scale)
#pragma unroll
        for (int c = 0; c < 3; c++) log_scale[3 * g + c] = fminf(log_scale[3 * g + c], logf(max_scale[3 * g + c]));
}

}  // namespace

int launch_activate(int64_t N, const float* log_scale, const float* logit_opac, float* scale, float* opac,
                    hipStream_t st) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(activate_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, N, log_scale, logit_opac, scale, opac);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_l1_grad(int64_t S, const float* feat, const float* dens, const float* aim_feat, const float* aim_dens,
                   float w, float* dfeat, float* ddens, double* loss, float* cell_loss, hipStream_t st) {
    if (S <= 0) return 0;
    hipLaunchKernelGGL(l1_grad_kernel, dim3(nblk(S, 256)), dim3(256), 0, st, S, feat, dens, aim_feat, aim_dens, w,
                       dfeat, ddens, loss, cell_loss);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_static_cells(int64_t nv, const int* valid, const int64_t* prefix, const int* list, const uint8_t* moved,
                        uint8_t* stat, hipStream_t st) {
    if (nv <= 0) return 0;
    hipLaunchKernelGGL(static_cells_kernel, dim3((unsigned)nv), dim3(64), 0, st, nv, valid, prefix, list, moved, stat);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_converge(int64_t nv, const float* cell_loss, float* cell_ref, const uint8_t* stat, float tol, int first,
                    uint8_t* conv, uint8_t* skip, hipStream_t st) {
    if (nv <= 0) return 0;
    hipLaunchKernelGGL(converge_kernel, dim3(nblk(nv, 256)), dim3(256), 0, st, nv, cell_loss, cell_ref, stat, tol,
                       first, conv, skip);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_expand_op_range(int64_t nv, const int* valid, int G, const uint8_t* stat, uint8_t* conv, uint8_t* skip,
                           uint8_t* map, hipStream_t st) {
    if (nv <= 0) return 0;
    const int64_t gc = (int64_t)G * G * G;
    hipLaunchKernelGGL(zero_f32_kernel, dim3(std::min<int64_t>(nblk((gc + 3) / 4, 256), 4096)), dim3(256), 0, st,
                       reinterpret_cast<float*>(map), (gc + 3) / 4);
    hipLaunchKernelGGL(expand_mark_kernel, dim3(nblk(nv, 256)), dim3(256), 0, st, nv, valid, conv, G, map);
    hipLaunchKernelGGL(expand_apply_kernel, dim3(nblk(nv, 256)), dim3(256), 0, st, nv, valid, map, stat, conv, skip);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_adam(int64_t N, const float* dpos, const float* drot, const float* dscale, const float* dopac,
                const float* dsh, const float* scale, const float* opac, float* pos, float* rot, float* log_scale,
                float* logit_opac, float* sh, float* m, float* v, const float lr[5], float b1, float b2, float eps,
                float bc1, float bc2, const uint8_t* moved, const float* max_scale, hipStream_t st) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, N, dpos, drot, dscale, dopac, dsh, scale,
                       opac, pos, rot, log_scale, logit_opac, sh, m, v, lr[0], lr[1], lr[2], lr[3], lr[4], b1, b2, eps,
                       bc1, bc2, moved, max_scale);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_zero_f32(float* p, int64_t n, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(zero_f32_kernel, dim3(std::min<unsigned>(nblk(n, 256), 8192)), dim3(256), 0, st, p, n);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_density_backward(int64_t nv, const int* valid, const int64_t* prefix, const int* list, const float* samples,
                            const float* pos, const float* rot, const float* scale, const float* opacity,
                            const float* sh, DensityFilter flt, const float* dfeat, const float* ddens, float* gsamp,
                            float* gdens, int64_t n_gauss, hipStream_t st, const uint8_t* skip) {
    if (nv <= 0) return 0;
    hipLaunchKernelGGL(density_backward_kernel, dim3((unsigned)nv), dim3(64), 0, st, nv, valid, prefix, list, samples,
                       pos, rot, scale, opacity, sh, flt, dfeat, ddens, gsamp, gdens, n_gauss, skip);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_density_attr_grad(int64_t N, const float* rot, const float* scale, const float* gd, DensityFilter flt, float* dpos,
                             float* drot, float* dscale, float* dopac, float* dsh, hipStream_t st) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(density_attr_grad_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, N, rot, scale, gd, flt, dpos,
                       drot, dscale, dopac, dsh);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_six_point_backward(int64_t N, int K, const NodeXf* table, const uint32_t* ends_idx, const float* ends_w,
                              const float* ends, const float* rot0, const float* scale0, const float* sh0,
                              const float* rot1, const float* scale1, const float* gr, const float* gd, DensityFilter flt,
                              float pad, float coeff, const ShRotTables* Tsh, float* gends, float* drot0,
                              float* dscale0, hipStream_t st) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(six_point_backward_kernel, dim3(nblk(N, 128)), dim3(128), 0, st, N, K, table, ends_idx,
                       ends_w, ends, rot0, scale0, sh0, rot1, scale1, gr, gd, flt, pad, coeff, Tsh, gends, drot0,
                       dscale0);
    HIP_RET(hipGetLastError());
    return 0;
}

size_t inverse_index_words(int64_t n_entries) {
    return (size_t)n_entries * 6 + radix_workspace_words(n_entries) + 64;
}

int build_inverse_index(int64_t n_end_entries, int64_t n_entries, int M, const uint32_t* ends_idx,
                        const uint32_t* sample_idx, uint32_t* ws, uint32_t* entries, uint2* rng, hipStream_t st) {
    if (n_entries <= 0 || n_entries > 0xffffffffll) return -1;
    uint64_t* keys = reinterpret_cast<uint64_t*>(ws);
    uint64_t* keys_alt = keys + n_entries;
    uint32_t* vals_alt = reinterpret_cast<uint32_t*>(keys_alt + n_entries);
    uint32_t* sort_ws = vals_alt + n_entries;
    hipLaunchKernelGGL(entry_keys_kernel, dim3(nblk(n_entries, 256)), dim3(256), 0, st, n_end_entries, n_entries,
                       ends_idx, sample_idx, keys, entries);
    HIP_RET(hipGetLastError());
    int bits = 1;
    while ((1ll << bits) < M) bits++;
    int rc = launch_radix_sort(keys, entries, keys_alt, vals_alt, n_entries, bits, sort_ws, st);
    if (rc) return rc;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(nblk(2 * (int64_t)M, 256)), dim3(256), 0, st, (uint32_t*)rng,
                       2 * (int64_t)M);
    hipLaunchKernelGGL(node_ranges_kernel, dim3(nblk(n_entries, 256)), dim3(256), 0, st, keys, n_entries, M, rng);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_node_gather(int M, int K, int64_t n_end_entries, const uint2* rng, const uint32_t* entries,
                       const NodeXf* table, const float* ends_w, const float* ends0, const float* gends,
                       const float* sample_w, const float* samples0, const float* gsamp, NodeXf* dtable,
                       hipStream_t st) {
    if (M <= 0) return 0;
    hipLaunchKernelGGL(node_gather_kernel, dim3((unsigned)M), dim3(256), 0, st, M, K, n_end_entries, rng, entries,
                       table, ends_w, ends0, gends, sample_w, samples0, gsamp, dtable);
    HIP_RET(hipGetLastError());
    return 0;
}

}  // namespace arapgs
