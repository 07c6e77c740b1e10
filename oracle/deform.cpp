// TEST INFRASTRUCTURE — the oracle (see common.hpp). Graph build (KNN,
// weights, edges, FPS) and the per-frame deformation of Gaussians and samples.
#include "common.hpp"

#include <algorithm>
#include <cfloat>
#include <limits>

using namespace orc;

namespace {

// FP32 distance exactly as Deform.hpp:164-165: float difference, float
// products summed left to right, float sqrt.
inline float knn_dist(const float* p, const float* n) {
    const float a = p[0] - n[0], b = p[1] - n[1], c = p[2] - n[2];
    return std::sqrt(a * a + b * b + c * c);
}

// DeformGraph::findNearestNodes (Deform.hpp:153-185): selection sort over the
// first k slots, scanning each slot's tail from the back with strict '<', then
// swapping slot i with the winner. Ties are therefore broken by the swap
// history, which this reproduces by running the same procedure.
void find_nearest(const float* pos, const float* nodes, int M, int k, double* d, uint32_t* index,
                  uint32_t* idx_out) {
    for (int j = 0; j < M; j++) {
        index[j] = (uint32_t)j;
        d[j] = (double)knn_dist(pos, nodes + 3 * (size_t)j);
    }
    for (int i = 0; i < k; i++) {
        int idx_min = i;
        for (int j = M - 1; j > i; j--)
            if (d[j] < d[idx_min]) idx_min = j;
        idx_out[i] = index[idx_min];
        const double td = d[i];
        const uint32_t t = index[i];
        d[i] = d[idx_min];
        index[i] = index[idx_min];
        d[idx_min] = td;
        index[idx_min] = t;
    }
}

}  // namespace

extern "C" {

// DeformGraph::computeWeights (Deform.hpp:187-208) for Q query points:
// idx[q*(k+1) + s] for s <= k (slot k is the normaliser node),
// w[q*k + j] = (1 - d_j/d_k)^2 / sum, FP64; k == 1 gives w = 1.
void orc_knn_weights(const float* nodes, int M, const float* queries, long Q, int k, uint32_t* idx,
                     double* w) {
#pragma omp parallel
    {
        std::vector<double> d(M);
        std::vector<uint32_t> index(M);
#pragma omp for schedule(dynamic, 256)
        for (long q = 0; q < Q; q++) {
            const float* p = queries + 3 * q;
            uint32_t* out = idx + q * (k + 1);
            const int kk = std::min(k + 1, M);
            find_nearest(p, nodes, M, kk, d.data(), index.data(), out);
            if (!w) continue;
            const double dmax = (double)knn_dist(p, nodes + 3 * (size_t)out[k]);
            double sum = 0.0;
            double* wq = w + q * k;
            for (int j = 0; j < k; j++) {
                const double dist = (double)knn_dist(p, nodes + 3 * (size_t)out[j]);
                const double x = 1.0 - dist / dmax;
                wq[j] = x * x;  // pow(x, 2.0)
                sum += wq[j];
            }
            if (k == 1) wq[0] = 1.0;
            else
                for (int j = 0; j < k; j++) wq[j] /= sum;
        }
    }
}

// DeformGraph::setupEdges (Deform.hpp:84-95): (k+1)-NN of every rest node,
// slot 0 dropped. nbr is M*k.
void orc_graph_edges(const float* nodes, int M, int k, uint32_t* nbr) {
    std::vector<uint32_t> idx((size_t)M * (k + 1));
    orc_knn_weights(nodes, M, nodes, M, k, idx.data(), nullptr);
    for (int i = 0; i < M; i++)
        for (int j = 1; j <= k; j++) nbr[(size_t)i * k + j - 1] = idx[(size_t)i * (k + 1) + j];
}

// farthest_control_points_sampling + FetchFirstNodeIdx (helper.cpp:139-195).
// Distances as pts_distance (helper.cpp:60-63): float differences squared and
// summed in double, double sqrt, rounded to float. Returns the count written;
// `out` needs min(M,N)+1 slots (M == 1 with a first node != 0 appends twice).
int orc_fps(const float* pos, long N, int M, int* out) {
    const int m = (int)std::min<long>(M, N);
    if (m <= 0) return 0;
    int first = 0;
    float best = -100000.0f;
    for (long i = 0; i < N; i++) {
        float s = 0.0f;
        for (int j = 0; j < 3; j++) s += pos[3 * i + j];
        if (s > best) { best = s; first = (int)i; }
    }
    int cnt = 0;
    out[cnt++] = first;
    std::vector<float> dist((size_t)N, std::numeric_limits<float>::max());
    for (int i = 0; i < m; i++) {
        if (first == i) continue;
        const float* last = pos + 3 * (size_t)out[cnt - 1];
#pragma omp parallel for schedule(static)
        for (long j = 0; j < N; j++) {
            const float dx = last[0] - pos[3 * j], dy = last[1] - pos[3 * j + 1], dz = last[2] - pos[3 * j + 2];
            const float d = (float)std::sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
            if (d < dist[j]) dist[j] = d;
        }
        long arg = 0;
        for (long j = 1; j < N; j++)
            if (dist[j] > dist[arg]) arg = j;
        out[cnt++] = (int)arg;
        if (cnt == m) break;
    }
    return cnt;
}

// GetEndPoints (GaussianView.cpp:4643-4668): e_{i,+/-} = mu +/- R(q) ((s_i+pad)*coeff) e_i,
// stored per Gaussian as [e0+, e0-, e1+, e1-, e2+, e2-] (x,y,z each).
void orc_end_points(const float* pos, const float* rot, const float* scale, long N, float pad, float coeff,
                    float* ends) {
#pragma omp parallel for schedule(static)
    for (long g = 0; g < N; g++) {
        float R[3][3];
        quat_to_mat(quat_normalized(Qf{rot[4 * g], rot[4 * g + 1], rot[4 * g + 2], rot[4 * g + 3]}), R);
        for (int i = 0; i < 3; i++) {
            const float len = (scale[3 * g + i] + pad) * coeff;
            for (int c = 0; c < 3; c++) {
                const float off = R[c][i] * len;
                ends[g * 18 + i * 6 + c] = pos[3 * g + c] + off;
                ends[g * 18 + i * 6 + 3 + c] = pos[3 * g + c] - off;
            }
        }
    }
}

// DeformGraph::predict_mesh (Deform.hpp:230-249): p' = sum_j w_j (R_j (p - g_j) + t_j + g_j)
// with p - g_j in float, the term in double and a float accumulator.
// rot9 is column-major per node (Deform.hpp:29-36).
void orc_blend(const float* pts, long P, const uint32_t* idx, const double* w, int k, const float* node_pos,
               const double* rot9, const double* t3, float* out) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < P; i++) {
        float o[3] = {0.0f, 0.0f, 0.0f};
        for (int j = 0; j < k; j++) {
            const uint32_t n = idx[i * k + j];
            const double wj = w[i * k + j];
            const float* g = node_pos + 3 * (size_t)n;
            const float tmp[3] = {pts[3 * i] - g[0], pts[3 * i + 1] - g[1], pts[3 * i + 2] - g[2]};
            const double* r = rot9 + 9 * (size_t)n;
            const double* t = t3 + 3 * (size_t)n;
            for (int c = 0; c < 3; c++)
                o[c] += wj * (r[c] * tmp[0] + r[3 + c] * tmp[1] + r[6 + c] * tmp[2] + t[c] + g[c]);
        }
        out[3 * i] = o[0]; out[3 * i + 1] = o[1]; out[3 * i + 2] = o[2];
    }
}

// UpdateAsSixPointsWithdrawBad (GaussianView.cpp:3081-3166) for the non-static
// Gaussians. M = P pinv(Q) reduces to M[:,i] = (e_i+ - e_i-)/2 because Q Q^T = 2I.
// Polar M = R S in FP64 (helper.cpp:429-440), q = quat(R) in float, normalised;
// s_i = S_ii / ((s0_i + pad) coeff) s0_i; mu = mean of the 6 end points;
// SH rotated by (q q_old^-1) with the odd-index sign wrapper.
void orc_six_point(long N, const float* ends, const float* scale0, const uint8_t* is_static, float pad,
                   float coeff, float* pos, float* rot, float* scale, float* sh) {
#pragma omp parallel for schedule(static)
    for (long g = 0; g < N; g++) {
        if (is_static && is_static[g]) continue;
        const float* e = ends + g * 18;
        float center[3];
        for (int c = 0; c < 3; c++) {
            float s = 0.0f;
            for (int p = 0; p < 6; p++) s += e[p * 3 + c];
            center[c] = s / 6.0f;
        }
        double M[3][3];
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++) {
                const float a = e[i * 6 + c] - center[c];
                const float b = e[i * 6 + 3 + c] - center[c];
                M[c][i] = 0.5 * ((double)a - (double)b);
            }
        double Rd[3][3], Sd[3][3];
        polar_fp64(M, Rd, Sd);
        float Rf[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Rf[i][j] = (float)Rd[i][j];
        const Qf q = quat_normalized(mat_to_quat(Rf));
        const Qf q_old{rot[4 * g], rot[4 * g + 1], rot[4 * g + 2], rot[4 * g + 3]};
        rot[4 * g] = q.w; rot[4 * g + 1] = q.x; rot[4 * g + 2] = q.y; rot[4 * g + 3] = q.z;
        for (int i = 0; i < 3; i++) {
            const float K = (float)Sd[i][i];
            const float s0 = scale0[3 * g + i];
            scale[3 * g + i] = K / ((s0 + pad) * coeff) * s0;
        }
        for (int c = 0; c < 3; c++) pos[3 * g + c] = center[c];
        float Rrel[3][3];
        quat_to_mat(quat_normalized(quat_mul(q, quat_inverse(q_old))), Rrel);
        sh_rotate(Rrel, sh + g * 48, true);
    }
}

// FastgetOthogonalMatrix (helper.cpp:506-517) + quaternion (GaussianView.cpp:3172-3180):
// Newton polar iteration M <- (M + M^-T)/2 in float until max |delta| < 1e-6.
// q out as (w,x,y,z).
void orc_node_quats(int M, const double* rot9, float* q) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < M; n++) {
        float A[3][3];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) A[r][c] = (float)rot9[9 * (size_t)n + c * 3 + r];
        for (int it = 0; it < 100; it++) {
            // inverse transpose via cofactors: (A^-T)_{ij} = cof_{ij} / det
            float cof[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                    cof[i][j] = A[i1][j1] * A[i2][j2] - A[i1][j2] * A[i2][j1];
                }
            const float det = A[0][0] * cof[0][0] + A[0][1] * cof[0][1] + A[0][2] * cof[0][2];
            float nx[3][3], mx = 0.0f;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    nx[i][j] = 0.5f * (A[i][j] + cof[i][j] / det);
                    mx = std::max(mx, std::fabs(A[i][j] - nx[i][j]));
                }
            if (mx < 1e-6f) break;
            std::memcpy(A, nx, sizeof(A));
        }
        const Qf r = quat_normalized(mat_to_quat(A));
        q[4 * n] = r.w; q[4 * n + 1] = r.x; q[4 * n + 2] = r.y; q[4 * n + 3] = r.z;
    }
}

// RotateSHs (cudakdtree.cu:201-222) with Q_SlerpCUDA (cudakdtree.cu:113-148):
// per non-static sample, a running slerp from identity towards each node's
// quaternion with t = w_j / (w_j + sum_prev) (float weights), ratio maths in
// double; then the 48-float feature is rotated with the odd-index wrapper.
void orc_sample_sh(long S, const float* w, const uint32_t* idx, const float* node_q, int k,
                   const uint8_t* is_static, float* feat) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < S; s++) {
        if (is_static && is_static[s]) continue;
        Qf acc{1.0f, 0.0f, 0.0f, 0.0f};
        float last = 0.0f;
        for (int j = 0; j < k; j++) {
            const float cw = w[s * k + j];
            const float t = cw / (cw + last);
            const float* nq = node_q + 4 * (size_t)idx[s * k + j];
            Qf e{nq[0], nq[1], nq[2], nq[3]};
            double cosang = (double)(acc.x * e.x + acc.y * e.y + acc.z * e.z + acc.w * e.w);
            if (cosang < 0) {
                e.x = -e.x; e.y = -e.y; e.z = -e.z; e.w = -e.w;
                cosang = -cosang;
            }
            double ra, rb;
            if (cosang > 0.99995f) {
                ra = 1.0f - (double)t;
                rb = (double)t;
            } else {
                const double sinang = std::sqrt(1.0f - cosang * cosang);
                const double ang = std::atan2(sinang, cosang);
                ra = std::sin((1.0f - (double)t) * ang) / sinang;
                rb = std::sin((double)t * ang) / sinang;
            }
            Qf l{(float)(ra * acc.w + rb * e.w), (float)(ra * acc.x + rb * e.x), (float)(ra * acc.y + rb * e.y),
                 (float)(ra * acc.z + rb * e.z)};
            acc = quat_normalized(l);
            last += cw;
        }
        float R[3][3];
        quat_to_mat(quat_normalized(acc), R);
        sh_rotate(R, feat + s * 48, true);
    }
}

}  // extern "C"
