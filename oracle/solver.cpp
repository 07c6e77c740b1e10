// TEST INFRASTRUCTURE — the oracle (see common.hpp). The embedded-deformation
// ("ARAP") Gauss-Newton solve of Deform.cpp:95-581 in FP64.
//
// Unknowns: 12 per free node in ascending node order (the std::set order of
// Deform.hpp:20), x[12f + 3c + r] = R(r,c) (column-major, Deform.cpp:388-390)
// and x[12f + 9 + c] = t_c. Residual rows, in order:
//   (i)   6 per free node: sqrt(w_rot) * {c0.c1, c0.c2, c1.c2, |c0|^2-1, |c1|^2-1, |c2|^2-1}
//   (ii)  3 per (free node i, graph neighbour k): sqrt(w_reg) * s_ik * (R_i(g_k-g_i)+g_i+t_i-g_k-t_k)
//   (iii) 3 per (static node i, graph neighbour k): sqrt(w_reg) * (-t_k)
//   (iv)  constraints: 3 per control node (or per control block in centre mode)
// The Jacobian keeps the reference's asymmetry: d/dt_i of (ii) omits s_ik
// (Deform.cpp:257 vs :261).
#include "common.hpp"

#include <algorithm>
#include <cstdio>
#include <iterator>
#include <set>
#include <vector>

namespace {

struct Problem {
    int M = 0, K = 0;
    const float* g = nullptr;          // node positions M*3
    const uint32_t* nbr = nullptr;     // M*K graph neighbours
    const uint32_t* src_idx = nullptr; // M*K KNN of each node's source vertex
    const double* src_w = nullptr;     // M*K
    const float* aim = nullptr;        // M*3
    std::vector<int> free_id;          // node -> rank among free nodes, -1 if static
    std::vector<int> free_nodes, static_nodes;
    std::vector<std::vector<int>> blocks;
    std::vector<int> btype;
    std::vector<std::vector<int>> key_nodes;  // centre mode: ascending key nodes per block
    std::vector<orc::V3f> aim_center;
    double sw_rot = 0, sw_reg = 0, sw_con = 0, wf = 1.0;
    bool center = false;
    long rows = 0;
    int n = 0;

    bool same_block(int a, int b) const {
        for (size_t i = 0; i < blocks.size(); i++) {
            if (btype[i] == -1) continue;
            bool fa = false, fb = false;
            for (int v : blocks[i]) { fa |= (v == a); fb |= (v == b); }
            if (fa && fb) return true;
        }
        return false;
    }
};

struct Trip { long r; int c; double v; };

// Centre-mode key controls (Deform.cpp:34-75). The reference runs FPS on the
// block's node positions but then keeps the FIRST min(20, |block|) nodes of the
// block list (c_indices[i], not c_indices[selected[i]]); reproduced here.
void select_key_controls(Problem& P) {
    P.key_nodes.assign(P.blocks.size(), {});
    P.aim_center.assign(P.blocks.size(), orc::V3f{0, 0, 0});
    for (size_t b = 0; b < P.blocks.size(); b++) {
        if (P.btype[b] == -1) continue;
        const int cnt = (int)std::min<size_t>(20, P.blocks[b].size());
        std::set<int> s;
        float cx = 0, cy = 0, cz = 0;
        for (int i = 0; i < cnt; i++) {
            const int v = P.blocks[b][i];
            s.insert(v);
            cx += P.aim[3 * v]; cy += P.aim[3 * v + 1]; cz += P.aim[3 * v + 2];
        }
        P.key_nodes[b].assign(s.begin(), s.end());
        const float fc = (float)cnt;
        P.aim_center[b] = orc::V3f{cx / fc, cy / fc, cz / fc};
    }
}

void build(Problem& P, int M, int K, const float* g, const uint32_t* nbr, const int* bptr, const int* bnodes,
           const int* btype, int nblocks, const uint32_t* src_idx, const double* src_w, const float* aim,
           double w_rot, double w_reg, double w_con, double wf, int center) {
    P.M = M; P.K = K; P.g = g; P.nbr = nbr; P.src_idx = src_idx; P.src_w = src_w; P.aim = aim;
    P.sw_rot = std::sqrt(w_rot); P.sw_reg = std::sqrt(w_reg); P.sw_con = std::sqrt(w_con);
    P.wf = wf; P.center = center != 0;
    P.blocks.resize(nblocks);
    P.btype.assign(btype, btype + nblocks);
    std::vector<char> stat(M, 0);
    for (int b = 0; b < nblocks; b++) {
        P.blocks[b].assign(bnodes + bptr[b], bnodes + bptr[b + 1]);
        if (btype[b] < 0)
            for (int v : P.blocks[b]) stat[v] = 1;
    }
    P.free_id.assign(M, -1);
    for (int i = 0; i < M; i++) {
        if (stat[i]) P.static_nodes.push_back(i);
        else { P.free_id[i] = (int)P.free_nodes.size(); P.free_nodes.push_back(i); }
    }
    P.n = 12 * (int)P.free_nodes.size();
    long ctrl_nodes = 0, ctrl_blocks = 0;
    for (int b = 0; b < nblocks; b++)
        if (btype[b] != -1) { ctrl_blocks++; ctrl_nodes += (long)P.blocks[b].size(); }
    P.rows = (long)P.free_nodes.size() * (6 + 3 * K) + (long)P.static_nodes.size() * 3 * K +
             3 * (P.center ? ctrl_blocks : ctrl_nodes);
    if (P.center) select_key_controls(P);
}

// Deform::CalcEnergyFunc (Deform.cpp:378-581).
void residual(const Problem& P, const double* x, double* f) {
    const int K = P.K;
    long row = 0;
    for (size_t i = 0; i < P.free_nodes.size(); i++) {
        const double* r = x + 12 * i;
        const double* c0 = r; const double* c1 = r + 3; const double* c2 = r + 6;
        auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        f[row++] = P.sw_rot * dot(c0, c1);
        f[row++] = P.sw_rot * dot(c0, c2);
        f[row++] = P.sw_rot * dot(c1, c2);
        f[row++] = P.sw_rot * (dot(c0, c0) - 1.0);
        f[row++] = P.sw_rot * (dot(c1, c1) - 1.0);
        f[row++] = P.sw_rot * (dot(c2, c2) - 1.0);
    }
    for (size_t fi = 0; fi < P.free_nodes.size(); fi++) {
        const int i = P.free_nodes[fi];
        const double* xi = x + 12 * fi;
        const double gi[3] = {P.g[3 * i], P.g[3 * i + 1], P.g[3 * i + 2]};
        for (int t = 0; t < K; t++) {
            const int k = (int)P.nbr[(size_t)i * K + t];
            const double gk[3] = {P.g[3 * k], P.g[3 * k + 1], P.g[3 * k + 2]};
            double tk[3] = {0, 0, 0};
            if (P.free_id[k] >= 0)
                for (int c = 0; c < 3; c++) tk[c] = x[12 * P.free_id[k] + 9 + c];
            const double s = P.same_block(i, k) ? P.wf : 1.0;
            const double d[3] = {gk[0] - gi[0], gk[1] - gi[1], gk[2] - gi[2]};
            for (int r = 0; r < 3; r++) {
                const double v = xi[r] * d[0] + xi[3 + r] * d[1] + xi[6 + r] * d[2] + gi[r] + xi[9 + r] - gk[r] - tk[r];
                f[row++] = P.sw_reg * v * s;
            }
        }
    }
    for (int i : P.static_nodes) {
        for (int t = 0; t < K; t++) {
            const int k = (int)P.nbr[(size_t)i * K + t];
            for (int c = 0; c < 3; c++) {
                const double tk = P.free_id[k] >= 0 ? x[12 * P.free_id[k] + 9 + c] : 0.0;
                f[row++] = P.sw_reg * -tk;
            }
        }
    }
    auto blend_node = [&](int c, double out[3]) {
        const double ve[3] = {P.g[3 * c], P.g[3 * c + 1], P.g[3 * c + 2]};
        out[0] = out[1] = out[2] = 0.0;
        for (int j = 0; j < K; j++) {
            const int nj = (int)P.src_idx[(size_t)c * K + j];
            const double w = P.src_w[(size_t)c * K + j];
            if (P.free_id[nj] < 0) {
                for (int r = 0; r < 3; r++) out[r] += w * ve[r];
            } else {
                const double* xj = x + 12 * P.free_id[nj];
                const double gj[3] = {P.g[3 * nj], P.g[3 * nj + 1], P.g[3 * nj + 2]};
                const double d[3] = {ve[0] - gj[0], ve[1] - gj[1], ve[2] - gj[2]};
                for (int r = 0; r < 3; r++)
                    out[r] += w * (xj[r] * d[0] + xj[3 + r] * d[1] + xj[6 + r] * d[2] + gj[r] + xj[9 + r]);
            }
        }
    };
    for (size_t b = 0; b < P.blocks.size(); b++) {
        if (P.btype[b] == -1) continue;
        if (P.center) {
            double acc[3] = {0, 0, 0};
            const double ac[3] = {P.aim_center[b].x, P.aim_center[b].y, P.aim_center[b].z};
            for (int c : P.key_nodes[b]) {
                double nv[3];
                blend_node(c, nv);
                for (int r = 0; r < 3; r++) acc[r] += P.sw_con * (nv[r] - ac[r]);
            }
            for (int r = 0; r < 3; r++) f[row++] = acc[r];
        } else {
            for (int c : P.blocks[b]) {
                double nv[3];
                blend_node(c, nv);
                for (int r = 0; r < 3; r++) f[row++] = P.sw_con * (nv[r] - (double)P.aim[3 * c + r]);
            }
        }
    }
}

// Deform::FastCalcJacobiMat (Deform.cpp:180-376) as triplets (duplicates add).
void jacobian(const Problem& P, const double* x, std::vector<Trip>& T) {
    T.clear();
    const int K = P.K;
    long row = 0;
    const double w = P.sw_rot;
    for (size_t i = 0; i < P.free_nodes.size(); i++) {
        const int k = 12 * (int)i;
        const double* r = x + k;
        // rows: c0.c1, c0.c2, c1.c2
        const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (int p = 0; p < 3; p++) {
            const int a = pairs[p][0], b = pairs[p][1];
            for (int c = 0; c < 3; c++) {
                T.push_back({row + p, k + 3 * a + c, r[3 * b + c] * w});
                T.push_back({row + p, k + 3 * b + c, r[3 * a + c] * w});
            }
        }
        for (int j = 0; j < 3; j++)
            for (int c = 0; c < 3; c++) T.push_back({row + 3 + j, k + 3 * j + c, 2.0 * r[3 * j + c] * w});
        row += 6;
    }
    for (size_t fi = 0; fi < P.free_nodes.size(); fi++) {
        const int i = P.free_nodes[fi];
        const int k1 = 12 * (int)fi;
        for (int t = 0; t < K; t++) {
            const int kk = (int)P.nbr[(size_t)i * K + t];
            const double s = P.same_block(i, kk) ? P.wf : 1.0;
            const double d[3] = {(double)P.g[3 * kk] - P.g[3 * i], (double)P.g[3 * kk + 1] - P.g[3 * i + 1],
                                 (double)P.g[3 * kk + 2] - P.g[3 * i + 2]};
            for (int j = 0; j < 3; j++) {
                T.push_back({row + j, k1 + j, d[0] * P.sw_reg * s});
                T.push_back({row + j, k1 + j + 3, d[1] * P.sw_reg * s});
                T.push_back({row + j, k1 + j + 6, d[2] * P.sw_reg * s});
                T.push_back({row + j, k1 + j + 9, P.sw_reg});
                if (P.free_id[kk] >= 0) T.push_back({row + j, 12 * P.free_id[kk] + j + 9, -P.sw_reg * s});
            }
            row += 3;
        }
    }
    for (int i : P.static_nodes) {
        for (int t = 0; t < K; t++) {
            const int kk = (int)P.nbr[(size_t)i * K + t];
            if (P.free_id[kk] >= 0)
                for (int j = 0; j < 3; j++) T.push_back({row + j, 12 * P.free_id[kk] + j + 9, -P.sw_reg});
            row += 3;
        }
    }
    auto con_terms = [&](int c, long r0) {
        const double ve[3] = {P.g[3 * c], P.g[3 * c + 1], P.g[3 * c + 2]};
        for (int j = 0; j < K; j++) {
            const int nj = (int)P.src_idx[(size_t)c * K + j];
            if (P.free_id[nj] < 0) continue;
            const double wei = P.src_w[(size_t)c * K + j];
            const int k1 = 12 * P.free_id[nj];
            const double d[3] = {ve[0] - P.g[3 * nj], ve[1] - P.g[3 * nj + 1], ve[2] - P.g[3 * nj + 2]};
            for (int r = 0; r < 3; r++) {
                T.push_back({r0 + r, k1 + r, wei * d[0] * P.sw_con});
                T.push_back({r0 + r, k1 + 3 + r, wei * d[1] * P.sw_con});
                T.push_back({r0 + r, k1 + 6 + r, wei * d[2] * P.sw_con});
                T.push_back({r0 + r, k1 + 9 + r, wei * P.sw_con});
            }
        }
    };
    for (size_t b = 0; b < P.blocks.size(); b++) {
        if (P.btype[b] == -1) continue;
        if (P.center) {
            for (int c : P.key_nodes[b]) con_terms(c, row);
            row += 3;
        } else {
            for (int c : P.blocks[b]) { con_terms(c, row); row += 3; }
        }
    }
}

// Dense FP64 Cholesky solve of (J^T J) h = rhs. The reference factors the same
// matrix with Eigen's SimplicialCholesky (Deform.cpp:103-133); the solution of
// an SPD system does not depend on the factorisation used.
bool dense_solve(int n, std::vector<double>& A, std::vector<double>& b) {
    for (int j = 0; j < n; j++) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; k++) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        A[(size_t)j * n + j] = d;
#pragma omp parallel for schedule(static)
        for (int i = j + 1; i < n; i++) {
            double s = A[(size_t)i * n + j];
            const double* ai = &A[(size_t)i * n];
            const double* aj = &A[(size_t)j * n];
            for (int k = 0; k < j; k++) s -= ai[k] * aj[k];
            A[(size_t)i * n + j] = s / d;
        }
    }
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= A[(size_t)i * n + k] * b[k];
        b[i] = s / A[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int k = i + 1; k < n; k++) s -= A[(size_t)k * n + i] * b[k];
        b[i] = s / A[(size_t)i * n + i];
    }
    return true;
}

// Sparse direct solve for large systems: a block Cholesky factorisation of
// J^T J over the 12x12 node blocks, with a minimum-degree ordering of the node
// graph — the class of solver of the reference's SimplicialCholesky (sparse
// LDL^T after an approximate-minimum-degree ordering, Deform.cpp:103-133). The
// solution of an SPD system does not depend on the factorisation used.
struct BlockChol {
    int nb = 0;
    std::vector<int> pos_of;  // node -> elimination position
    std::vector<int> cp, ri;  // strictly-lower block columns by position: rows ri[cp[p]..cp[p+1]) ascending
    std::vector<double> L;    // 144 doubles (row-major 12x12) per stored block
    std::vector<double> D;    // 144 doubles per diagonal block (lower Cholesky factor after factor())

    // Exact minimum-degree elimination of the node graph (ties: lowest node
    // id). The neighbours of a node when it is eliminated are the rows of its
    // column of L, so the symbolic factorisation falls out of the ordering.
    void analyse(int nblocks, std::vector<std::vector<int>> adj) {
        nb = nblocks;
        pos_of.assign(nb, -1);
        std::vector<int> order;
        std::vector<std::vector<int>> reach(nb);
        std::vector<int> merged;
        for (int step = 0; step < nb; step++) {
            int best = -1;
            size_t bd = (size_t)-1;
            for (int i = 0; i < nb; i++)
                if (pos_of[i] < 0 && adj[i].size() < bd) { bd = adj[i].size(); best = i; }
            pos_of[best] = step;
            order.push_back(best);
            reach[best] = adj[best];
            const std::vector<int>& nv = reach[best];
            for (int u : nv) {
                merged.clear();
                std::set_union(adj[u].begin(), adj[u].end(), nv.begin(), nv.end(), std::back_inserter(merged));
                adj[u].clear();
                for (int w : merged)
                    if (w != u && w != best) adj[u].push_back(w);
            }
            adj[best].clear();
            adj[best].shrink_to_fit();
        }
        cp.assign(nb + 1, 0);
        ri.clear();
        for (int p = 0; p < nb; p++) {
            std::vector<int> rows;
            for (int u : reach[order[p]]) rows.push_back(pos_of[u]);
            std::sort(rows.begin(), rows.end());
            ri.insert(ri.end(), rows.begin(), rows.end());
            cp[p + 1] = (int)ri.size();
        }
        L.assign(ri.size() * 144, 0.0);
        D.assign((size_t)nb * 144, 0.0);
    }

    double* block(int prow, int pcol) {
        if (prow == pcol) return &D[(size_t)prow * 144];
        const int* b = &ri[cp[pcol]];
        const int* e = &ri[cp[pcol + 1]];
        const int* it = std::lower_bound(b, e, prow);
        return &L[(size_t)(it - ri.data()) * 144];
    }

    // A = sum over merged J rows of row^T row (lower triangle of the permuted matrix).
    void assemble(const std::vector<std::vector<std::pair<int, double>>>& rows) {
        std::fill(L.begin(), L.end(), 0.0);
        std::fill(D.begin(), D.end(), 0.0);
        for (const auto& r : rows)
            for (const auto& a : r)
                for (const auto& b : r) {
                    const int pa = pos_of[a.first / 12], pb = pos_of[b.first / 12];
                    if (pa < pb) continue;
                    block(pa, pb)[(a.first % 12) * 12 + b.first % 12] += a.second * b.second;
                }
    }

    // Right-looking block Cholesky: L_pp = chol(A_pp), L_up = A_up L_pp^-T,
    // A_uw -= L_up L_wp^T for u >= w in the column's rows.
    bool factor() {
        for (int p = 0; p < nb; p++) {
            double* d = &D[(size_t)p * 144];
            for (int j = 0; j < 12; j++) {
                double s = d[j * 12 + j];
                for (int k = 0; k < j; k++) s -= d[j * 12 + k] * d[j * 12 + k];
                if (!(s > 0.0)) return false;
                s = std::sqrt(s);
                d[j * 12 + j] = s;
                for (int i = j + 1; i < 12; i++) {
                    double t = d[i * 12 + j];
                    for (int k = 0; k < j; k++) t -= d[i * 12 + k] * d[j * 12 + k];
                    d[i * 12 + j] = t / s;
                }
                for (int k = j + 1; k < 12; k++) d[j * 12 + k] = 0.0;
            }
            const int c0 = cp[p], c1 = cp[p + 1];
#pragma omp parallel for schedule(static)
            for (int q = c0; q < c1; q++) {
                double* l = &L[(size_t)q * 144];
                for (int i = 0; i < 12; i++)
                    for (int j = 0; j < 12; j++) {
                        double t = l[i * 12 + j];
                        for (int k = 0; k < j; k++) t -= l[i * 12 + k] * d[j * 12 + k];
                        l[i * 12 + j] = t / d[j * 12 + j];
                    }
            }
#pragma omp parallel for schedule(dynamic, 1)
            for (int qw = c0; qw < c1; qw++) {
                const int w = ri[qw];
                const double* lw = &L[(size_t)qw * 144];
                for (int qu = qw; qu < c1; qu++) {
                    const int u = ri[qu];
                    const double* lu = &L[(size_t)qu * 144];
                    double* t = block(u, w);
                    for (int i = 0; i < 12; i++)
                        for (int j = 0; j < 12; j++) {
                            double s = 0.0;
                            for (int k = 0; k < 12; k++) s += lu[i * 12 + k] * lw[j * 12 + k];
                            t[i * 12 + j] -= s;
                        }
                }
            }
        }
        return true;
    }

    // y = A^-1 y in the permuted order (forward then backward substitution).
    void solve(std::vector<double>& y) const {
        for (int p = 0; p < nb; p++) {
            const double* d = &D[(size_t)p * 144];
            double* yp = &y[(size_t)p * 12];
            for (int i = 0; i < 12; i++) {
                double s = yp[i];
                for (int k = 0; k < i; k++) s -= d[i * 12 + k] * yp[k];
                yp[i] = s / d[i * 12 + i];
            }
            for (int q = cp[p]; q < cp[p + 1]; q++) {
                const double* l = &L[(size_t)q * 144];
                double* yu = &y[(size_t)ri[q] * 12];
                for (int i = 0; i < 12; i++) {
                    double s = 0.0;
                    for (int k = 0; k < 12; k++) s += l[i * 12 + k] * yp[k];
                    yu[i] -= s;
                }
            }
        }
        for (int p = nb - 1; p >= 0; p--) {
            double* yp = &y[(size_t)p * 12];
            for (int q = cp[p]; q < cp[p + 1]; q++) {
                const double* l = &L[(size_t)q * 144];
                const double* yu = &y[(size_t)ri[q] * 12];
                for (int k = 0; k < 12; k++) {
                    double s = 0.0;
                    for (int i = 0; i < 12; i++) s += l[i * 12 + k] * yu[i];
                    yp[k] -= s;
                }
            }
            const double* d = &D[(size_t)p * 144];
            for (int i = 11; i >= 0; i--) {
                double s = yp[i];
                for (int k = i + 1; k < 12; k++) s -= d[k * 12 + i] * yp[k];
                yp[i] = s / d[i * 12 + i];
            }
        }
    }
};

// Node adjacency of J^T J from merged rows of J.
std::vector<std::vector<int>> node_graph(int nb, const std::vector<std::vector<std::pair<int, double>>>& rows) {
    std::vector<std::vector<int>> adj(nb);
    for (const auto& r : rows)
        for (size_t a = 0; a < r.size(); a++)
            for (size_t b = 0; b < r.size(); b++) {
                const int u = r[a].first / 12, w = r[b].first / 12;
                if (u != w) adj[u].push_back(w);
            }
    for (auto& a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    return adj;
}

// When set, every Gauss-Newton iteration repeats the reference's work pattern
// (Deform.cpp:131-132: SimplicialCholesky::compute, i.e. ordering + symbolic +
// numeric factorisation, then factorize() again); the result is the same.
int g_reference_cost = 0;

// Solve (J^T J) h = rhs from merged rows of J; the ordering and the symbolic
// factorisation are computed on the first call for a structure and reused.
bool sparse_solve(int n, const std::vector<std::vector<std::pair<int, double>>>& rows, std::vector<double>& rhs,
                  BlockChol& F) {
    const int nb = n / 12;
    if (F.nb != nb || g_reference_cost) F.analyse(nb, node_graph(nb, rows));
    F.assemble(rows);
    if (!F.factor()) return false;
    if (g_reference_cost) {
        F.assemble(rows);
        if (!F.factor()) return false;
    }
    std::vector<double> y((size_t)n);
    for (int c = 0; c < n; c++) y[(size_t)F.pos_of[c / 12] * 12 + c % 12] = rhs[c];
    F.solve(y);
    for (int c = 0; c < n; c++) rhs[c] = y[(size_t)F.pos_of[c / 12] * 12 + c % 12];
    return true;
}

}  // namespace

extern "C" {

void orc_gn_reference_cost(int on) { g_reference_cost = on; }

struct orc_gn_args {
    int M, K;
    const float* node_pos;
    const uint32_t* nbr;
    const int* block_ptr;
    const int* block_nodes;
    const int* block_type;
    int nblocks;
    const uint32_t* src_idx;
    const double* src_w;
    const float* aim;
    double w_rot, w_reg, w_con, wf;
    int center;
};

long orc_gn_sizes(const orc_gn_args* a, int* n_out) {
    Problem P;
    build(P, a->M, a->K, a->node_pos, a->nbr, a->block_ptr, a->block_nodes, a->block_type, a->nblocks, a->src_idx,
          a->src_w, a->aim, a->w_rot, a->w_reg, a->w_con, a->wf, a->center);
    *n_out = P.n;
    return P.rows;
}

void orc_gn_residual(const orc_gn_args* a, const double* x, double* f) {
    Problem P;
    build(P, a->M, a->K, a->node_pos, a->nbr, a->block_ptr, a->block_nodes, a->block_type, a->nblocks, a->src_idx,
          a->src_w, a->aim, a->w_rot, a->w_reg, a->w_con, a->wf, a->center);
    residual(P, x, f);
}

// Dense Jacobian (rows x n, row-major), for finite-difference checks.
void orc_gn_jacobian_dense(const orc_gn_args* a, const double* x, double* J) {
    Problem P;
    build(P, a->M, a->K, a->node_pos, a->nbr, a->block_ptr, a->block_nodes, a->block_type, a->nblocks, a->src_idx,
          a->src_w, a->aim, a->w_rot, a->w_reg, a->w_con, a->wf, a->center);
    std::vector<Trip> T;
    jacobian(P, x, T);
    std::memset(J, 0, sizeof(double) * (size_t)P.rows * P.n);
    for (const Trip& t : T) J[(size_t)t.r * P.n + t.c] += t.v;
}

// Jacobian triplets (duplicates not merged). Call with rows==NULL to get the count.
long orc_gn_jacobian_triplets(const orc_gn_args* a, const double* x, long* rows, int* cols, double* vals) {
    Problem P;
    build(P, a->M, a->K, a->node_pos, a->nbr, a->block_ptr, a->block_nodes, a->block_type, a->nblocks, a->src_idx,
          a->src_w, a->aim, a->w_rot, a->w_reg, a->w_con, a->wf, a->center);
    std::vector<Trip> T;
    jacobian(P, x, T);
    if (rows)
        for (size_t i = 0; i < T.size(); i++) { rows[i] = T[i].r; cols[i] = T[i].c; vals[i] = T[i].v; }
    return (long)T.size();
}

// Deform::real_time_deform + optimize (Deform.cpp:77-169): x starts at identity
// rotations and zero translations; each iteration solves (J^T J) h = -J^T f,
// halves h until |f(x+h)|^2 <= |f(x)|^2 (alpha down to 1e-15), and stops when
// |h| < (|x| + 1e-6) 1e-6 if stop_rule is set. energies[it] = |f(x_it)|^2 at the
// start of iteration it; energies[iters_done] = energy of the final x.
// Returns the number of iterations run, or -1 if a factorisation failed.
int orc_gn_solve(const orc_gn_args* a, int max_iters, int stop_rule, double* x, double* energies) {
    Problem P;
    build(P, a->M, a->K, a->node_pos, a->nbr, a->block_ptr, a->block_nodes, a->block_type, a->nblocks, a->src_idx,
          a->src_w, a->aim, a->w_rot, a->w_reg, a->w_con, a->wf, a->center);
    const int n = P.n;
    const long m = P.rows;
    for (int i = 0; i < n; i++) x[i] = 0.0;
    for (int i = 0; i < n / 12; i++) x[12 * i] = x[12 * i + 4] = x[12 * i + 8] = 1.0;
    const bool sparse = n > 1536;
    std::vector<double> f(m), f1(m), xt(n), A(sparse ? 0 : (size_t)n * n), g(n);
    std::vector<Trip> T;
    BlockChol chol;
    int it = 0;
    for (; it < max_iters; it++) {
        jacobian(P, x, T);
        residual(P, x, f.data());
        // J^T J and g = -J^T f via row-grouped triplets
        std::fill(A.begin(), A.end(), 0.0);
        std::vector<std::vector<std::pair<int, double>>> jrows;
        std::fill(g.begin(), g.end(), 0.0);
        std::vector<Trip> Ts(T);
        std::stable_sort(Ts.begin(), Ts.end(), [](const Trip& p, const Trip& q) { return p.r < q.r; });
        size_t s = 0;
        while (s < Ts.size()) {
            size_t e = s;
            while (e < Ts.size() && Ts[e].r == Ts[s].r) e++;
            // merge duplicates within the row
            std::vector<std::pair<int, double>> rowv;
            for (size_t q = s; q < e; q++) rowv.push_back({Ts[q].c, Ts[q].v});
            std::sort(rowv.begin(), rowv.end());
            std::vector<std::pair<int, double>> mrg;
            for (auto& pv : rowv) {
                if (!mrg.empty() && mrg.back().first == pv.first) mrg.back().second += pv.second;
                else mrg.push_back(pv);
            }
            const double fr = f[Ts[s].r];
            for (auto& p1 : mrg) {
                g[p1.first] -= p1.second * fr;
                if (!sparse)
                    for (auto& p2 : mrg) A[(size_t)p1.first * n + p2.first] += p1.second * p2.second;
            }
            if (sparse) jrows.push_back(std::move(mrg));
            s = e;
        }
        double old_e = 0.0;
        for (long r = 0; r < m; r++) old_e += f[r] * f[r];
        energies[it] = old_e;
        std::vector<double> h(g);
        if (sparse ? !sparse_solve(n, jrows, h, chol) : !dense_solve(n, A, h)) return -1;
        double normv = 0.0;
        for (int i = 0; i < n; i++) normv += x[i] * x[i];
        normv = std::sqrt(normv);
        for (double alpha = 1.0; alpha > 1e-15; alpha *= 0.5) {
            for (int i = 0; i < n; i++) xt[i] = x[i] + h[i];
            residual(P, xt.data(), f1.data());
            double ne = 0.0;
            for (long r = 0; r < m; r++) ne += f1[r] * f1[r];
            if (ne > old_e) {
                for (int i = 0; i < n; i++) h[i] *= 0.5;
            } else {
                std::copy(xt.begin(), xt.end(), x);
                break;
            }
        }
        double normh = 0.0;
        for (int i = 0; i < n; i++) normh += h[i] * h[i];
        normh = std::sqrt(normh);
        if (stop_rule && normh < (normv + 1e-6) * 1e-6) { it++; break; }
    }
    residual(P, x, f.data());
    double e = 0.0;
    for (long r = 0; r < m; r++) e += f[r] * f[r];
    energies[it] = e;
    return it;
}

}  // extern "C"
