// Host-side construction of the Gauss-Newton problem structure (once per edit,
// i.e. whenever blocks / graph change — the per-frame inputs are only node
// positions and aims). Mirrors the bookkeeping of Deform::Deform,
// GetControlNums and BelongsToSameBlock (Deform.hpp:414-455, Deform.cpp:6-32)
// and UpdateIndicies (GaussianView.cpp:1995-2021), precomputed as flat arrays.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "solver.hpp"

namespace arapgs {

namespace {

template <typename T>
T* upload(SolverWork& W, const std::vector<T>& v, int& rc) {
    if (rc) return nullptr;
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) { rc = -2; return nullptr; }
    W.allocs.push_back(p);
    if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) rc = -2;
    return static_cast<T*>(p);
}

template <typename T>
T* alloc(SolverWork& W, size_t n, int& rc) {
    if (rc) return nullptr;
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) { rc = -2; return nullptr; }
    if (hipMemset(p, 0, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) { rc = -2; return nullptr; }
    W.allocs.push_back(p);
    return static_cast<T*>(p);
}

}  // namespace

void solver_free(SolverWork& W) {
    for (void* p : W.allocs) (void)hipFree(p);
    W.allocs.clear();
    if (W.built_ev) (void)hipEventDestroy(W.built_ev);
    W.built_ev = nullptr;
}

int solver_build(const SolverDesc& d, SolverWork& W, std::string& err) {
    const int M = d.M, K = d.K;
    if (M <= 0 || K <= 0 || K > KNN_MAX_T) { err = "solver: need M > 0 and 1 <= K <= 12"; return -1; }
    // static = nodes of blocks with type < 0 (UpdateIndicies)
    std::vector<char> is_static(M, 0);
    std::vector<std::vector<int>> node_blocks(M);
    for (int b = 0; b < d.nblocks; b++) {
        for (int q = d.block_ptr[b]; q < d.block_ptr[b + 1]; q++) {
            const int v = d.block_nodes[q];
            if (v < 0 || v >= M) { err = "solver: block node out of range"; return -1; }
            if (d.block_type[b] < 0) is_static[v] = 1;
            if (d.block_type[b] != -1) node_blocks[v].push_back(b);
        }
    }
    std::vector<int> free_nodes, static_nodes, free_id(M, -1), static_id(M, -1);
    for (int i = 0; i < M; i++) {
        if (is_static[i]) { static_id[i] = (int)static_nodes.size(); static_nodes.push_back(i); }
        else { free_id[i] = (int)free_nodes.size(); free_nodes.push_back(i); }
    }
    const int nF = (int)free_nodes.size(), nS = (int)static_nodes.size();
    if (nF == 0) { err = "solver: no free nodes"; return -1; }
    for (int64_t e = 0; e < (int64_t)M * K; e++)
        if (d.nbr[e] >= (uint32_t)M || d.src_idx[e] >= (uint32_t)M) { err = "solver: neighbour index out of range"; return -1; }
    // BelongsToSameBlock (Deform.cpp:20-32) per edge
    std::vector<float> edge_scale((size_t)M * K, 1.0f);
    for (int i = 0; i < M; i++)
        for (int t = 0; t < K; t++) {
            const int k = (int)d.nbr[(size_t)i * K + t];
            bool same = false;
            for (int b : node_blocks[i])
                if (std::find(node_blocks[k].begin(), node_blocks[k].end(), b) != node_blocks[k].end()) same = true;
            if (same) edge_scale[(size_t)i * K + t] = d.wf;
        }
    // in-edges of free nodes
    std::vector<std::vector<int>> ins(nF);
    for (int i = 0; i < M; i++)
        for (int t = 0; t < K; t++) {
            const int k = (int)d.nbr[(size_t)i * K + t];
            if (free_id[k] >= 0) ins[free_id[k]].push_back(i * K + t);
        }
    std::vector<int> in_ptr(nF + 1, 0), in_edge;
    for (int f = 0; f < nF; f++) {
        in_ptr[f + 1] = in_ptr[f] + (int)ins[f].size();
        in_edge.insert(in_edge.end(), ins[f].begin(), ins[f].end());
    }
    // constraint groups of the non-excluded blocks (type != -1), in block order:
    // plain mode, one per block node (CalcEnergyFunc, Deform.cpp:531-579);
    // centre mode, one per block (Deform.cpp:483-529) whose members are the key
    // controls of SelectKeyControls (Deform.cpp:34-75): the reference samples
    // min(20, |block|) nodes by FPS but keeps the FIRST that many block nodes
    // (c_indices[i]), as an ascending set; the aim centre averages their aims.
    std::vector<int> con_ptr{0}, con_mem, caim_ptr{0}, caim_node, multi_con;
    for (int b = 0; b < d.nblocks; b++) {
        if (d.block_type[b] == -1) continue;
        const int beg = d.block_ptr[b], end = d.block_ptr[b + 1];
        if (!d.center) {
            for (int q = beg; q < end; q++) {
                con_mem.push_back(d.block_nodes[q]);
                con_ptr.push_back((int)con_mem.size());
                caim_node.push_back(d.block_nodes[q]);
                caim_ptr.push_back((int)caim_node.size());
            }
        } else {
            if (end == beg) { err = "solver: empty control block in centre mode"; return -1; }
            const int cnt = std::min(20, end - beg);
            std::vector<int> keys(d.block_nodes + beg, d.block_nodes + beg + cnt);
            caim_node.insert(caim_node.end(), keys.begin(), keys.end());
            std::sort(keys.begin(), keys.end());
            keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
            if (keys.size() > 1) multi_con.push_back((int)con_ptr.size() - 1);
            con_mem.insert(con_mem.end(), keys.begin(), keys.end());
            con_ptr.push_back((int)con_mem.size());
            caim_ptr.push_back((int)caim_node.size());
        }
    }
    const int nC = (int)con_ptr.size() - 1;
    // memberships: one per (free node, group), listing the source-weight
    // entries member*K + slot through which the group's rows reach the node
    std::vector<std::vector<std::pair<int, int>>> cms(nF);
    for (int c = 0; c < nC; c++)
        for (int q = con_ptr[c]; q < con_ptr[c + 1]; q++)
            for (int j = 0; j < K; j++) {
                const int e = con_mem[q] * K + j;
                const int nj = (int)d.src_idx[e];
                if (free_id[nj] >= 0) cms[free_id[nj]].push_back({c, e});
            }
    std::vector<int> cm_ptr(nF + 1, 0), cm_con, cm_sbeg, cm_src;
    for (int f = 0; f < nF; f++) {
        auto& v = cms[f];
        std::stable_sort(v.begin(), v.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
            return a.first < b.first;
        });
        for (size_t q = 0; q < v.size(); q++) {
            if (q == 0 || v[q].first != v[q - 1].first) {
                cm_con.push_back(v[q].first);
                cm_sbeg.push_back((int)cm_src.size());
            }
            cm_src.push_back(v[q].second);
        }
        cm_ptr[f + 1] = (int)cm_con.size();
    }
    cm_sbeg.push_back((int)cm_src.size());
    // Assembled-operator pairs (plain mode only: a multi-member group couples
    // every pair of its sources). For free owner f: every free j != f sharing a
    // residual group with f, with that group's terms in a fixed order.
    std::vector<int> pair_ptr(nF + 1, 0), pair_j, term_ptr{0};
    std::vector<int4> terms;
    if (multi_con.empty()) {
        // constraint groups of each free node (plain mode: one member, K slots)
        std::vector<std::vector<int>> con_of(nF);
        for (int c = 0; c < nC; c++) {
            std::vector<int> seen;
            for (int q = con_ptr[c]; q < con_ptr[c + 1]; q++)
                for (int j = 0; j < K; j++) {
                    const int fj = free_id[(int)d.src_idx[(size_t)con_mem[q] * K + j]];
                    if (fj >= 0 && std::find(seen.begin(), seen.end(), fj) == seen.end()) seen.push_back(fj);
                }
            for (int f : seen) con_of[f].push_back(c);
        }
        for (int f = 0; f < nF; f++) {
            const int i = free_nodes[f];
            std::vector<std::pair<int, int4>> tl;  // (free j, term)
            for (int t = 0; t < K; t++) {          // edges (i, t)
                const int fk = free_id[(int)d.nbr[(size_t)i * K + t]];
                if (fk >= 0 && fk != f) tl.push_back({fk, make_int4(0, i, t, 0)});
            }
            for (int q = in_ptr[f]; q < in_ptr[f + 1]; q++) {  // edges (src, t) into i
                const int e = in_edge[q], src = e / K, t = e % K;
                const int fs = free_id[src];
                if (fs >= 0 && fs != f) tl.push_back({fs, make_int4(1, src, t, 0)});
            }
            for (int c : con_of[f]) {              // constraint groups reaching i
                std::vector<int> js;
                for (int q = con_ptr[c]; q < con_ptr[c + 1]; q++)
                    for (int j = 0; j < K; j++) {
                        const int fj = free_id[(int)d.src_idx[(size_t)con_mem[q] * K + j]];
                        if (fj >= 0 && fj != f && std::find(js.begin(), js.end(), fj) == js.end()) js.push_back(fj);
                    }
                for (int fj : js) tl.push_back({fj, make_int4(2, c, 0, 0)});
            }
            std::stable_sort(tl.begin(), tl.end(),
                             [](const std::pair<int, int4>& a, const std::pair<int, int4>& b) { return a.first < b.first; });
            for (size_t q = 0; q < tl.size(); q++) {
                if (q == 0 || tl[q].first != tl[q - 1].first) {
                    if (q > 0) term_ptr.push_back((int)terms.size());
                    pair_j.push_back(tl[q].first);
                }
                terms.push_back(tl[q].second);
            }
            if (!tl.empty()) term_ptr.push_back((int)terms.size());
            pair_ptr[f + 1] = (int)pair_j.size();
        }
    }
    // aggregates: farthest-point seeds over the free nodes' positions, then
    // nearest-seed assignment
    const int n_agg = std::max(1, std::min(d.n_agg > 0 ? d.n_agg : 128, nF));
    if (!multi_con.empty() && n_agg > 2048) { err = "solver: centre mode supports at most 2048 aggregates"; return -1; }
    std::vector<int> seeds;
    {
        std::vector<float> dist(nF, 3.4e38f);
        int cur = 0;
        for (int s = 0; s < n_agg; s++) {
            seeds.push_back(cur);
            const float* pc = d.node_pos + 3 * (size_t)free_nodes[cur];
            int best = 0;
            float bd = -1.0f;
            for (int f = 0; f < nF; f++) {
                const float* pf = d.node_pos + 3 * (size_t)free_nodes[f];
                const float dx = pf[0] - pc[0], dy = pf[1] - pc[1], dz = pf[2] - pc[2];
                dist[f] = std::min(dist[f], dx * dx + dy * dy + dz * dz);
                if (dist[f] > bd) { bd = dist[f]; best = f; }
            }
            cur = best;
        }
    }
    std::vector<int> agg_of(nF, 0);
    for (int f = 0; f < nF; f++) {
        const float* pf = d.node_pos + 3 * (size_t)free_nodes[f];
        float bd = 3.4e38f;
        for (int a = 0; a < n_agg; a++) {
            const float* ps = d.node_pos + 3 * (size_t)free_nodes[seeds[a]];
            const float dx = pf[0] - ps[0], dy = pf[1] - ps[1], dz = pf[2] - ps[2];
            const float dd = dx * dx + dy * dy + dz * dz;
            if (dd < bd) { bd = dd; agg_of[f] = a; }
        }
    }
    std::vector<int> agg_ptr(n_agg + 1, 0), agg_member;
    for (int a = 0; a < n_agg; a++) {
        for (int f = 0; f < nF; f++)
            if (agg_of[f] == a) agg_member.push_back(f);
        agg_ptr[a + 1] = (int)agg_member.size();
    }

    int rc = 0;
    SolverDev& S = W.dev;
    S.M = M; S.K = K; S.nF = nF; S.nS = nS; S.nC = nC;
    S.n_groups = (int64_t)nF + (int64_t)nF * K + (int64_t)nS * K + nC;
    S.n_agg = n_agg;
    S.nc_pad = ((12 * n_agg + 63) / 64) * 64;
    S.sw_rot = std::sqrt(d.w_rot); S.sw_reg = std::sqrt(d.w_reg); S.sw_con = std::sqrt(d.w_con);
    S.free_nodes = upload(W, free_nodes, rc);
    S.static_nodes = upload(W, static_nodes, rc);
    S.free_id = upload(W, free_id, rc);
    S.static_id = upload(W, static_id, rc);
    S.nbr = upload(W, std::vector<uint32_t>(d.nbr, d.nbr + (size_t)M * K), rc);
    S.edge_scale = upload(W, edge_scale, rc);
    S.src_idx = upload(W, std::vector<uint32_t>(d.src_idx, d.src_idx + (size_t)M * K), rc);
    S.src_w = upload(W, std::vector<float>(d.src_w, d.src_w + (size_t)M * K), rc);
    S.con_ptr = upload(W, con_ptr, rc);
    S.con_mem = upload(W, con_mem, rc);
    S.caim_ptr = upload(W, caim_ptr, rc);
    S.caim_node = upload(W, caim_node, rc);
    S.in_ptr = upload(W, in_ptr, rc);
    S.in_edge = upload(W, in_edge, rc);
    S.cm_ptr = upload(W, cm_ptr, rc);
    S.cm_con = upload(W, cm_con, rc);
    S.cm_sbeg = upload(W, cm_sbeg, rc);
    S.cm_src = upload(W, cm_src, rc);
    S.multi_con = upload(W, multi_con, rc);
    S.n_multi = (int)multi_con.size();
    {
        // walk tables: static parts here, per-frame parts in walk_table_kernel
        std::vector<int> out_nb((size_t)nF * K), in_nb(in_edge.size()), con_nb(con_mem.size() * (size_t)K);
        std::vector<int2> cm_rng(cm_con.size());
        for (int f = 0; f < nF; f++)
            for (int t = 0; t < K; t++) out_nb[(size_t)f * K + t] = free_id[d.nbr[(size_t)free_nodes[f] * K + t]];
        for (size_t q = 0; q < in_edge.size(); q++) in_nb[q] = free_id[in_edge[q] / K];
        for (size_t e = 0; e < cm_con.size(); e++) cm_rng[e] = make_int2(con_ptr[cm_con[e]], con_ptr[cm_con[e] + 1]);
        for (size_t m = 0; m < con_mem.size(); m++)
            for (int j = 0; j < K; j++)
                con_nb[m * K + j] = free_id[d.src_idx[(size_t)con_mem[m] * K + j]];
        std::vector<int2> out_ik((size_t)nF * K), in_is(in_edge.size());
        std::vector<int> cm_i(cm_con.size());
        for (int f = 0; f < nF; f++) {
            const int i = free_nodes[f];
            for (int t = 0; t < K; t++) out_ik[(size_t)f * K + t] = make_int2(i, (int)d.nbr[(size_t)i * K + t]);
            for (int q = in_ptr[f]; q < in_ptr[f + 1]; q++) in_is[q] = make_int2(i, in_edge[q] / K);
            for (int e = cm_ptr[f]; e < cm_ptr[f + 1]; e++) cm_i[e] = i;
        }
        S.n_in = (int)in_edge.size();
        S.n_cm = (int)cm_con.size();
        S.wt_out_ik = upload(W, out_ik, rc);
        S.wt_in_is = upload(W, in_is, rc);
        S.wt_cm_i = upload(W, cm_i, rc);
        S.n_mem = (int)con_mem.size();
        S.wt_out_nb = upload(W, out_nb, rc);
        S.wt_in_nb = upload(W, in_nb, rc);
        S.wt_cm_rng = upload(W, cm_rng, rc);
        S.wt_con_nb = upload(W, con_nb, rc);
        S.wt_out = alloc<float4>(W, std::max<size_t>(1, out_nb.size()), rc);
        S.wt_in = alloc<float4>(W, std::max<size_t>(1, in_nb.size()), rc);
        S.wt_cm = alloc<float4>(W, std::max<size_t>(1, cm_rng.size()), rc);
        S.wt_con = alloc<float4>(W, std::max<size_t>(1, con_nb.size()), rc);
    }
    S.n_pairs = (int)pair_j.size();
    S.pair_ptr = upload(W, pair_ptr, rc);
    S.pair_j = upload(W, pair_j, rc);
    S.term_ptr = upload(W, term_ptr, rc);
    S.terms = upload(W, terms, rc);
    S.agg_ptr = upload(W, agg_ptr, rc);
    S.agg_member = upload(W, agg_member, rc);
    S.agg_of = upload(W, agg_of, rc);
    S.agg_center = alloc<float>(W, 3 * (size_t)n_agg, rc);
    S.g = nullptr;
    S.aim = nullptr;

    W.n = 12 * nF;
    W.x = alloc<float>(W, W.n, rc);
    W.hsol = alloc<float>(W, W.n, rc);
    W.r = alloc<float>(W, W.n, rc);
    W.z = alloc<float>(W, W.n, rc);
    W.p = alloc<float>(W, W.n, rc);
    W.q = alloc<float>(W, W.n, rc);
    W.binv = alloc<float>(W, (size_t)144 * nF, rc);
    W.dmat = alloc<float>(W, (size_t)144 * nF, rc);
    W.pair_m = alloc<float>(W, (size_t)16 * std::max<size_t>(1, pair_j.size()), rc);
    // rc_k, rc_{k+1} and P^T q (the two-kernel CG form; the three-kernel form uses the first)
    W.rc = alloc<float>(W, 3 * (size_t)S.nc_pad, rc);
    W.acinv = alloc<float>(W, (size_t)S.nc_pad * S.nc_pad, rc);
    W.acinv_alt = alloc<float>(W, (size_t)S.nc_pad * S.nc_pad, rc);
    W.agg_center_alt = alloc<float>(W, 3 * (size_t)n_agg, rc);
    W.g_snap = alloc<float>(W, 3 * (size_t)d.M, rc);
    {
        std::vector<float> ident((size_t)W.n, 0.0f);
        for (int i = 0; i < nF; i++) ident[12 * (size_t)i] = ident[12 * (size_t)i + 4] = ident[12 * (size_t)i + 8] = 1.0f;
        W.x_ident = upload(W, ident, rc);
    }
    if (!rc && hipEventCreateWithFlags(&W.built_ev, hipEventDisableTiming) != hipSuccess) rc = -2;
    W.ac64 = alloc<unsigned long long>(W, (size_t)S.nc_pad * S.nc_pad, rc);
    W.gj_piv = alloc<float>(W, 2 * 64 * 64, rc);                  // two pivot inverses
    W.gj_col = alloc<float>(W, 4 * (size_t)S.nc_pad * 64, rc);    // column and row panels, two each
    const int nb_owner = (nF + 15) / 16;
    // p.q partials of the CG apply at the smallest CG block size (64: 4 owners per
    // workgroup) or one per aggregate (two-kernel form)
    W.partial = alloc<double>(W, std::max(std::max((nF + 3) / 4, n_agg), 1), rc);
    W.partial2 = alloc<double>(W, n_agg, rc);
    W.rzp = alloc<double>(W, 2 * (size_t)n_agg, rc);
    W.e_blocks = (int)std::min<int64_t>(512, (S.n_groups + 255) / 256);
    W.n_cand = 50;
    W.e_partial = alloc<double>(W, (size_t)W.e_blocks * W.n_cand, rc);
    W.cand_e = alloc<double>(W, W.n_cand, rc);
    W.acc_blocks = (int)std::min<int64_t>(128, std::max<int64_t>(1, (W.n + 255) / 256));
    W.acc_partial = alloc<double>(W, 2 * (size_t)W.acc_blocks, rc);
    W.warm = 0;
    W.warm_slots = 8;
    W.hwarm = alloc<float>(W, (size_t)W.warm_slots * W.n, rc);
    W.cg = alloc<CgScalars>(W, 1, rc);
    W.gs = alloc<GnScalars>(W, 1, rc);
    W.gn_done = alloc<int>(W, 1, rc);
    W.cg_done = alloc<int>(W, 1, rc);
    W.gbar = alloc<int>(W, 4, rc);
    W.energies = alloc<double>(W, GN_MAX_ITERS + 1, rc);
    W.coop_grid = 0;
    W.cg_max_iters = 100;
    W.cg_tol = 1e-4;
    if (rc) { err = "solver: device allocation failed"; solver_free(W); return rc; }
    return 0;
}

}  // namespace arapgs
