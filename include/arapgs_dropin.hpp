/*
 * arapgs_dropin.hpp — C++ classes with the shape of the reference's per-frame
 * interfaces, implemented over the arapgs C ABI (arapgs.h). A viewer written
 * against the reference's DeformGraph / Deform / Rasterizer::forward switches by
 * including this header and linking libarapgs.so and libamdhip64.so.
 *
 *   arapgs::dropin::DeformGraph   DeformGraph (Deform.hpp:13-376): nodes,
 *                                 back_up_nodes, free/static/control index
 *                                 sets, indices_blocks, rot / trans (column-
 *                                 major, Deform.hpp:28-37), w_rot/w_reg/w_con,
 *                                 k_nearest; setupEdges, computeWeights (batched
 *                                 too), putFreeInputs, predict_mesh.
 *   arapgs::dropin::Deform        Deform (Deform.hpp:378-455): same constructor
 *                                 arguments; real_time_deform() / optimize()
 *                                 (Deform.cpp:77-169) leave the unknowns in m_x.
 *   arapgs::dropin::Rasterizer    CudaRasterizer::Rasterizer::forward with the
 *                                 positional arguments of GaussianView.cpp:
 *                                 1067-1096 (device pointers).
 *
 * Differences a caller sees: Eigen / SIBR types are replaced by the small POD
 * types below with the same field names; errors (which the reference does not
 * report) throw arapgs::dropin::Error; the three scratch-buffer allocators of
 * Rasterizer::forward are accepted and not called (the context owns growable
 * workspaces).
 */
#ifndef ARAPGS_DROPIN_HPP
#define ARAPGS_DROPIN_HPP

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "arapgs.h"

namespace arapgs {
namespace dropin {

constexpr int KNN_MAX = 12;  // helper.hpp:48
constexpr int MAX_ITERS = 30;  // Deform.cpp:3

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

/* sibr::Vector3f stand-in: three floats, p(i) access as in the reference. */
struct Pos {
    float v[3] = {0.0f, 0.0f, 0.0f};
    Pos() = default;
    Pos(float x, float y, float z) : v{x, y, z} {}
    float& operator()(int i) { return v[i]; }
    float operator()(int i) const { return v[i]; }
    Pos operator-(const Pos& o) const { return Pos(v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]); }
    Pos operator+(const Pos& o) const { return Pos(v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]); }
};

/* helper.hpp:122-152 */
struct Vertex {
    unsigned int index = 0;
    unsigned int Neighbor_Nodes[KNN_MAX] = {};
    std::array<double, KNN_MAX> Neighbor_Weights{};
};

struct Node {
    Pos Position;
    Pos Color;
    unsigned int Neighbor[KNN_MAX] = {};
    unsigned int Vertex_index = 0;
};

struct SamplePoint {
    unsigned int aim_index = 0;
    unsigned int Neighbor_Nodes[KNN_MAX] = {};
    std::array<double, KNN_MAX> Neighbor_Weights{};
};

namespace detail {
inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

/* device buffer of T, freed on destruction */
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { resize(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void resize(size_t count) {
        if (count <= n && p) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hip_check(hipMalloc(&p, sizeof(T) * (count ? count : 1)), "hipMalloc");
        n = count;
    }
    void upload(const T* h, size_t count) {
        resize(count);
        if (count) hip_check(hipMemcpy(p, h, sizeof(T) * count, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    void download(T* h, size_t count) const {
        if (count) hip_check(hipMemcpy(h, p, sizeof(T) * count, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
};
}  // namespace detail

/* One arapgs context (device, stream). */
class Context {
public:
    explicit Context(int device = 0) {
        if (arapgs_create(device, &h_) != 0) throw Error("arapgs_create failed");
    }
    ~Context() {
        if (h_) arapgs_destroy(h_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    arapgs_ctx* get() const { return h_; }
    void check(int rc, const char* what) const {
        if (rc != 0) throw Error(std::string(what) + ": " + arapgs_last_error(h_));
    }
    void synchronize() const { check(arapgs_synchronize(h_), "synchronize"); }

    /* The context Rasterizer::forward uses (created on first use, device 0). */
    static Context& default_context() {
        static Context c(0);
        return c;
    }

private:
    arapgs_ctx* h_ = nullptr;
};

/* DeformGraph (Deform.hpp:13-376), the members the per-frame path reads. */
class DeformGraph {
public:
    std::vector<Node> nodes;
    std::vector<Node> back_up_nodes;
    std::set<unsigned int> free_indices;
    std::set<unsigned int> static_indices;
    std::set<unsigned int> control_indices;
    std::vector<std::vector<unsigned int>> indices_blocks;
    std::vector<std::array<double, 9>> rot;   // column-major, Deform.hpp:28-37
    std::vector<std::array<double, 3>> trans;
    double w_rot = 1.0, w_reg = 10.0, w_con = 100.0;
    int k_nearest = 8;

    DeformGraph(Context& ctx, std::vector<Node> nodes_in, int knn_k) : k_nearest(knn_k), ctx_(&ctx) {
        if (knn_k < 1 || knn_k > KNN_MAX) throw Error("DeformGraph: knn_k must be 1..12");
        nodes = std::move(nodes_in);
        for (unsigned int i = 0; i < nodes.size(); i++) {
            nodes[i].Color = Pos(0.6f, 0.0f, 0.0f);
            free_indices.insert(i);
        }
        back_up_nodes = nodes;
        rot.assign(nodes.size(), std::array<double, 9>{1, 0, 0, 0, 1, 0, 0, 0, 1});
        trans.assign(nodes.size(), std::array<double, 3>{0, 0, 0});
        setupEdges();
    }

    /* (K+1)-NN of every rest node, slot 0 dropped (Deform.hpp:84-95), on the GPU. */
    void setupEdges() {
        const int m = (int)back_up_nodes.size(), k = k_nearest;
        if (m < k + 1) throw Error("setupEdges: need more nodes than k_nearest");
        std::vector<float> pos(3 * (size_t)m);
        for (int i = 0; i < m; i++)
            for (int c = 0; c < 3; c++) pos[3 * i + c] = back_up_nodes[i].Position(c);
        detail::DevBuf<float> dpos;
        dpos.upload(pos.data(), pos.size());
        detail::DevBuf<uint32_t> dnbr((size_t)m * k);
        ctx_->check(arapgs_graph_edges(ctx_->get(), dpos.p, m, k, dnbr.p), "setupEdges");
        std::vector<uint32_t> nbr((size_t)m * k);
        dnbr.download(nbr.data(), nbr.size());
        for (int i = 0; i < m; i++)
            for (int j = 0; j < k; j++) {
                nodes[i].Neighbor[j] = nbr[(size_t)i * k + j];
                back_up_nodes[i].Neighbor[j] = nbr[(size_t)i * k + j];
            }
    }

    /* computeWeights (Deform.hpp:187-208) for many points at once: ids (K+1 per
     * point, slot K the normaliser) and weights (K per point). */
    void computeWeights(const std::vector<Pos>& pts, std::vector<std::array<unsigned int, KNN_MAX + 1>>& idx,
                        std::vector<std::array<double, KNN_MAX>>& weights) const {
        const int m = (int)back_up_nodes.size(), k = k_nearest;
        const int64_t q = (int64_t)pts.size();
        std::vector<float> np(3 * (size_t)m), qp(3 * (size_t)q);
        for (int i = 0; i < m; i++)
            for (int c = 0; c < 3; c++) np[3 * i + c] = back_up_nodes[i].Position(c);
        for (int64_t i = 0; i < q; i++)
            for (int c = 0; c < 3; c++) qp[3 * i + c] = pts[i](c);
        detail::DevBuf<float> dn, dq, dw((size_t)q * k);
        dn.upload(np.data(), np.size());
        dq.upload(qp.data(), qp.size());
        detail::DevBuf<uint32_t> di((size_t)q * (k + 1));
        ctx_->check(arapgs_knn(ctx_->get(), dn.p, m, dq.p, q, k, di.p, dw.p), "computeWeights");
        std::vector<uint32_t> hi((size_t)q * (k + 1));
        std::vector<float> hw((size_t)q * k);
        di.download(hi.data(), hi.size());
        dw.download(hw.data(), hw.size());
        idx.assign((size_t)q, {});
        weights.assign((size_t)q, {});
        for (int64_t i = 0; i < q; i++) {
            for (int j = 0; j <= k; j++) idx[i][j] = hi[(size_t)i * (k + 1) + j];
            for (int j = 0; j < k; j++) weights[i][j] = hw[(size_t)i * k + j];
        }
    }

    /* single-point form of the reference signature */
    void computeWeights(const Pos pos, std::array<double, KNN_MAX>& weights,
                        std::array<unsigned int, KNN_MAX + 1>& idx) const {
        std::vector<std::array<unsigned int, KNN_MAX + 1>> i1;
        std::vector<std::array<double, KNN_MAX>> w1;
        computeWeights(std::vector<Pos>{pos}, i1, w1);
        idx = i1[0];
        weights = w1[0];
    }

    /* GaussianView::UpdateIndicies (GaussianView.cpp:1996-2021): every node is
     * free except those of blocks whose type is negative, which become static. */
    void UpdateIndicies(const std::vector<int>& blocks_type) {
        if (blocks_type.size() != indices_blocks.size()) throw Error("UpdateIndicies: one type per block expected");
        free_indices.clear();
        static_indices.clear();
        control_indices.clear();
        for (unsigned int i = 0; i < nodes.size(); i++) free_indices.insert(i);
        for (size_t b = 0; b < indices_blocks.size(); b++)
            if (blocks_type[b] < 0)
                for (unsigned int n : indices_blocks[b]) {
                    free_indices.erase(n);
                    static_indices.insert(n);
                }
    }

    /* Deform.hpp:140-151 */
    void putFreeInputs(const double* x, const std::vector<Pos>&) {
        int index = 0;
        for (unsigned int n : free_indices) {
            for (int j = 0; j < 9; j++) rot[n][j] = x[index * 12 + j];
            for (int j = 0; j < 3; j++) trans[n][j] = x[index * 12 + 9 + j];
            index++;
        }
    }

    /* Deform.hpp:230-247 on the host (one point; the batched path is arapgs_blend_points). */
    Pos predict_mesh(const Vertex& v, const std::vector<Pos>&, Pos cur_mesh_v) const {
        Pos out(0.0f, 0.0f, 0.0f);
        for (int j = 0; j < k_nearest; j++) {
            const Pos position = nodes[v.Neighbor_Nodes[j]].Position;
            const Pos temp = cur_mesh_v - position;
            const auto& r = rot[v.Neighbor_Nodes[j]];
            const auto& t = trans[v.Neighbor_Nodes[j]];
            // float += double: the sum is formed in double and rounded once, as with the reference's Pos
            for (int c = 0; c < 3; c++)
                out(c) += v.Neighbor_Weights[j] *
                          (r[c] * temp(0) + r[3 + c] * temp(1) + r[6 + c] * temp(2) + t[c] + position(c));
        }
        return out;
    }

    Context& context() const { return *ctx_; }

private:
    Context* ctx_;
};

/* Settings of the linear solve inside each GN iteration (the reference solves
 * J^T J h = -J^T f exactly with Eigen; here it is preconditioned CG). Zero
 * selects the library defaults (100 iterations, relative residual 1e-4, 128
 * coarse aggregates). */
struct SolverOptions {
    int cg_max_iters = 0;
    float cg_tol = 0.0f;
    int n_aggregates = 0;
};

/* Deform (Deform.hpp:378-455): the GN solve of one frame. */
class Deform {
public:
    std::vector<double> m_x;  // 12 per free node, Deform.cpp:83-93 layout
    double energy = 0.0;      // final |f|^2 (optimize's return value)
    int iterations = 0;

    Deform(const std::vector<Vertex>& verts, const std::vector<Pos>& pos, const std::set<unsigned int>& control_idx,
           const std::set<unsigned int>& static_idx, const DeformGraph& deform_graph, std::vector<Pos> aim_pos,
           bool cons_on_center, bool show_efficiency, int knn_k, const std::vector<int>& blocks_ts,
           float weighting_factor, const SolverOptions& options = SolverOptions())
        : graph_(&deform_graph), aims_(std::move(aim_pos)) {
        (void)pos;
        (void)control_idx;
        (void)static_idx;
        (void)show_efficiency;
        if (knn_k != deform_graph.k_nearest) throw Error("Deform: knn_k differs from the graph's k_nearest");
        const auto& nodes = deform_graph.nodes;
        const int m = (int)nodes.size(), k = knn_k;
        if ((int)aims_.size() != m) throw Error("Deform: one aim position per node expected");
        std::vector<uint32_t> nbr((size_t)m * k), sidx((size_t)m * k);
        std::vector<float> sw((size_t)m * k), npos(3 * (size_t)m);
        for (int i = 0; i < m; i++) {
            const Vertex& v = verts.at(nodes[i].Vertex_index);
            for (int j = 0; j < k; j++) {
                nbr[(size_t)i * k + j] = nodes[i].Neighbor[j];
                sidx[(size_t)i * k + j] = v.Neighbor_Nodes[j];
                sw[(size_t)i * k + j] = (float)v.Neighbor_Weights[j];
            }
            for (int c = 0; c < 3; c++) npos[3 * i + c] = nodes[i].Position(c);
        }
        std::vector<int> bptr{0}, bnodes, btype;
        const auto& blocks = deform_graph.indices_blocks;
        if (blocks.size() != blocks_ts.size()) throw Error("Deform: one block type per block expected");
        for (size_t b = 0; b < blocks.size(); b++) {
            for (unsigned int n : blocks[b]) bnodes.push_back((int)n);
            bptr.push_back((int)bnodes.size());
            btype.push_back(blocks_ts[b]);
        }
        arapgs_solver_desc d = {};
        d.m = m;
        d.k = k;
        d.nbr_h = nbr.data();
        d.block_ptr_h = bptr.data();
        d.block_nodes_h = bnodes.data();
        d.block_type_h = btype.data();
        d.nblocks = (int)blocks.size();
        d.src_idx_h = sidx.data();
        d.src_w_h = sw.data();
        d.node_pos_h = npos.data();
        d.w_rot = (float)deform_graph.w_rot;
        d.w_reg = (float)deform_graph.w_reg;
        d.w_con = (float)deform_graph.w_con;
        d.weighting_factor = weighting_factor;
        d.constraints_on_center = cons_on_center ? 1 : 0;
        d.cg_max_iters = options.cg_max_iters;
        d.cg_tol = options.cg_tol;
        d.n_aggregates = options.n_aggregates;
        Context& ctx = deform_graph.context();
        ctx.check(arapgs_solver_create(ctx.get(), &d, &solver_), "Deform");
        int nfree = 0, n = 0;
        arapgs_solver_dims(solver_, &nfree, &n);
        n_ = n;
        if ((size_t)nfree != deform_graph.free_indices.size()) {
            arapgs_solver_destroy(solver_);
            solver_ = nullptr;
            throw Error("Deform: the graph's free_indices do not match its static blocks (call UpdateIndicies)");
        }
        npos_ = std::move(npos);
    }
    ~Deform() {
        if (solver_) arapgs_solver_destroy(solver_);
    }
    Deform(const Deform&) = delete;
    Deform& operator=(const Deform&) = delete;

    /* Deform.cpp:77-81 */
    void real_time_deform() { energy = optimize(); }

    /* Deform.cpp:95-169: GN from identity, at most MAX_ITERS iterations, reference stop rule. */
    double optimize() {
        Context& ctx = graph_->context();
        const int m = (int)graph_->nodes.size();
        std::vector<float> aims(3 * (size_t)m);
        for (int i = 0; i < m; i++)
            for (int c = 0; c < 3; c++) aims[3 * i + c] = aims_[i](c);
        detail::DevBuf<float> dpos, daim, dx((size_t)n_);
        dpos.upload(npos_.data(), npos_.size());
        daim.upload(aims.data(), aims.size());
        detail::DevBuf<double> de(MAX_ITERS + 1);
        detail::DevBuf<int> dit(1);
        ctx.check(arapgs_gn_solve(ctx.get(), solver_, dpos.p, daim.p, MAX_ITERS, 1, dx.p, de.p, dit.p), "optimize");
        std::vector<float> x((size_t)n_);
        std::vector<double> e(MAX_ITERS + 1);
        int it = 0;
        dx.download(x.data(), x.size());
        de.download(e.data(), e.size());
        dit.download(&it, 1);
        if (it < 0) throw Error("optimize: the solve failed");
        iterations = it;
        m_x.assign(x.begin(), x.end());
        return e[it];
    }

private:
    const DeformGraph* graph_;
    std::vector<Pos> aims_;
    std::vector<float> npos_;
    arapgs_solver* solver_ = nullptr;
    int n_ = 0;
};

/* CudaRasterizer::Rasterizer::forward, argument order of GaussianView.cpp:1067-1096.
 * Runs on the context given to use_context (else a device-0 context made on
 * first use) and on that context's stream; returns the number of (tile,
 * Gaussian) instances like the reference. */
struct Rasterizer {
    static Context*& bound_context() {
        static Context* c = nullptr;
        return c;
    }
    static void use_context(Context& c) { bound_context() = &c; }

    static int forward(std::function<char*(size_t)> geometryBuffer, std::function<char*(size_t)> binningBuffer,
                       std::function<char*(size_t)> imageBuffer, const int P, int D, int M, const float* background,
                       const int width, int height, const float* means3D, const float* shs,
                       const float* colors_precomp, const float* opacities, const float* scales,
                       const float scale_modifier, const float* rotations, const float* cov3D_precomp,
                       const float* viewmatrix, const float* projmatrix, const float* cam_pos, const float tan_fovx,
                       float tan_fovy, const bool prefiltered, float* out_color, int* surface = nullptr,
                       int* radii = nullptr, int* rects = nullptr, float* boxmin = nullptr, float* boxmax = nullptr,
                       bool to_ortho = false, float ortho_scale = 1.0f) {
        (void)geometryBuffer;
        (void)binningBuffer;
        (void)imageBuffer;
        arapgs_render_desc d = {P, D, M, background, width, height, means3D, shs, colors_precomp, opacities, scales,
                                scale_modifier, rotations, cov3D_precomp, viewmatrix, projmatrix, cam_pos, tan_fovx,
                                tan_fovy, prefiltered ? 1 : 0, out_color, surface, radii, rects, boxmin, boxmax,
                                to_ortho ? 1 : 0, ortho_scale, 16};
        Context& ctx = bound_context() ? *bound_context() : Context::default_context();
        int n = 0;
        ctx.check(arapgs_render(ctx.get(), &d, &n), "Rasterizer::forward");
        return n;
    }
};

}  // namespace dropin
}  // namespace arapgs

#endif  // ARAPGS_DROPIN_HPP
