// Reference-shaped C++ caller of libarapgs: drives one drag frame through the
// classes of arapgs_dropin.hpp exactly as GaussianView drives the reference's
// (DeformGraph construction + setupEdges, computeWeights of each node's source
// vertex, Deform + real_time_deform, putFreeInputs, predict_mesh,
// Rasterizer::forward), reading its inputs from and writing its outputs to
// raw files in a directory. tests/test_gpu_dropin.py writes the inputs and
// checks the outputs against the oracle and the Python binding.
//
// usage: dropin_check <dir>
//   in : meta.txt  "m K nblocks weighting_factor P W H tan_fovx tan_fovy D
//        cg_max_iters cg_tol"
//        nodes.f32 (m*3) aims.f32 (m*3) bptr.i32 (nblocks+1) bnodes.i32 btype.i32
//        gpos.f32 grot.f32 gscale.f32 gopac.f32 gsh.f32 view.f32 proj.f32 campos.f32 bg.f32
//   out: nbr.u32 (m*K) knn_idx.u32 (m*(K+1)) knn_w.f64 (m*K) x.f64 (12*free)
//        pred.f32 (m*3) image.f32 (3*H*W) result.txt "energy iterations num_rendered"
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "arapgs_dropin.hpp"

using namespace arapgs::dropin;

template <typename T>
static std::vector<T> read_raw(const std::string& path, size_t count) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error("cannot open " + path);
    std::vector<T> v(count);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(sizeof(T) * count));
    if ((size_t)f.gcount() != sizeof(T) * count) throw Error("short read " + path);
    return v;
}

template <typename T>
static void write_raw(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(sizeof(T) * v.size()));
    if (!f) throw Error("cannot write " + path);
}

template <typename T>
static detail::DevBuf<T> to_device(const std::vector<T>& h) {
    detail::DevBuf<T> d;
    d.upload(h.data(), h.size());
    return d;
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: dropin_check <dir>\n";
        return 2;
    }
    const std::string dir = std::string(argv[1]) + "/";
    try {
        int m, K, nblocks, P, W, H, D;
        float wf, tanfx, tanfy;
        SolverOptions opt;
        {
            std::ifstream meta(dir + "meta.txt");
            if (!(meta >> m >> K >> nblocks >> wf >> P >> W >> H >> tanfx >> tanfy >> D >> opt.cg_max_iters >>
                  opt.cg_tol))
                throw Error("bad meta.txt");
        }
        const auto nodes_f = read_raw<float>(dir + "nodes.f32", 3 * (size_t)m);
        const auto aims_f = read_raw<float>(dir + "aims.f32", 3 * (size_t)m);
        const auto bptr = read_raw<int>(dir + "bptr.i32", (size_t)nblocks + 1);
        const auto bnodes = read_raw<int>(dir + "bnodes.i32", (size_t)bptr[nblocks]);
        const auto btype = read_raw<int>(dir + "btype.i32", (size_t)nblocks);

        Context ctx(0);

        // DeformGraph(nodes, knn_k): setupEdges runs in the constructor (Deform.hpp:62-82)
        std::vector<Node> nodes((size_t)m);
        std::vector<Pos> aim_pos((size_t)m), node_pos((size_t)m);
        for (int i = 0; i < m; i++) {
            nodes[i].Position = Pos(nodes_f[3 * i], nodes_f[3 * i + 1], nodes_f[3 * i + 2]);
            nodes[i].Vertex_index = (unsigned)i;  // each node's source vertex is the node itself
            node_pos[i] = nodes[i].Position;
            aim_pos[i] = Pos(aims_f[3 * i], aims_f[3 * i + 1], aims_f[3 * i + 2]);
        }
        DeformGraph graph(ctx, nodes, K);
        std::vector<uint32_t> nbr((size_t)m * K);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < K; j++) nbr[(size_t)i * K + j] = graph.nodes[i].Neighbor[j];
        write_raw(dir + "nbr.u32", nbr);

        // computeWeights of the source vertices (GaussianView.cpp:2880-2918)
        std::vector<std::array<unsigned int, KNN_MAX + 1>> idx;
        std::vector<std::array<double, KNN_MAX>> wts;
        graph.computeWeights(node_pos, idx, wts);
        std::vector<Vertex> verts((size_t)m);
        std::vector<uint32_t> knn_idx((size_t)m * (K + 1));
        std::vector<double> knn_w((size_t)m * K);
        for (int i = 0; i < m; i++) {
            verts[i].index = (unsigned)i;
            for (int j = 0; j <= K; j++) knn_idx[(size_t)i * (K + 1) + j] = idx[i][j];
            for (int j = 0; j < K; j++) {
                verts[i].Neighbor_Nodes[j] = idx[i][j];
                verts[i].Neighbor_Weights[j] = wts[i][j];
                knn_w[(size_t)i * K + j] = wts[i][j];
            }
        }
        write_raw(dir + "knn_idx.u32", knn_idx);
        write_raw(dir + "knn_w.f64", knn_w);

        // the edit blocks (indices_blocks + blocks_type, GaussianView.cpp:1995-2021)
        std::vector<int> blocks_ts((size_t)nblocks);
        graph.indices_blocks.assign((size_t)nblocks, {});
        for (int b = 0; b < nblocks; b++) {
            for (int j = bptr[b]; j < bptr[b + 1]; j++) graph.indices_blocks[b].push_back((unsigned)bnodes[j]);
            blocks_ts[b] = btype[b];
        }
        graph.UpdateIndicies(blocks_ts);

        // Deform + real_time_deform (GaussianView.cpp:1487-1499)
        Deform deform(verts, node_pos, graph.control_indices, graph.static_indices, graph, aim_pos, false, false, K,
                      blocks_ts, wf, opt);
        deform.real_time_deform();
        write_raw(dir + "x.f64", deform.m_x);

        // putFreeInputs + predict_mesh of every node (the node update of UpdatePosition)
        graph.putFreeInputs(deform.m_x.data(), node_pos);
        std::vector<float> pred(3 * (size_t)m);
        for (int i = 0; i < m; i++) {
            const Pos p = graph.predict_mesh(verts[i], node_pos, node_pos[i]);
            for (int c = 0; c < 3; c++) pred[3 * i + c] = p(c);
        }
        write_raw(dir + "pred.f32", pred);

        // Rasterizer::forward on device buffers (GaussianView.cpp:1067-1096)
        const int M = (D + 1) * (D + 1);
        auto gpos = to_device(read_raw<float>(dir + "gpos.f32", 3 * (size_t)P));
        auto grot = to_device(read_raw<float>(dir + "grot.f32", 4 * (size_t)P));
        auto gscale = to_device(read_raw<float>(dir + "gscale.f32", 3 * (size_t)P));
        auto gopac = to_device(read_raw<float>(dir + "gopac.f32", (size_t)P));
        auto gsh = to_device(read_raw<float>(dir + "gsh.f32", 48 * (size_t)P));
        auto view = to_device(read_raw<float>(dir + "view.f32", 16));
        auto proj = to_device(read_raw<float>(dir + "proj.f32", 16));
        auto campos = to_device(read_raw<float>(dir + "campos.f32", 3));
        auto bg = to_device(read_raw<float>(dir + "bg.f32", 3));
        detail::DevBuf<float> image(3 * (size_t)W * H);
        detail::DevBuf<int> rects(2 * (size_t)P);
        auto no_alloc = [](size_t) -> char* { return nullptr; };
        Rasterizer::use_context(ctx);
        const int num_rendered =
            Rasterizer::forward(no_alloc, no_alloc, no_alloc, P, D, M, bg.p, W, H, gpos.p, gsh.p, nullptr, gopac.p,
                                gscale.p, 1.0f, grot.p, nullptr, view.p, proj.p, campos.p, tanfx, tanfy, false,
                                image.p, nullptr, nullptr, rects.p, nullptr, nullptr, false, 1.0f);
        std::vector<float> img(3 * (size_t)W * H);
        image.download(img.data(), img.size());
        write_raw(dir + "image.f32", img);

        // a bad argument is reported as an exception, not a crash
        bool threw = false;
        try {
            Rasterizer::forward(no_alloc, no_alloc, no_alloc, P, D, M, bg.p, W, H, gpos.p, gsh.p, nullptr, gopac.p,
                                gscale.p, 1.0f, grot.p, gsh.p /* cov3D_precomp: unsupported */, view.p, proj.p,
                                campos.p, tanfx, tanfy, false, image.p);
        } catch (const Error&) {
            threw = true;
        }
        if (!threw) throw Error("an unsupported argument was accepted");

        FILE* r = std::fopen((dir + "result.txt").c_str(), "w");
        if (!r) throw Error("cannot write result.txt");
        std::fprintf(r, "%.17g %d %d\n", deform.energy, deform.iterations, num_rendered);
        std::fclose(r);
        Rasterizer::bound_context() = nullptr;
    } catch (const std::exception& e) {
        std::cerr << "dropin_check: " << e.what() << "\n";
        return 1;
    }
    std::cout << "dropin_check: ok\n";
    return 0;
}
