/*
 * arapgs_history.hpp — the edit history of a deformation session and its text
 * file, deform.txt, for C++ callers (host only, no GPU).
 *
 *   arapgs::dropin::DeformHistory     class DeformHistory (helper.hpp:172-197)
 *   arapgs::dropin::RecordDeformation GaussianView::RecordDeformation
 *                                     (GaussianView.cpp:4859-4935): writes the
 *                                     node type, the graph nodes' vertex ids and
 *                                     the history
 *   arapgs::dropin::LoadDeformation   GaussianView::LoadDeformation
 *                                     (GaussianView.cpp:4961-5100): reads them
 *                                     back; throws on a malformed file (the
 *                                     reference reads on regardless)
 *
 * The file is one line of tokens, every value followed by one space; floats
 * go through ostream's default formatting (6 significant digits), so a file
 * read and written again is identical byte for byte.
 */
#ifndef ARAPGS_HISTORY_HPP
#define ARAPGS_HISTORY_HPP

#include <array>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace arapgs {
namespace dropin {

struct DeformHistory {
    int total_operations = 0;
    int move_operations = 0;
    std::vector<int> operation_types;  // 0 add a block, -(i+1) remove block i, 1..4 move
    std::vector<std::vector<unsigned int>> block_nodes;
    std::vector<std::vector<std::array<float, 3>>> mouse_movements;
    std::vector<std::vector<int>> blocks_types_moves;
    std::vector<int> energy_on_centers;
    std::vector<std::array<float, 4>> twist_axis;
};

inline void WriteDeformation(std::ostream& out, bool nodes_on_mesh, const std::vector<unsigned int>& node_vertex_ids,
                             const DeformHistory& h) {
    out << (nodes_on_mesh ? 1 : 0) << " ";
    out << "Nodes: " << node_vertex_ids.size() << " ";
    for (unsigned int v : node_vertex_ids) out << v << " ";
    out << "Total_Operations: " << h.total_operations << " ";
    out << "Move_Operations: " << h.move_operations << " ";
    out << "Operation_Types: " << h.operation_types.size() << " ";
    for (int v : h.operation_types) out << v << " ";
    out << "Block_Nodes: " << h.block_nodes.size() << " ";
    for (const auto& b : h.block_nodes) {
        out << b.size() << " ";
        for (unsigned int v : b) out << v << " ";
    }
    out << "Mouse_Movements: " << h.mouse_movements.size() << " ";
    for (const auto& m : h.mouse_movements) {
        out << m.size() << " ";
        for (const auto& p : m) out << p[0] << " " << p[1] << " " << p[2] << " ";
    }
    out << "Blocks_Types_Moves: " << h.blocks_types_moves.size() << " ";
    for (const auto& b : h.blocks_types_moves) {
        out << b.size() << " ";
        for (int v : b) out << v << " ";
    }
    out << "Energy_on_Center: " << h.energy_on_centers.size() << " ";
    for (int v : h.energy_on_centers) out << v << " ";
    out << "Twist_Axis: " << h.twist_axis.size() << " ";
    for (const auto& a : h.twist_axis) out << a[0] << " " << a[1] << " " << a[2] << " " << a[3] << " ";
}

inline void RecordDeformation(const std::string& path, bool nodes_on_mesh,
                              const std::vector<unsigned int>& node_vertex_ids, const DeformHistory& h) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("RecordDeformation: cannot create " + path);
    WriteDeformation(out, nodes_on_mesh, node_vertex_ids, h);
}

inline void ReadDeformation(std::istream& in, bool& nodes_on_mesh, std::vector<unsigned int>& node_vertex_ids,
                            DeformHistory& h) {
    auto fail = [](const std::string& what) { throw std::runtime_error("deform.txt: " + what); };
    auto label = [&](const char* name) {
        std::string s;
        if (!(in >> s) || s != name) fail(std::string("expected ") + name);
    };
    auto count = [&]() {
        long long n = -1;
        if (!(in >> n) || n < 0) fail("bad count");
        return (size_t)n;
    };
    int type = 0;
    if (!(in >> type)) fail("missing node type");
    nodes_on_mesh = type == 1;
    h = DeformHistory();
    label("Nodes:");
    node_vertex_ids.resize(count());
    for (auto& v : node_vertex_ids)
        if (!(in >> v)) fail("bad node id");
    label("Total_Operations:");
    if (!(in >> h.total_operations)) fail("bad Total_Operations");
    label("Move_Operations:");
    if (!(in >> h.move_operations)) fail("bad Move_Operations");
    label("Operation_Types:");
    h.operation_types.resize(count());
    for (auto& v : h.operation_types)
        if (!(in >> v)) fail("bad operation type");
    label("Block_Nodes:");
    h.block_nodes.resize(count());
    for (auto& b : h.block_nodes) {
        b.resize(count());
        for (auto& v : b)
            if (!(in >> v)) fail("bad block node");
    }
    label("Mouse_Movements:");
    h.mouse_movements.resize(count());
    for (auto& m : h.mouse_movements) {
        m.resize(count());
        for (auto& p : m)
            if (!(in >> p[0] >> p[1] >> p[2])) fail("bad mouse movement");
    }
    label("Blocks_Types_Moves:");
    h.blocks_types_moves.resize(count());
    for (auto& b : h.blocks_types_moves) {
        b.resize(count());
        for (auto& v : b)
            if (!(in >> v)) fail("bad block type");
    }
    label("Energy_on_Center:");
    h.energy_on_centers.resize(count());
    for (auto& v : h.energy_on_centers)
        if (!(in >> v)) fail("bad Energy_on_Center value");
    label("Twist_Axis:");
    h.twist_axis.resize(count());
    for (auto& a : h.twist_axis)
        if (!(in >> a[0] >> a[1] >> a[2] >> a[3])) fail("bad twist axis");
}

inline void LoadDeformation(const std::string& path, bool& nodes_on_mesh, std::vector<unsigned int>& node_vertex_ids,
                            DeformHistory& h) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("LoadDeformation: cannot open " + path);
    ReadDeformation(in, nodes_on_mesh, node_vertex_ids, h);
}

}  // namespace dropin
}  // namespace arapgs

#endif  // ARAPGS_HISTORY_HPP
