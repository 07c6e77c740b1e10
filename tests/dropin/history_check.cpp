// Reads a deform.txt with arapgs::dropin::LoadDeformation and writes it back
// with RecordDeformation (include/arapgs_history.hpp); prints the counts.
#include <cstdio>

#include "arapgs_history.hpp"

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: history_check in.txt out.txt\n");
        return 2;
    }
    bool on_mesh = false;
    std::vector<unsigned int> ids;
    arapgs::dropin::DeformHistory h;
    try {
        arapgs::dropin::LoadDeformation(argv[1], on_mesh, ids, h);
        arapgs::dropin::RecordDeformation(argv[2], on_mesh, ids, h);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    size_t steps = 0;
    for (const auto& m : h.mouse_movements) steps += m.size();
    std::printf("nodes %zu operations %d moves %d blocks %zu steps %zu\n", ids.size(), h.total_operations,
                h.move_operations, h.block_nodes.size(), steps);
    return 0;
}
