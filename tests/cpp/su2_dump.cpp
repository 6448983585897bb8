// Reads an SU2 file with the C++ mirror's reader (mesh::read_su2, no GPU) and writes its arrays, so
// that a test can compare them with the Python reader's.
//
//   su2_dump <mesh.su2> <out.bin>
//
// out.bin: int64 N, E, nn, G, B, nb, GB; then N x 2 doubles, E x nn int32, G + 1 int64, B x nb int32,
// GB + 1 int64; the group names and boundary names follow as text lines on stdout.
#include <nonlocal/nonlocal_amd.hpp>

#include <cstdio>
#include <fstream>

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s mesh.su2 out.bin\n", argv[0]);
        return 2;
    }
    try {
        std::ifstream in(argv[1]);
        const auto d = nonlocal::mesh::read_su2(in);
        std::ofstream out(argv[2], std::ios::binary);
        auto put = [&out](const void* p, size_t bytes) { out.write(static_cast<const char*>(p), std::streamsize(bytes)); };
        const int64_t head[7] = {int64_t(d.nodes.size()), int64_t(d.elements.size()) / d.nodes_per_element, d.nodes_per_element,
                                 int64_t(d.group_names.size()), int64_t(d.boundary_edges.size()) / d.nodes_per_edge,
                                 d.nodes_per_edge, int64_t(d.boundary_names.size())};
        put(head, sizeof(head));
        put(d.nodes.data(), d.nodes.size() * sizeof(d.nodes[0]));
        put(d.elements.data(), d.elements.size() * sizeof(int32_t));
        put(d.group_begin.data(), d.group_begin.size() * sizeof(int64_t));
        put(d.boundary_edges.data(), d.boundary_edges.size() * sizeof(int32_t));
        put(d.boundary_begin.data(), d.boundary_begin.size() * sizeof(int64_t));
        for (const auto& g : d.group_names) std::printf("group %s\n", g.c_str());
        for (const auto& g : d.boundary_names) std::printf("boundary %s\n", g.c_str());
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
