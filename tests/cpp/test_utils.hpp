// Helpers of the C++ drivers in the place of the reference's test utilities: norm_error (tests/utils/
// error.hpp:22-26: max |actual - expected| / max |expected| over the nodes), a check that prints
// its result, and the text of a committed mesh file (the reference embeds its meshes at build time).
#pragma once

#include <nonlocal/nonlocal_amd.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace nonlocal::unit_tests {

template<class T, class I, class Expected>
T norm_error(const std::vector<T>& actual, const mesh::mesh_container_2d<T, I>& mesh, const Expected& function) {
    T err = 0, norm = 0;
    for (const size_t node : mesh.nodes()) {
        const T e = function(mesh.node_coord(node));
        err = std::max(err, std::abs(actual[node] - e));
        norm = std::max(norm, std::abs(e));
    }
    return err / norm;
}

inline std::string file_text(const std::string& path) {
    std::ifstream in(path);
    std::stringstream s;
    s << in.rdbuf();
    return s.str();
}

// expect(approx(value, expected, epsilon)) of boost::ut, reporting the test by name
struct checks {
    int failed = 0;
    void approx(const char* name, double value, double expected, double epsilon) {
        const bool ok = std::abs(value - expected) <= epsilon;
        std::printf("%s %s: %.3e (expected %.3e +- %.1e)\n", ok ? "PASS" : "FAIL", name, value, expected, epsilon);
        failed += !ok;
    }
};

}  // namespace nonlocal::unit_tests
