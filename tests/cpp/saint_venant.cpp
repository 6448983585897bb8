// tests/solvers_2d/mechanical_2d/saint_venant.cpp of the reference, written against the C++ mirror of
// its API: the 10 x 1 plate of its SU2 mesh (80 x 8 bilinear quads, h = 0.125), E = 1, nu = 0.3;
// u_x = 0 on "Vertical", u_y = 0 on "Horizontal", x traction -4|y| on "Left" and +4|y| on "Right".
// The checks and bars are the reference's: u_x = 0 (2e-14) and sigma_xx = 1 (1.3e-9) on the centre
// line, the integrals of u_x (0, 7.5e-13) and sigma_xx (10, 1e-9).  Exit status 0 when all pass.
#include "test_utils.hpp"

#include <nonlocal/nonlocal_amd.hpp>

#include <cmath>
#include <limits>
#include <memory>
#include <sstream>

#ifndef NLF_MESH_DIR
#define NLF_MESH_DIR "tests/golden/meshes"
#endif

namespace {

using T = double;
using I = int64_t;
using namespace nonlocal;
using namespace unit_tests;
using namespace mesh;
using namespace solver_2d::mechanical;

// the reference embeds the file at build time (embedded_files/plate_10x1_h0_125_su2.h)
const std::string plate_10x1_h0_125_su2_data = file_text(NLF_MESH_DIR "/plate_10x1_h0.125.su2");

std::vector<T> get_values_on_center_line(const mesh_container_2d<T, I>& mesh, const std::vector<T>& solution) {
    std::vector<T> values;
    for (const size_t node : mesh.nodes())
        if (std::abs(mesh.node_coord(node)[X]) < std::numeric_limits<T>::epsilon())
            values.push_back(solution[node]);
    return values;
}

}  // namespace

int main() try {
    checks expect;
    std::stringstream stream{plate_10x1_h0_125_su2_data};
    const auto mesh = std::make_shared<mesh_2d<T, I>>(stream, mesh_format::SU2);
    const raw_mechanical_parameters<T> parameters = {
        {"DEFAULT", { .physical = { .elastic = isotropic_elastic_parameters<T>{ .young_modulus = 1., .poissons_ratio = 0.3 } } }}
    };
    mechanical_boundaries_conditions_2d<T> boundaries_conditions;
    boundaries_conditions["Vertical"] = {
        std::make_unique<displacement_2d<T>>(T{0}),
        std::make_unique<pressure_2d<T>>(T{0})
    };
    boundaries_conditions["Horizontal"] = {
        std::make_unique<pressure_2d<T>>(T{0}),
        std::make_unique<displacement_2d<T>>(T{0})
    };
    boundaries_conditions["Left"] = {
        std::make_unique<pressure_2d<T>>([](const std::array<T, 2>& point) { return -4 * std::abs(point[Y]); }),
        std::make_unique<pressure_2d<T>>(T{0})
    };
    boundaries_conditions["Right"] = {
        std::make_unique<pressure_2d<T>>([](const std::array<T, 2>& point) { return  4 * std::abs(point[Y]); }),
        std::make_unique<pressure_2d<T>>(T{0})
    };
    auto solution = equilibrium_equation<I>(mesh, parameters, boundaries_conditions);
    std::printf("iterations %ju\n", solution.iterations);

    double worst = 0;
    for (const T value : get_values_on_center_line(mesh->container(), solution.displacement()[X]))
        worst = std::max(worst, std::abs(value));
    expect.approx("displacement_x_on_center_line", worst, 0, 2e-14);
    worst = 0;
    for (const T value : get_values_on_center_line(mesh->container(), solution.stress()[0]))
        worst = std::max(worst, std::abs(value - 1));
    expect.approx("stress_xx_on_center_line", 1 + worst, 1, 1.3e-9);
    expect.approx("displacement_x_integral", mesh::utils::integrate(*mesh, solution.displacement()[X]), 0, 7.5e-13);
    expect.approx("stress_xx_integral", mesh::utils::integrate(*mesh, solution.stress()[0]), 10, 1e-9);
    return expect.failed ? 1 : 0;
} catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
}
