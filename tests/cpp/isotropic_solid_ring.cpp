// tests/solvers_2d/thermal_2d/isotropic_solid_ring.cpp of the reference, written against the C++ mirror
// of its API: the solid ring (radii 0.5 and 1) of its SU2 mesh, T = 0 on "Inner" and T = 1 on "Outer",
// lambda = 1, local theory; the temperature and both flux components against the closed forms, at the
// reference's bars (1.4e-3 and 1.7e-2).  Exit status 0 when every check passes.
#include "test_utils.hpp"

#include <nonlocal/nonlocal_amd.hpp>

#include <cmath>
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
using namespace solver_2d::thermal;

// the reference embeds the file at build time (embedded_files/solid_ring_su2.h)
const std::string solid_ring_su2_data = file_text(NLF_MESH_DIR "/solid_ring.su2");

constexpr T Expected_Error = T{0};
constexpr T Inner_Temperature = T{0};
constexpr T Outer_Temperature = T{1};
constexpr T Norm_Temperature = Outer_Temperature / (Outer_Temperature - Inner_Temperature);
constexpr T Inner_Radius = T{0.5};
constexpr T Outer_Radius = T{1};
constexpr T Norm_Radius = Inner_Radius / Outer_Radius;
const T Coeff = T{1} / std::log(T{1} / Norm_Radius);

}  // namespace

int main() try {
    checks expect;
    std::stringstream stream{solid_ring_su2_data};
    const auto mesh = std::make_shared<mesh_2d<T, I>>(stream, mesh_format::SU2);
    const parameters_2d<T> parameters = {{"DEFAULT", {.physical = {.conductivity = T{1}}}}};
    thermal_boundaries_conditions_2d<T> boundaries_conditions;
    boundaries_conditions["Inner"] = std::make_unique<temperature_2d<T>>([](const std::array<T, 2>&) { return Inner_Temperature; });
    boundaries_conditions["Outer"] = std::make_unique<temperature_2d<T>>([](const std::array<T, 2>&) { return Outer_Temperature; });
    const auto solution = stationary_heat_equation_solver_2d<I>(mesh, parameters, boundaries_conditions, {});

    {   // "temperature"_test
        static constexpr auto Expected_Temperature = [](const std::array<T, 2>& point) {
            return Norm_Temperature + Coeff * std::log(std::hypot(point[X], point[Y]));
        };
        static constexpr T Epsilon = 1.4e-3;
        const T error = norm_error(solution.temperature(), mesh->container(), Expected_Temperature);
        expect.approx("temperature", error, Expected_Error, Epsilon);
    }
    {   // "flux_x"_test
        static constexpr auto Expected_Flux_X = [](const std::array<T, 2>& point) {
            const auto& [x, y] = point;
            return -x * Coeff / (x * x + y * y);
        };
        static constexpr T Epsilon = 1.7e-2;
        const T error = norm_error(solution.flux()[X], mesh->container(), Expected_Flux_X);
        expect.approx("flux_x", error, Expected_Error, Epsilon);
    }
    {   // "flux_y"_test
        static constexpr auto Expected_Flux_Y = [](const std::array<T, 2>& point) {
            const auto& [x, y] = point;
            return -y * Coeff / (x * x + y * y);
        };
        static constexpr T Epsilon = 1.7e-2;
        const T error = norm_error(solution.flux()[Y], mesh->container(), Expected_Flux_Y);
        expect.approx("flux_y", error, Expected_Error, Epsilon);
    }
    return expect.failed ? 1 : 0;
} catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
}
