// Stationary nonlocal heat conduction on an n x n Q4 mesh through the C++ mirror of the reference API,
// set up the way the reference's NonLocFEM driver does it: the mesh, the parameters with a bell
// influence, the neighbour search adopted by the mesh, the boundary conditions, the solver.
//
//   heat_q4 <n> <p1> <r> <out.bin> [distance]
//
// distance: l2 (default), n1, n3, n1.5, inf (powered_distance with that exponent) or rotation
// (powered_distance_with_rotation), used by the search and the influence alike.  The mesh spans
// [0, 1]^2, or [-0.5, 0.5]^2 for the rotated ellipse.
// T = 0 on x = 0, unit inward flux on x = 1, insulated elsewhere, lambda = 1.  Writes the
// temperature, then the node fluxes x and y ((n+1)^2 doubles each) to out.bin and prints iterations
// and residuals.
#include <nonlocal/nonlocal_amd.hpp>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s n p1 r out.bin [l2|n1|n3|n1.5|inf|rotation]\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]);
    const double p1 = std::atof(argv[2]), r = std::atof(argv[3]);
    const std::string kind = argc > 5 ? argv[5] : "l2";
    using T = double;
    using I = int64_t;
    using namespace nonlocal;
    using namespace nonlocal::mesh;
    using namespace nonlocal::solver_2d::thermal;
    try {
        mesh_2d_data d;
        auto v = [n](int i, int j) { return int32_t(j * (n + 1) + i); };
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i)
                d.nodes.push_back({double(i) / n - (kind == "rotation" ? 0.5 : 0.0), double(j) / n - (kind == "rotation" ? 0.5 : 0.0)});
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int32_t id : {v(i, j), v(i + 1, j), v(i + 1, j + 1), v(i, j + 1)}) d.elements.push_back(id);
        d.group_names = {"material"};
        d.group_begin = {0, int64_t(n) * n};
        d.boundary_names = {"left", "right", "bottom", "top"};
        for (int j = 0; j < n; ++j) { d.boundary_edges.push_back(v(0, j)); d.boundary_edges.push_back(v(0, j + 1)); }
        for (int j = 0; j < n; ++j) { d.boundary_edges.push_back(v(n, j)); d.boundary_edges.push_back(v(n, j + 1)); }
        for (int i = 0; i < n; ++i) { d.boundary_edges.push_back(v(i, 0)); d.boundary_edges.push_back(v(i + 1, 0)); }
        for (int i = 0; i < n; ++i) { d.boundary_edges.push_back(v(i, n)); d.boundary_edges.push_back(v(i + 1, n)); }
        d.boundary_begin = {0, n, 2 * int64_t(n), 3 * int64_t(n), 4 * int64_t(n)};
        const auto mesh = std::make_shared<mesh_2d<T, I>>(d);

        using solver_2d::influence::polynomial;
        using size_t_or = metamath::types::size_t_or<T>;
        const std::array<T, 2> radius = {r, r};
        influence_function_2d<T> bell;
        distance_function<T> distance;
        if (kind == "rotation") {
            // an ellipse of radii r (radial) and 0.7 r (tangential) turning with the position (the
            // symmetric reading of this asymmetric kernel is positive definite here; at r / 2 it is not)
            const std::array<T, 2> ellipse = {r, T{0.7} * r};
            bell = polynomial<T, powered_distance_with_rotation<T>>(ellipse, size_t{2}, size_t{2}, size_t{2});
            distance = powered_distance_with_rotation<T>{};
            mesh->neighbours(find_neighbours(*mesh, influences<T>{{"material", {.distance = distance, .radius = ellipse}}}));
        } else {
            const size_t_or exponent = kind == "n1" ? size_t_or{size_t{1}} : kind == "n3" ? size_t_or{size_t{3}} :
                                       kind == "n1.5" ? size_t_or{T{1.5}} :
                                       kind == "inf" ? size_t_or{metamath::constants::Infinity<size_t>} : size_t_or{size_t{2}};
            bell = polynomial<T>(radius, exponent, size_t{2}, size_t{2});
            distance = powered_distance<T>{exponent};
            mesh->neighbours(find_neighbours(*mesh, influences<T>{{"material", {.distance = distance, .radius = radius}}}));
        }
        const parameters_2d<T> parameters = {
            {"material", {.model = {.influence = bell, .local_weight = p1}, .physical = {.conductivity = T{1}}}}};
        thermal_boundaries_conditions_2d<T> boundaries_conditions;
        boundaries_conditions["left"] = std::make_unique<temperature_2d<T>>(T{0});
        boundaries_conditions["right"] = std::make_unique<flux_2d<T>>(T{1});
        const auto sol = stationary_heat_equation_solver_2d<I>(mesh, parameters, boundaries_conditions, {});
        std::ofstream out(argv[4], std::ios::binary);
        auto write = [&out](const std::vector<double>& v) {
            out.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(double)));
        };
        write(sol.temperature());
        write(sol.flux()[X]);
        write(sol.flux()[Y]);
        std::printf("iterations %ju residual %.3e true_residual %.3e\n", sol.iterations, sol.residual, sol.true_residual);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
