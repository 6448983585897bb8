// Nonstationary heat conduction through the C++ mirror of the reference API
// (nonstationary_heat_equation_solver_2d: compute, then calc_step), on an n x n Q4 mesh of the unit
// square: T = 0 on x = 0 and x = 1, insulated elsewhere, T0 = sin(pi x), lambda = 1, rho c = 2,
// a source s = 1 + y, bell influence (p1, r).
//
//   transient_q4 <n> <p1> <r> <tau> <steps> <out.bin>
//
// Writes the temperature after the last step ((n+1)^2 doubles) to out.bin.
#include <nonlocal/nonlocal_amd.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s n p1 r tau steps out.bin\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]), steps = std::atoi(argv[5]);
    const double p1 = std::atof(argv[2]), r = std::atof(argv[3]), tau = std::atof(argv[4]);
    using namespace nonlocal;
    try {
        mesh::mesh_2d_data d;
        auto v = [n](int i, int j) { return int32_t(j * (n + 1) + i); };
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) d.nodes.push_back({double(i) / n, double(j) / n});
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
        d.boundary_begin = {0, n, 2 * n, 3 * n, 4 * int64_t(n)};
        const auto mesh = std::make_shared<mesh::mesh_2d<double, int64_t>>(d);
        using namespace nonlocal::solver_2d::thermal;
        const parameters_2d<double> parameters = {
            {"material", {.model = {.influence = solver_2d::influence::polynomial<double>({r, r}, size_t{2}, size_t{2}, size_t{2}), .local_weight = p1},
                          .physical = {.conductivity = 1.0, .capacity = 2.0}}}};
        mesh->neighbours(mesh::find_neighbours(*mesh, mesh::influences<double>{{"material", {.radius = {r, r}}}}));
        thermal_boundaries_conditions_2d<double> conditions;
        conditions["left"] = std::make_unique<temperature_2d<double>>(0.0);
        conditions["right"] = std::make_unique<temperature_2d<double>>(0.0);
        nonstationary_heat_equation_solver_2d<double, int64_t, int64_t> solver(mesh, tau);
        solver.compute(parameters, conditions, [](const std::array<double, 2>& x) { return std::sin(M_PI * x[0]); });
        for (int k = 0; k < steps; ++k)
            solver.calc_step(conditions, [](const std::array<double, 2>& x) { return 1.0 + x[1]; });
        const auto T = solver.temperature();
        std::ofstream(argv[6], std::ios::binary).write(reinterpret_cast<const char*>(T.data()), std::streamsize(T.size() * sizeof(double)));
        std::printf("steps %d last CG iterations %ju\n", steps, solver.iterations());
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
