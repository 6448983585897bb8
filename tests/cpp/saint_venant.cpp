// saint_venant.cpp of the reference (tests/solvers_2d/mechanical_2d/saint_venant.cpp:26-77) through the
// C++ mirror of its API: a 10 x 1 plate of 80 x 8 bilinear quads (h = 0.125, x in [-5, 5], y in
// [-0.5, 0.5]), E = 1, nu = 0.3; u_x = 0 on the vertical centre line, u_y = 0 on the horizontal centre
// line, x traction -4|y| on the left edge and +4|y| on the right edge, as functions of the point.
//
//   saint_venant
//
// Prints the largest |u_x| and |sigma_xx - 1| on the centre line x = 0, and the iterations.
#include <nonlocal/nonlocal_amd.hpp>

#include <cmath>
#include <cstdio>

int main() {
    using namespace nonlocal;
    try {
        const int nx = 80, ny = 8;
        const double h = 0.125;
        mesh::mesh_2d_data d;
        auto v = [nx](int i, int j) { return int32_t(j * (nx + 1) + i); };
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) d.nodes.push_back({-5.0 + i * h, -0.5 + j * h});
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                for (int32_t id : {v(i, j), v(i + 1, j), v(i + 1, j + 1), v(i, j + 1)}) d.elements.push_back(id);
        d.group_names = {"DEFAULT"};
        d.group_begin = {0, int64_t(nx) * ny};
        d.boundary_names = {"Left", "Right", "Vertical", "Horizontal"};
        for (int j = 0; j < ny; ++j) { d.boundary_edges.push_back(v(0, j)); d.boundary_edges.push_back(v(0, j + 1)); }
        for (int j = 0; j < ny; ++j) { d.boundary_edges.push_back(v(nx, j)); d.boundary_edges.push_back(v(nx, j + 1)); }
        for (int j = 0; j < ny; ++j) { d.boundary_edges.push_back(v(nx / 2, j)); d.boundary_edges.push_back(v(nx / 2, j + 1)); }
        for (int i = 0; i < nx; ++i) { d.boundary_edges.push_back(v(i, ny / 2)); d.boundary_edges.push_back(v(i + 1, ny / 2)); }
        d.boundary_begin = {0, ny, 2 * ny, 3 * ny, 3 * ny + nx};
        auto m = std::make_shared<mesh::mesh_2d<double, int64_t>>(d);

        mechanical::parameters_2d parameters;
        parameters["DEFAULT"] = mechanical::parameter_2d{model_parameters{}, mechanical::isotropic_elastic_parameters(1.0, 0.3)};
        using bc = mechanical::boundary_condition;
        mechanical::boundaries_conditions_2d conditions;
        conditions["Vertical"] = {bc::displacement(0.0), bc::pressure(0.0)};
        conditions["Horizontal"] = {bc::pressure(0.0), bc::displacement(0.0)};
        conditions["Left"] = {bc::pressure([](const std::array<double, 2>& p) { return -4 * std::abs(p[1]); }), bc::pressure(0.0)};
        conditions["Right"] = {bc::pressure([](const std::array<double, 2>& p) { return 4 * std::abs(p[1]); }), bc::pressure(0.0)};
        const auto sol = mechanical::equilibrium_equation(m, parameters, conditions, 1e-15, 100000);

        double ux = 0.0, sxx = 0.0;
        for (int j = 0; j <= ny; ++j) {
            const size_t node = size_t(v(nx / 2, j));
            ux = std::max(ux, std::abs(sol.displacement()[0][node]));
            sxx = std::max(sxx, std::abs(sol.stress()[0][node] - 1.0));
        }
        std::printf("iterations %ju ux_centre %.3e sxx_centre %.3e\n", sol.iterations, ux, sxx);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
