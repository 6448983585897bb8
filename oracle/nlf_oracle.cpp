// ORACLE -- test infrastructure only.
//
// FP64 OpenMP C++ restatement of NonLocFEM's 2D nonlocal hot path (VIXER13/nonlocal), used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker and the CPU
// baseline ("reference algorithm, restated").  It is never linked into or called by the product
// library.  Each function cites the reference file:line it follows.
//
// Parity status of each part (what pins it) is listed in DESIGN.md; in short:
//   * element tables, mesh precompute, centres, neighbour search, pattern: pinned by closed forms,
//     brute force and invariants (tests/test_oracle_*.py);
//   * assembly: pinned by the textbook Q4 element matrix (p1 = 1), the constant-influence rank-2
//     closed form, K*1 = 0, symmetry and dense brute force;
//   * IC(0): our own choice of incomplete Cholesky (Eigen::IncompleteCholesky is not vendored) --
//     its factor is parity unpinned; only the solution it leads to is checked.
//
// Compile with -ffp-contract=off: the neighbour predicate must be evaluated op-for-op like the
// reference (search_function.hpp:28-63 -> distance.hpp:17-23 -> norm.hpp:15-22).

#include "../include/nlf.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>
#include <omp.h>

namespace {

using i64 = int64_t;
using i32 = int32_t;

// ---------------------------------------------------------------------------------------------
// Finite-element tables.  Quadrature points follow element_2d_integrate::set_quadrature
// (element_2d_integrate.hpp:69-106): q = ix*ny + iy, x = -1 + (g + 1)*1, w = wx*1*wy*1.
// Gauss rules: gaussian_quadrature.hpp:22-40.  Bases: serendipity_2d.hpp:36-114 (p = 2/9).
// ---------------------------------------------------------------------------------------------
struct Tables2D {
    int nn = 0, nq = 0;
    std::vector<double> w, qN, qNxi, qNeta;   // [i*nq + q]
    std::vector<double> N0;                   // N_i(0, 0) (centre weights)
    std::vector<int> nearest;                 // nearest quadrature point of each node (element_2d_integrate.hpp:87-105)
};

void gauss_rule(int k, std::vector<double>& x, std::vector<double>& w) {
    if (k == 1) { x = {0.0}; w = {2.0}; }
    else if (k == 2) { const double s = std::sqrt(3.0); x = {-1.0 / s, 1.0 / s}; w = {1.0, 1.0}; }
    else if (k == 3) { const double s = std::sqrt(3.0 / 5.0); x = {-s, 0.0, s}; w = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0}; }
    else throw std::domain_error("unsupported Gauss rule");
}

// Q4 bilinear basis (serendipity<T,1>).
void q4_basis(double x, double y, double* N, double* Nx, double* Ny) {
    N[0] = (1 - x) * (1 - y) / 4; N[1] = (1 + x) * (1 - y) / 4; N[2] = (1 + x) * (1 + y) / 4; N[3] = (1 - x) * (1 + y) / 4;
    Nx[0] = -(1 - y) / 4; Nx[1] = (1 - y) / 4; Nx[2] = (1 + y) / 4; Nx[3] = -(1 + y) / 4;
    Ny[0] = -(1 - x) / 4; Ny[1] = -(1 + x) / 4; Ny[2] = (1 + x) / 4; Ny[3] = (1 - x) / 4;
}

// Q8 serendipity basis with parameter p (serendipity<T,2>, p = 2/9): closed-form derivatives.
void q8_basis(double x, double y, double* N, double* Nx, double* Ny) {
    const double p = 2.0 / 9.0;
    const double a = 9 * p - 1, b = 9 * p + 3, c = 9 * p - 5;
    // corners: s = (1 -/+ x)(1 -/+ y)/16 * (a*(1 + sx*x + sy*y) + sx*sy*b*x*y) with signs of the corner
    const int sxs[4] = {-1, 1, 1, -1}, sys[4] = {-1, -1, 1, 1};
    const int cid[4] = {0, 2, 4, 6};
    for (int k = 0; k < 4; ++k) {
        // corner (sx, sy): (1 + sx x)(1 + sy y)(a (1 - sx x - sy y) + b sx sy x y)/16, e.g. corner
        // (-1,-1) is (1-x)(1-y)((9p-1)(1+x+y) + (9p+3)xy)/16 (serendipity_2d.hpp:102)
        const double sx = sxs[k], sy = sys[k];
        const double X = 1 + sx * x, Y = 1 + sy * y;
        const double gg = a * (1 - sx * x - sy * y) + b * sx * sy * x * y;
        N[cid[k]] = X * Y * gg / 16;
        const double dgx = -a * sx + b * sx * sy * y, dgy = -a * sy + b * sx * sy * x;
        Nx[cid[k]] = (sx * Y * gg + X * Y * dgx) / 16;
        Ny[cid[k]] = (X * sy * gg + X * Y * dgy) / 16;
    }
    // midsides on y = -1 (m01) and y = +1 (m23): -(1-x^2)(1 -/+ y)(c +/- b*y)/16
    {   // m01: -(1-x^2)(1-y)(c + b y)/16
        const double f = -(1 - x * x), Y = 1 - y, g = c + b * y;
        N[1] = f * Y * g / 16; Nx[1] = (2 * x) * Y * g / 16; Ny[1] = f * (-g + Y * b) / 16;
    }
    {   // m23: -(1-x^2)(1+y)(c - b y)/16
        const double f = -(1 - x * x), Y = 1 + y, g = c - b * y;
        N[5] = f * Y * g / 16; Nx[5] = (2 * x) * Y * g / 16; Ny[5] = f * (g - Y * b) / 16;
    }
    {   // m12: -(1-y^2)(1+x)(c - b x)/16
        const double f = -(1 - y * y), X = 1 + x, g = c - b * x;
        N[3] = f * X * g / 16; Ny[3] = (2 * y) * X * g / 16; Nx[3] = f * (g - X * b) / 16;
    }
    {   // m30: -(1-y^2)(1-x)(c + b x)/16
        const double f = -(1 - y * y), X = 1 - x, g = c + b * x;
        N[7] = f * X * g / 16; Ny[7] = (2 * y) * X * g / 16; Nx[7] = f * (-g + X * b) / 16;
    }
}

void basis(int nn, double x, double y, double* N, double* Nx, double* Ny) {
    if (nn == 4) q4_basis(x, y, N, Nx, Ny); else q8_basis(x, y, N, Nx, Ny);
}

Tables2D make_tables(int nn) {
    Tables2D t;
    t.nn = nn;
    std::vector<double> g, gw;
    gauss_rule(nn == 4 ? 2 : 3, g, gw);
    const int k = int(g.size());
    t.nq = k * k;
    t.w.resize(t.nq); t.qN.resize(nn * t.nq); t.qNxi.resize(nn * t.nq); t.qNeta.resize(nn * t.nq);
    std::vector<double> px(k), py(k);
    for (int i = 0; i < k; ++i) { px[i] = -1.0 + (g[i] - (-1.0)) * 1.0; py[i] = px[i]; }
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) t.w[i * k + j] = gw[i] * 1.0 * gw[j] * 1.0;
    double N[8], Nx[8], Ny[8];
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
            basis(nn, px[i], py[j], N, Nx, Ny);
            for (int a = 0; a < nn; ++a) {
                t.qN[a * t.nq + i * k + j] = N[a];
                t.qNxi[a * t.nq + i * k + j] = Nx[a];
                t.qNeta[a * t.nq + i * k + j] = Ny[a];
            }
        }
    // nearest_qnode: the first Gauss point at the smallest Euclidean distance from node i of the
    // reference element (strict comparison keeps the first minimum)
    static const double q4_nodes[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    static const double q8_nodes[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}};
    t.nearest.assign(nn, 0);
    for (int a = 0; a < nn; ++a) {
        const double* node = nn == 4 ? q4_nodes[a] : q8_nodes[a];
        double length = std::numeric_limits<double>::max();
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) {
                const double dx = node[0] - px[i], dy = node[1] - py[j];
                const double curr = std::sqrt(dx * dx + dy * dy);
                if (length > curr) { length = curr; t.nearest[a] = i * k + j; }
            }
    }
    basis(nn, 0.0, 0.0, N, Nx, Ny);
    t.N0.assign(N, N + nn);
    if (nn == 8) {   // exact reference weights (9p-1)/16 and -(9p-5)/16 at xi0 = 0
        const double p = 2.0 / 9.0;
        for (int a = 0; a < 8; a += 2) { t.N0[a] = (9 * p - 1) / 16; t.N0[a + 1] = -(9 * p - 5) / 16; }
    }
    return t;
}

// 1D boundary elements: linear with Gauss 1, quadratic (nodes -1, 0, 1) with Gauss 2
// (vtk_elements_set.hpp:43-52; lagrangian_elements_1d.hpp).
struct Tables1D { int nb = 0, nq = 0; std::vector<double> w, qN, qNxi; };

Tables1D make_tables_1d(int nb) {
    Tables1D t;
    t.nb = nb;
    std::vector<double> g, gw;
    gauss_rule(nb == 2 ? 1 : 2, g, gw);
    t.nq = int(g.size());
    t.w = gw;
    t.qN.resize(nb * t.nq); t.qNxi.resize(nb * t.nq);
    for (int q = 0; q < t.nq; ++q) {
        const double x = g[q];
        if (nb == 2) {
            t.qN[0 * t.nq + q] = (1 - x) / 2; t.qN[1 * t.nq + q] = (1 + x) / 2;
            t.qNxi[0 * t.nq + q] = -0.5; t.qNxi[1 * t.nq + q] = 0.5;
        } else {
            t.qN[0 * t.nq + q] = x * (x - 1) / 2; t.qN[1 * t.nq + q] = 1 - x * x; t.qN[2 * t.nq + q] = x * (x + 1) / 2;
            t.qNxi[0 * t.nq + q] = x - 0.5; t.qNxi[1 * t.nq + q] = -2 * x; t.qNxi[2 * t.nq + q] = x + 0.5;
        }
    }
    return t;
}

// ---------------------------------------------------------------------------------------------
// Mesh + precompute (mesh_2d::init, mesh_2d.hpp:96-108; mesh_container_2d_utils.hpp:17-132)
// ---------------------------------------------------------------------------------------------
struct Mesh {
    i64 N = 0, E = 0;
    int nn = 0, nb = 0;
    std::vector<double> xy;              // 2N
    std::vector<i32> el;                 // E*nn
    std::vector<i64> gbeg;               // G+1
    std::vector<i32> be;                 // B*nb
    std::vector<i64> bbeg;               // GB+1
    Tables2D t;
    Tables1D t1;
    std::vector<i64> ne_off;             // node -> elements (ascending)
    std::vector<i32> ne_ids;
    std::vector<double> xq;              // 2*E*nq
    std::vector<double> jac;             // E*nq
    std::vector<double> D;               // 2*E*nn*nq, layout [e][i][q] (mesh_2d.hpp:161-179)
    std::vector<double> centres;         // 2E
    // neighbours
    std::vector<i64> nb_off;
    std::vector<i32> nb_ids;
    std::vector<double> expanded;        // 2G
    std::vector<int> group_of;           // E

    int G() const { return int(gbeg.size()) - 1; }
    const i32* nodes(i64 e) const { return &el[size_t(e) * nn]; }
    int local(i64 e, i64 node) const {
        for (int i = 0; i < nn; ++i) if (el[size_t(e) * nn + i] == node) return i;
        return -1;
    }
    double Dx(i64 e, int i, int q) const { return D[2 * ((size_t(e) * nn + i) * t.nq + q)]; }
    double Dy(i64 e, int i, int q) const { return D[2 * ((size_t(e) * nn + i) * t.nq + q) + 1]; }
    const double* xqp(i64 e, int q) const { return &xq[2 * (size_t(e) * t.nq + q)]; }
};

void mesh_init(Mesh& m) {
    m.t = make_tables(m.nn);
    m.t1 = make_tables_1d(m.nb);
    const int nq = m.t.nq, nn = m.nn;
    // node_elements_2d (mesh_container_2d_utils.hpp:17-26)
    m.ne_off.assign(m.N + 1, 0);
    for (i64 e = 0; e < m.E; ++e) for (int i = 0; i < nn; ++i) m.ne_off[m.el[e * nn + i] + 1]++;
    for (i64 n = 0; n < m.N; ++n) m.ne_off[n + 1] += m.ne_off[n];
    m.ne_ids.resize(m.ne_off[m.N]);
    std::vector<i64> fill(m.ne_off.begin(), m.ne_off.end() - 1);
    for (i64 e = 0; e < m.E; ++e) for (int i = 0; i < nn; ++i) m.ne_ids[fill[m.el[e * nn + i]]++] = i32(e);
    m.group_of.resize(m.E);
    for (int g = 0; g < m.G(); ++g) for (i64 e = m.gbeg[g]; e < m.gbeg[g + 1]; ++e) m.group_of[e] = g;
    // quad coords, Jacobi matrices, |det J|, derivatives (mesh_container_2d_utils.hpp:75-123;
    // element_data_2d::quad_coord / jacobi_matrix, mesh_container_2d.hpp:141-160)
    m.xq.assign(2 * m.E * nq, 0.0);
    m.jac.assign(m.E * nq, 0.0);
    m.D.assign(2 * m.E * nn * nq, 0.0);
    m.centres.assign(2 * m.E, 0.0);
#pragma omp parallel for schedule(static)
    for (i64 e = 0; e < m.E; ++e) {
        const i32* v = m.nodes(e);
        for (int q = 0; q < nq; ++q) {
            double cx = 0, cy = 0, J00 = 0, J01 = 0, J10 = 0, J11 = 0;
            for (int i = 0; i < nn; ++i) {
                const double x = m.xy[2 * v[i]], y = m.xy[2 * v[i] + 1];
                const double N = m.t.qN[i * nq + q], Nx = m.t.qNxi[i * nq + q], Ny = m.t.qNeta[i * nq + q];
                cx = cx + x * N; cy = cy + y * N;
                J00 = J00 + x * Nx; J01 = J01 + x * Ny;
                J10 = J10 + y * Nx; J11 = J11 + y * Ny;
            }
            m.xq[2 * (e * nq + q)] = cx; m.xq[2 * (e * nq + q) + 1] = cy;
            m.jac[e * nq + q] = std::abs(J00 * J11 - J01 * J10);
            for (int i = 0; i < nn; ++i) {
                const double Nx = m.t.qNxi[i * nq + q], Ny = m.t.qNeta[i * nq + q];
                m.D[2 * ((e * nn + i) * nq + q)] = Nx * J11 - Ny * J10;
                m.D[2 * ((e * nn + i) * nq + q) + 1] = -Nx * J01 + Ny * J00;
            }
        }
        // centre (mesh_container_2d.hpp:130-139): xi0 = 0 for every quad
        double cx = 0, cy = 0;
        for (int i = 0; i < nn; ++i) {
            cx = cx + m.xy[2 * v[i]] * m.t.N0[i];
            cy = cy + m.xy[2 * v[i] + 1] * m.t.N0[i];
        }
        m.centres[2 * e] = cx; m.centres[2 * e + 1] = cy;
    }
}

// ---------------------------------------------------------------------------------------------
// Neighbour search (find_neighbours.hpp:49-160)
// ---------------------------------------------------------------------------------------------
double powered_distance2(const double* a, const double* b) {          // distance.hpp:7-10, norm.hpp:15-22
    const double z0 = a[0] - b[0], z1 = a[1] - b[1];
    double s = 0.0;
    s = s + z0 * z0;
    s = s + z1 * z1;
    return s;
}

double powered_distance2r(const double* a, const double* b, const double* r) {   // distance.hpp:17-23
    const double z0 = (a[0] - b[0]) / r[0];
    const double z1 = (a[1] - b[1]) / r[1];
    double s = 0.0;
    s = s + z0 * z0;
    s = s + z1 * z1;
    return s;
}

double ipow(double x, i64 e);

// powered_distance (search_function.hpp:28-73) with n held like metamath::types::size_t_or<T>
// (an integer, Infinity<size_t> among them, or a real), and powered_distance_with_rotation (:75-90).
// x: the point the distance is measured from (centre of eL / quadrature point of eL), y: the other,
// r: the radii.
struct Distance {
    int kind = NLF_DISTANCE_LP;
    bool real = false, infinity = false;
    i64 n_int = 2;
    double n_real = 0.0;

    double operator()(const double* x, const double* y, const double* r) const {
        if (kind == NLF_DISTANCE_ELLIPSE_WITH_ROTATION) {
            const auto square = [](double v) { return v * v; };            // power<2> (power.hpp:9-23)
            double norm_x = 0.0;                                            // powered_norm<2>(x) (norm.hpp:15-22)
            norm_x = norm_x + square(x[0]);
            norm_x = norm_x + square(x[1]);
            return (square(r[0] * (x[0] * y[1] - x[1] * y[0])) + square(r[1] * (x[0] * (x[0] - y[0]) + x[1] * (x[1] - y[1])))) /
                   (square(r[0] * r[1]) * norm_x);
        }
        if (!real && !infinity && n_int == 2) return powered_distance2r(x, y, r);   // calculate<2>
        const double z[2] = {(x[0] - y[0]) / r[0], (x[1] - y[1]) / r[1]};   // distance.hpp:17-31
        if (infinity) {                                                     // distance<Infinity> -> norm<Infinity> (norm.hpp:29-40)
            const double* it = std::max_element(z, z + 2, [](double a, double b) { return std::abs(a) < std::abs(b); });
            return std::abs(*it);
        }
        double sum = 0.0;                                                   // powered_norm (norm.hpp:15-29)
        for (const double v : z) {
            if (real) sum = sum + std::pow(std::abs(v), n_real);            // power(x, floating exp) = std::pow
            else if (n_int == 1) sum = sum + std::abs(v);                   // power<1>(|x|)
            else sum = sum + ipow(std::abs(v), n_int);                      // power(|x|, integral exp)
        }
        return sum;
    }
};

// read_distance_2d / read_influence_parameter (read_model.hpp:97-131): n > 0 or infinity; the
// rotated ellipse ignores n.
Distance make_distance(int kind, int n, double n_real) {
    Distance d;
    d.kind = kind;
    if (kind == NLF_DISTANCE_ELLIPSE_WITH_ROTATION) return d;
    if (kind != NLF_DISTANCE_LP) throw std::domain_error("unknown distance");
    if (!(n_real >= 0.0) || !std::isfinite(n_real)) throw std::domain_error("distance exponent");
    if (n_real > 0.0) { d.real = true; d.n_real = n_real; return d; }
    if (n == NLF_DISTANCE_N_INFINITY) { d.infinity = true; return d; }
    if (n < 0) throw std::domain_error("distance exponent");
    d.n_int = n == 0 ? 2 : n;
    return d;
}

double expansion(const Mesh& m, int g, int strategy) {                  // find_neighbours.hpp:49-72
    if (strategy == NLF_EXPAND_NO) return 0.0;
    double diameter = strategy == NLF_EXPAND_MIN ? std::numeric_limits<double>::max() : 0.0;
    std::set<std::pair<i64, i64>> pairs;
    for (i64 e = m.gbeg[g]; e < m.gbeg[g + 1]; ++e)
        for (int i = 0; i < m.nn; ++i) {
            const i64 node = m.el[e * m.nn + i];
            for (i64 k = m.ne_off[node]; k < m.ne_off[node + 1]; ++k) {
                const i64 a = m.ne_ids[k];
                const double d = powered_distance2(&m.centres[2 * e], &m.centres[2 * a]);
                if (strategy == NLF_EXPAND_MAX) diameter = std::max(diameter, d);
                else if (strategy == NLF_EXPAND_MIN) diameter = std::min(diameter, d);
                else if (!pairs.count({e, a}) && !pairs.count({a, e})) { pairs.insert({e, a}); diameter += std::sqrt(d); }
            }
        }
    return strategy == NLF_EXPAND_MEAN ? diameter / double(pairs.size()) : std::sqrt(diameter);
}

i64 find_neighbours(Mesh& m, const double* radii, const nlf_distance* distances, int strategy, double* expanded_out) {
    m.nb_off.assign(m.E + 1, 0);
    std::vector<std::vector<i32>> lists(m.E);
    m.expanded.assign(2 * m.G(), 0.0);
    for (int g = 0; g < m.G(); ++g) {
        const double add = expansion(m, g, strategy);                   // increase_influence :74-84
        const double r[2] = {radii[2 * g] + add, radii[2 * g + 1] + add};
        m.expanded[2 * g] = r[0]; m.expanded[2 * g + 1] = r[1];
        if (radii[2 * g] <= 0.0 || radii[2 * g + 1] <= 0.0) {
            // A group whose influence radius is 0 is local: the reference's influences map only holds
            // nonlocal groups after fix_parameters (read_model.hpp:186-192), so it gets no lists.
            m.expanded[2 * g] = 0.0; m.expanded[2 * g + 1] = 0.0;
            continue;
        }
        if (r[0] <= 0.0 || r[1] <= 0.0) continue;                       // :131-134
        const Distance distance = distances ? make_distance(distances[g].kind, distances[g].n, distances[g].n_real) : Distance{};
        // group_corners :40-47
        double left = std::numeric_limits<double>::max(), down = left;
        double right = std::numeric_limits<double>::lowest(), up = right;
        for (i64 e = m.gbeg[g]; e < m.gbeg[g + 1]; ++e)
            for (int i = 0; i < m.nn; ++i) {
                const i64 node = m.el[e * m.nn + i];
                left = std::min(left, m.xy[2 * node]); down = std::min(down, m.xy[2 * node + 1]);
                right = std::max(right, m.xy[2 * node]); up = std::max(up, m.xy[2 * node + 1]);
            }
        const double R = std::max(r[0], r[1]);
        const size_t rows = size_t((right - left) / R) + 1, cols = size_t((up - down) / R) + 1;
        std::vector<std::vector<i32>> cells(rows * cols);               // split_elements_by_subareas :95-107
        auto cell_of = [&](i64 e) {
            const size_t a = size_t((m.centres[2 * e] - left) / R);
            const size_t b = size_t((m.centres[2 * e + 1] - down) / R);
            return std::pair<size_t, size_t>{a, b};
        };
        for (i64 e = m.gbeg[g]; e < m.gbeg[g + 1]; ++e) {
            const auto [a, b] = cell_of(e);
            cells[a * cols + b].push_back(i32(e));
        }
        for (i64 eL = m.gbeg[g]; eL < m.gbeg[g + 1]; ++eL) {             // :140-157
            const auto [row, col] = cell_of(eL);
            const size_t r0 = row ? row - 1 : row, r1 = row != rows - 1 ? row + 2 : row + 1;
            const size_t c0 = col ? col - 1 : col, c1 = col != cols - 1 ? col + 2 : col + 1;
            auto& out = lists[eL];
            for (size_t i = r0; i < r1; ++i)
                for (size_t j = c0; j < c1; ++j)
                    for (const i32 eNL : cells[i * cols + j])
                        if (distance(&m.centres[2 * eL], &m.centres[2 * eNL], r) <= 1.0)
                            out.push_back(eNL);
        }
    }
    for (i64 e = 0; e < m.E; ++e) m.nb_off[e + 1] = m.nb_off[e] + i64(lists[e].size());
    m.nb_ids.resize(m.nb_off[m.E]);
    for (i64 e = 0; e < m.E; ++e) std::copy(lists[e].begin(), lists[e].end(), m.nb_ids.begin() + m.nb_off[e]);
    if (expanded_out) std::copy(m.expanded.begin(), m.expanded.end(), expanded_out);
    return m.nb_off[m.E];
}

// ---------------------------------------------------------------------------------------------
// Influence functions (influence_functions_2d.hpp:39-148), selected like read_influence_2d
// (read_model.hpp:144-165): polynomial with n = p = 2, q = 1 -> fast_polynomial.
// ---------------------------------------------------------------------------------------------
struct Influence {
    int kind = NLF_INFLUENCE_POLYNOMIAL;
    bool fast = false;
    double r[2] = {0, 0};
    double norm = 1.0;
    // polynomial exponent p/n and q: integer if exactly integral (metamath size_t_or)
    bool p_int = true, q_int = true;
    i64 p_i = 1, q_i = 1;
    double p_d = 1.0, q_d = 1.0;
    double exp_q = 0.5;
    Distance distance;
};

bool is_integral(double v) { return v >= 0 && std::floor(v) == v && v < 1e9; }

double ipow(double x, i64 e) {                       // metamath/functions/power.hpp:25-40 (integral exponent)
    if (e < 0) return 1.0 / ipow(x, -e);
    double result = 1.0;
    while (e > 0) {
        if (e & 1) { result *= x; --e; }
        else { x *= x; e >>= 1; }
    }
    return result;
}

double exp_pow(double v, bool is_int, i64 ei, double ed) {   // _influence_function_2d::power (influence_functions_2d.hpp:19-27)
    if (!is_int) return std::pow(v, ed);
    if (ei == 1) return v;
    return ipow(v, ei);
}

// The influence of a model (read_influence_2d, read_model.hpp:144-165; influence_functions_2d.hpp).
// n enters as get_value(n): Infinity<size_t> converted to double.  The exponent is
// _influence_function_2d::division(p, n) (:8-17).
Influence make_influence(const nlf_model& md) {
    Influence f;
    f.kind = md.influence;
    f.r[0] = md.radius[0]; f.r[1] = md.radius[1];
    f.distance = make_distance(md.distance, md.distance_n, md.distance_real);
    const Distance& d = f.distance;
    const bool rotation = d.kind == NLF_DISTANCE_ELLIPSE_WITH_ROTATION;   // its n is 2zu (search_function.hpp:77)
    const bool infinity = !rotation && d.infinity;
    const bool real = !rotation && d.real;
    const i64 n_int = rotation ? 2 : d.n_int;
    const double n = real ? d.n_real : infinity ? double(std::numeric_limits<size_t>::max()) : double(n_int);
    const double p = md.p, q = md.q;
    auto division = [&](double v, bool& is_int, i64& vi, double& vd) {
        if (infinity) {                                         // "used infinity norm": p unchanged
            is_int = is_integral(v);
            if (is_int) vi = i64(v); else vd = v;
        } else if (!real && is_integral(v) && i64(v) % n_int == 0) {
            is_int = true; vi = i64(v) / n_int;
        } else {
            is_int = false; vd = v / n;
        }
    };
    if (f.kind == NLF_INFLUENCE_CONSTANT) {                                    // :44-48
        f.norm = std::tgamma(1 + 2 / n) / (4 * f.r[0] * f.r[1] * std::pow(std::tgamma(1 + 1 / n), 2));
    } else if (f.kind == NLF_INFLUENCE_POLYNOMIAL) {
        if (!real && !infinity && n_int == 2 && p == 2.0 && q == 1.0) {       // fast_polynomial (read_model.hpp:160-161)
            f.fast = true;
            f.norm = 2.0 / (M_PI * f.r[0] * f.r[1]);
        } else {                                                               // :70-75
            division(p, f.p_int, f.p_i, f.p_d);
            if (is_integral(q)) { f.q_int = true; f.q_i = i64(q); } else { f.q_int = false; f.q_d = q; }
            f.norm = infinity ? p / (8 * f.r[0] * f.r[1] * std::beta(2 / p, q + 1))
                              : n * p / (4 * f.r[0] * f.r[1] * std::beta(1 / n, 1 / n) * std::beta(2 / p, q + 1));
        }
    } else if (f.kind == NLF_INFLUENCE_EXPONENTIAL) {                          // :103-109
        division(p, f.p_int, f.p_i, f.p_d);
        f.exp_q = -q;
        f.norm = infinity ? p * std::pow(q, 2 / p) / (8 * f.r[0] * f.r[1] * std::tgamma(2 / p))
                          : (n * p * std::pow(4.0, 1 / n) * std::pow(q, 2 / p)) /
                                (8 * f.r[0] * f.r[1] * std::tgamma(2 / p) * std::beta(0.5, 1 / n));
    } else {
        throw std::domain_error("unknown influence");
    }
    return f;
}

double influence_eval(const Influence& f, const double* x, const double* y) {
    const double dist = f.distance(x, y, f.r);
    if (f.kind == NLF_INFLUENCE_CONSTANT) return dist <= 1.0 ? f.norm : 0.0;
    if (f.kind == NLF_INFLUENCE_EXPONENTIAL) return f.norm * std::exp(f.exp_q * exp_pow(dist, f.p_int, f.p_i, f.p_d));
    if (f.fast) return dist < 1.0 ? f.norm * (1.0 - dist) : 0.0;
    return dist < 1.0 ? f.norm * exp_pow(1.0 - exp_pow(dist, f.p_int, f.p_i, f.p_d), f.q_int, f.q_i, f.q_d) : 0.0;
}

// ---------------------------------------------------------------------------------------------
// CSR (Eigen::SparseMatrix<T, RowMajor, int64> layout: outer offsets + inner indices + values)
// ---------------------------------------------------------------------------------------------
struct CSR {
    i64 rows = 0, cols = 0;
    std::vector<i64> outer;
    std::vector<i64> inner;
    std::vector<double> val;
    double& coeffRef(i64 r, i64 c) {     // binary search within the row (Eigen compressed coeffRef)
        auto b = inner.begin() + outer[r], e = inner.begin() + outer[r + 1];
        auto it = std::lower_bound(b, e, c);
        if (it == e || *it != c) throw std::logic_error("coeffRef outside the pattern");
        return val[size_t(it - inner.begin())];
    }
    double coeff(i64 r, i64 c) const {
        auto b = inner.begin() + outer[r], e = inner.begin() + outer[r + 1];
        auto it = std::lower_bound(b, e, c);
        return (it == e || *it != c) ? 0.0 : val[size_t(it - inner.begin())];
    }
};

enum Part { INNER = 0, BOUND = 1, NO = 2 };

struct Separator {                                   // matrix_separator_base::part :50-58
    const std::vector<uint8_t>* inner;
    bool symmetric;
    Part part(i64 row, i64 col) const {
        if ((*inner)[col]) {
            if ((!symmetric || row <= col) && (*inner)[row]) return INNER;
        } else if (row != col) return BOUND;
        return NO;
    }
};

struct Group {
    bool nonlocal = false;
    double p1 = 1.0;
    // coefficients per quadrature point (evaluated_parameters' Variable alternative,
    // equation_parameters.hpp; evaluate_conductivity.hpp:16-37): cq[((e - begin) * nq + q) * ncomp + k]
    // in the layout of the constant (conductivity c[], or the Hooke matrix of the kind), or null
    const double* cq = nullptr;
    i64 begin = 0;
    int ncomp = 0;
    const double* coeff(i64 e, int q, int nq, const double* constant) const {
        return cq ? cq + ((e - begin) * nq + q) * ncomp : constant;
    }
    Influence inf;
    nlf_thermal_material th{};
    // Hooke matrix in the reference's index order of its kind (elastic_parameters.hpp:9-11):
    // isotropic {11, 12, 66}, orthotropic {11, 12, 66, 22}, anisotropic {11, 12, 66, 22, 16, 26}
    int hkind = NLF_ELASTIC_ISOTROPIC;
    double hooke[6] = {0, 0, 0, 0, 0, 0};
    int ekind = NLF_EXPANSION_NONE;  // thermal expansion (thermal_expansion_parameters.hpp:7-18)
    double alpha[3] = {0, 0, 0};
    double capacity = 0.0;         // density * capacity (heat capacity matrix)
};

struct Assembly {
    int dof = 1;
    bool symmetric = true;
    CSR A[2];                                        // inner, bound
};

// mesh_run traversal (mesh_2d_utils.hpp:13-37): omp dynamic over nodes, elements of the node in
// ascending order, then LOCAL (jL) or NONLOCAL (eNL in neighbours, jNL).
template<class Runner>
void mesh_run(const Mesh& m, const std::vector<Group>& groups, Runner proto) {
#pragma omp parallel firstprivate(proto)
    {
        Runner runner = proto;
        runner.init();
#pragma omp for schedule(dynamic)
        for (i64 node = 0; node < m.N; ++node) {
            runner.reset(node);
            for (i64 k = m.ne_off[node]; k < m.ne_off[node + 1]; ++k) {
                const i64 eL = m.ne_ids[k];
                const int iL = m.local(eL, node);
                const int g = m.group_of[eL];
                if (!groups[g].nonlocal) {
                    for (int jL = 0; jL < m.nn; ++jL) runner(g, eL, eL, iL, jL, true);
                } else {
                    for (i64 s = m.nb_off[eL]; s < m.nb_off[eL + 1]; ++s) {
                        const i64 eNL = m.nb_ids[s];
                        for (int jNL = 0; jNL < m.nn; ++jNL) runner(g, eL, eNL, iL, jNL, false);
                    }
                }
            }
        }
    }
}

// matrix_indexator_base flags (matrix_indexator_base.hpp:34-61).  `faithful_reset` clears the
// whole tail of the flag vectors per node like the reference (O(rows) per node); otherwise only the
// flags that were set are cleared -- same result, used by the tests for speed.
struct Flags {
    i64 rows = 0;
    int dof = 1;
    bool symmetric = true, faithful = false;
    std::vector<std::vector<uint8_t>> f;             // [part*dof + d][col]
    std::vector<std::vector<i64>> touched;
    void init_flags() {
        f.assign(2 * dof, std::vector<uint8_t>(rows, 0));
        touched.assign(2 * dof, {});
    }
    void reset(i64 node) {
        for (int d = 0; d < dof; ++d) {
            if (faithful) {
                std::fill(f[BOUND * dof + d].begin(), f[BOUND * dof + d].end(), 0);
                std::fill(f[INNER * dof + d].begin() + (symmetric ? dof * node : 0), f[INNER * dof + d].end(), 0);
            } else {
                for (int p = 0; p < 2; ++p) {
                    for (i64 c : touched[p * dof + d]) f[p * dof + d][c] = 0;
                    touched[p * dof + d].clear();
                }
            }
        }
    }
    bool check(Part part, int d, i64 col) {
        if (part == NO) return false;
        auto& v = f[part * dof + d];
        if (v[col]) return false;
        v[col] = 1;
        if (!faithful) touched[part * dof + d].push_back(col);
        return true;
    }
};

struct ShiftRunner {                                 // shift_initializer.hpp:36-57
    const Mesh* m; Assembly* a; Separator sep; Flags fl;
    void init() { fl.init_flags(); }
    void reset(i64 node) { fl.reset(node); }
    void operator()(int, i64 eL, i64 eNL, int iL, int jNL, bool) {
        const int D = a->dof;
        const i64 rg = i64(D) * m->el[eL * m->nn + iL], cg = i64(D) * m->el[eNL * m->nn + jNL];
        for (int rl = 0; rl < D; ++rl)
            for (int cl = 0; cl < D; ++cl) {
                const i64 row = rg + rl, col = cg + cl;
                const Part p = sep.part(row, col);
                if (fl.check(p, rl, col)) {
#pragma omp atomic
                    a->A[p].outer[row + 1]++;
                }
            }
    }
};

struct IndexRunner {                                 // matrix_index_initializer.hpp:95-124
    const Mesh* m; Assembly* a; Separator sep; Flags fl;
    i64 cur[2][2] = {{0, 0}, {0, 0}};
    void init() { fl.init_flags(); }
    void reset(i64 node) {
        for (int p = 0; p < 2; ++p)
            for (int d = 0; d < a->dof; ++d) cur[p][d] = a->A[p].outer[a->dof * node + d];
        fl.reset(node);
    }
    void operator()(int, i64 eL, i64 eNL, int iL, int jNL, bool) {
        const int D = a->dof;
        const i64 rg = i64(D) * m->el[eL * m->nn + iL], cg = i64(D) * m->el[eNL * m->nn + jNL];
        for (int rl = 0; rl < D; ++rl)
            for (int cl = 0; cl < D; ++cl) {
                const i64 row = rg + rl, col = cg + cl;
                const Part p = sep.part(row, col);
                if (fl.check(p, rl, col)) a->A[p].inner[cur[p][rl]++] = col;
            }
    }
};

// Local and nonlocal integrands.
// conductivity_matrix_2d.hpp:106-164, stiffness_matrix_2d.hpp:44-123.
// conductivity_matrix_2d::integrate_local (conductivity_matrix_2d.hpp:106-130), the conductivity at q
void thermal_local(const Mesh& m, const Group& g, i64 e, int i, int j, double* out) {
    double integral = 0.0;
    for (int q = 0; q < m.t.nq; ++q) {
        const double ix = m.Dx(e, i, q), iy = m.Dy(e, i, q), jx = m.Dx(e, j, q), jy = m.Dy(e, j, q);
        const double* c = g.coeff(e, q, m.t.nq, g.th.c);
        double value;
        if (g.th.kind == NLF_CONDUCTIVITY_ISOTROPIC) value = c[0] * (ix * jx + iy * jy);
        else if (g.th.kind == NLF_CONDUCTIVITY_ORTHOTROPIC) value = c[0] * ix * jx + c[1] * iy * jy;
        else value = ix * (c[0] * jx + c[2] * jy) + iy * (c[2] * jx + c[1] * jy);
        integral += m.t.w[q] * value / m.jac[e * m.t.nq + q];
    }
    out[0] = g.p1 * integral;
}

void thermal_nonlocal(const Mesh& m, const Group& g, i64 eL, i64 eNL, int iL, int jNL, double* out) {
    double integral = 0.0;
    const int nq = m.t.nq;
    for (int qL = 0; qL < nq; ++qL) {
        double in0 = 0.0, in1 = 0.0;
        const double* xl = m.xqp(eL, qL);
        const double ix = m.Dx(eL, iL, qL), iy = m.Dy(eL, iL, qL);
        for (int qN = 0; qN < nq; ++qN) {
            const double jx = m.Dx(eNL, jNL, qN), jy = m.Dy(eNL, jNL, qN);
            const double iw = m.t.w[qN] * influence_eval(g.inf, xl, m.xqp(eNL, qN));
            if (g.th.kind == NLF_CONDUCTIVITY_ISOTROPIC) {
                const double s = iw * g.th.c[0];
                in0 += s * jx; in1 += s * jy;
            } else if (g.th.kind == NLF_CONDUCTIVITY_ORTHOTROPIC) {
                in0 += iw * g.th.c[0] * jx; in1 += iw * g.th.c[1] * jy;
            } else {
                in0 += iw * (g.th.c[0] * jx + g.th.c[2] * jy);
                in1 += iw * (g.th.c[2] * jx + g.th.c[1] * jy);
            }
        }
        integral += m.t.w[qL] * (ix * in0 + iy * in1);
    }
    out[0] = (1.0 - g.p1) * integral;
}

// heat_capacity_matrix_2d::integrate_basic_pair (heat_capacity_matrix_2d.hpp:31-37) times density *
// capacity (:55-56)
void capacity_local(const Mesh& m, const Group& g, i64 e, int i, int j, double* out) {
    double integral = 0.0;
    for (int q = 0; q < m.t.nq; ++q)
        integral += m.t.w[q] * m.t.qN[i * m.t.nq + q] * m.t.qN[j * m.t.nq + q] * m.jac[e * m.t.nq + q];
    out[0] = g.capacity * integral;
}

// stiffness_matrix::integrate_local (stiffness_matrix_2d.hpp:44-80), the branch of the group's Hooke kind
void elastic_local(const Mesh& m, const Group& g, i64 e, int i, int j, double* out) {
    double b[4] = {0, 0, 0, 0};
    for (int q = 0; q < m.t.nq; ++q) {
        const double* h = g.coeff(e, q, m.t.nq, g.hooke);
        const double f = m.t.w[q] / m.jac[e * m.t.nq + q];
        const double wx = f * m.Dx(e, i, q), wy = f * m.Dy(e, i, q);
        const double jx = m.Dx(e, j, q), jy = m.Dy(e, j, q);
        if (g.hkind != NLF_ELASTIC_ANISOTROPIC) {               // indices 11 12 66 (22)
            const double s1 = h[2] * wy, s2 = h[2] * wx;
            b[0] += h[0] * wx * jx + s1 * jy;
            b[1] += h[1] * wx * jy + s1 * jx;
            b[2] += h[1] * wy * jx + s2 * jy;
            b[3] += (g.hkind == NLF_ELASTIC_ISOTROPIC ? h[0] * wy * jy : h[3] * wy * jy) + s2 * jx;
        } else {                                                 // 11 12 66 22 16 26
            const double h16 = h[4] * wx, h26 = h[5] * wy;
            const double s1 = h[2] * wy + h16, s2 = h[2] * wx + h26;
            b[0] += (h[0] * wx + h[4] * wy) * jx + s1 * jy;
            b[1] += (h[1] * wx + h26) * jy + s1 * jx;
            b[2] += (h[1] * wy + h16) * jx + s2 * jy;
            b[3] += (h[3] * wy + h[5] * wx) * jy + s2 * jx;
        }
    }
    for (int k = 0; k < 4; ++k) out[k] = g.p1 * b[k];
}

// stiffness_matrix::integrate_nonlocal (stiffness_matrix_2d.hpp:82-123); Hooke at the nonlocal point
void elastic_nonlocal(const Mesh& m, const Group& g, i64 eL, i64 eNL, int iL, int jNL, double* out) {
    double b[4] = {0, 0, 0, 0};
    const double* h = g.hooke;
    const int nq = m.t.nq;
    for (int qL = 0; qL < nq; ++qL) {
        const double* xl = m.xqp(eL, qL);
        const double ix = m.t.w[qL] * m.Dx(eL, iL, qL), iy = m.t.w[qL] * m.Dy(eL, iL, qL);
        for (int qN = 0; qN < nq; ++qN) {
            const double wgt = m.t.w[qN] * influence_eval(g.inf, xl, m.xqp(eNL, qN));
            const double jx = wgt * m.Dx(eNL, jNL, qN), jy = wgt * m.Dy(eNL, jNL, qN);
            if (g.hkind != NLF_ELASTIC_ANISOTROPIC) {
                const double h22 = g.hkind == NLF_ELASTIC_ISOTROPIC ? h[0] : h[3];
                b[0] += h[0] * ix * jx + h[2] * iy * jy;
                b[1] += h[1] * ix * jy + h[2] * iy * jx;
                b[2] += h[1] * iy * jx + h[2] * ix * jy;
                b[3] += h22 * iy * jy + h[2] * ix * jx;
            } else {
                b[0] += (h[0] * ix + h[4] * iy) * jx + (h[4] * ix + h[2] * iy) * jy;
                b[1] += (h[4] * ix + h[2] * iy) * jx + (h[1] * ix + h[5] * iy) * jy;
                b[2] += (h[4] * ix + h[1] * iy) * jx + (h[2] * ix + h[5] * iy) * jy;
                b[3] += (h[2] * ix + h[5] * iy) * jx + (h[5] * ix + h[3] * iy) * jy;
            }
        }
    }
    for (int k = 0; k < 4; ++k) out[k] = (1.0 - g.p1) * b[k];
}

// Hooke matrix of a material (elastic_parameters.hpp:62-133), in the index order of its kind
void hooke_matrix(const nlf_elastic_material& e, int& kind, double h[6]) {
    kind = e.kind;
    if (e.kind == NLF_ELASTIC_ISOTROPIC) {                       // :70-75
        const double E = e.young, nu = e.poisson;
        const double value = E / (1.0 - nu * nu);
        h[0] = value; h[1] = nu * value; h[2] = 0.5 * value * (1.0 - nu);
        return;
    }
    if (e.kind != NLF_ELASTIC_ORTHOTROPIC && e.kind != NLF_ELASTIC_ANISOTROPIC)
        throw std::domain_error("unknown elastic material kind");
    // orthotropic :89-95: E = {young, young_y}, nu = {poisson, poisson_y}
    const double value = 1.0 / (1.0 - e.poisson * e.poisson_y);
    const double m11 = e.young * value, m12 = e.young * e.poisson_y * value, m66 = e.shear, m22 = e.young_y * value;
    if (e.kind == NLF_ELASTIC_ORTHOTROPIC) {
        h[0] = m11; h[1] = m12; h[2] = m66; h[3] = m22;
        return;
    }
    // anisotropic_elastic_parameters::rotate :102-123
    const double angle_2 = e.angle + e.angle;
    const double sin_2 = std::sin(angle_2);
    const double cos_2 = std::cos(angle_2);
    const double cos_4 = 0.125 * (cos_2 * cos_2 - sin_2 * sin_2);
    const double stiff_diff = m11 - m22;
    const double stiff_core = m11 + m22 - 2 * m12 - 4 * m66;
    const double main_part_plus = stiff_core * (cos_4 + 0.125);
    const double main_part_minus = stiff_core * (0.125 - cos_4);
    const double cos_diff = stiff_diff * (cos_2 + 0.5);
    const double minor_part = stiff_core * cos_2;
    const double sin_2_025 = 0.25 * sin_2;
    h[0] = main_part_plus + 0.5 * (m12 + m22 + cos_diff) + m66;
    h[1] = main_part_minus + m12;
    h[2] = main_part_minus + m66;
    h[3] = main_part_plus + 0.5 * (m12 + m11 - cos_diff) + m66;
    h[4] = sin_2_025 * (stiff_diff + minor_part);
    h[5] = sin_2_025 * (stiff_diff - minor_part);
}

struct ValueRunner {                                 // integrator.hpp:36-91
    const Mesh* m; Assembly* a; Separator sep; const std::vector<Group>* groups; int physics;
    void init() {}
    void reset(i64) {}
    void operator()(int gi, i64 eL, i64 eNL, int iL, int jNL, bool local_theory) {
        const int D = a->dof;
        const i64 rg = i64(D) * m->el[eL * m->nn + iL], cg = i64(D) * m->el[eNL * m->nn + jNL];
        const Group& g = (*groups)[gi];
        double block[4] = {0, 0, 0, 0};
        bool have = false;
        for (int rl = 0; rl < D; ++rl) {
            double* val = nullptr;
            for (int cl = 0; cl < D; ++cl) {
                const i64 row = rg + rl, col = cg + cl;
                const Part p = sep.part(row, col);
                if (p == NO) continue;
                if (!have) {
                    have = true;
                    if (physics == 2) {
                        capacity_local(*m, g, eL, iL, jNL, block);
                    } else if (local_theory) {
                        if (physics == 0) thermal_local(*m, g, eL, iL, jNL, block); else elastic_local(*m, g, eL, iL, jNL, block);
                    } else {
                        if (physics == 0) thermal_nonlocal(*m, g, eL, eNL, iL, jNL, block); else elastic_nonlocal(*m, g, eL, eNL, iL, jNL, block);
                        if (eL == eNL) {
                            double lb[4];
                            if (physics == 0) thermal_local(*m, g, eL, iL, jNL, lb); else elastic_local(*m, g, eL, iL, jNL, lb);
                            for (int k = 0; k < D * D; ++k) block[k] += lb[k];
                        }
                    }
                }
                if (p == BOUND) a->A[BOUND].coeffRef(row, col) += block[rl * D + cl];
                else {
                    if (!val) val = &a->A[INNER].coeffRef(row, col);
                    *val += block[rl * D + cl];
                    ++val;
                }
            }
        }
    }
};

Group make_group(const nlf_model& md, int physics, const nlf_thermal_material* th, const nlf_elastic_material* el, bool only_local) {
    Group g;
    double p1 = md.local_weight;
    double r0 = md.radius[0], r1 = md.radius[1];
    // fix_parameters (read_model.hpp:186-192)
    if (p1 >= 0.999) { r0 = 0; r1 = 0; }
    if (r0 < std::numeric_limits<double>::epsilon() || r1 < std::numeric_limits<double>::epsilon()) p1 = 1.0;
    g.p1 = p1;
    g.nonlocal = !only_local && p1 < 0.999;
    if (g.nonlocal) {
        nlf_model m2 = md; m2.radius[0] = r0; m2.radius[1] = r1;
        g.inf = make_influence(m2);
    }
    if (physics == 2) return g;
    if (physics == 0) g.th = *th;
    else {
        hooke_matrix(*el, g.hkind, g.hooke);
        g.ekind = el->expansion_kind;
        for (int k = 0; k < 3; ++k) g.alpha[k] = el->expansion[k];
    }
    return g;
}

Assembly* assemble(const Mesh& m, int physics, const std::vector<Group>& groups, const std::vector<uint8_t>& is_inner,
                   bool symmetric, bool faithful_reset) {
    auto* a = new Assembly();
    a->dof = physics == 1 ? 2 : 1;
    a->symmetric = symmetric;
    const i64 rows = i64(a->dof) * m.N;
    for (int p = 0; p < 2; ++p) { a->A[p].rows = rows; a->A[p].cols = rows; a->A[p].outer.assign(rows + 1, 0); }
    Separator sep{&is_inner, symmetric};
    Flags fl; fl.rows = rows; fl.dof = a->dof; fl.symmetric = symmetric; fl.faithful = faithful_reset;
    // create_matrix_portrait (conductivity_matrix_2d.hpp:39-72) -> init_shifts (matrix_assembler_2d.hpp:144-156)
    mesh_run(m, groups, ShiftRunner{&m, a, sep, fl});
    for (i64 r = 0; r < rows; ++r) if (!is_inner[r]) a->A[INNER].outer[r + 1]++;   // first_kind_filler
    for (int p = 0; p < 2; ++p) for (i64 r = 0; r < rows; ++r) a->A[p].outer[r + 1] += a->A[p].outer[r];
    for (int p = 0; p < 2; ++p) {                    // allocate_matrix (solvers/base/utils.hpp:32-38)
        a->A[p].inner.assign(a->A[p].outer[rows], 0);
        a->A[p].val.assign(a->A[p].outer[rows], 0.0);
    }
    mesh_run(m, groups, IndexRunner{&m, a, sep, fl});                // init_indices :158-175
    for (i64 r = 0; r < rows; ++r) if (!is_inner[r]) a->A[INNER].inner[a->A[INNER].outer[r]] = r;
    for (int p = 0; p < 2; ++p) {                    // sort_indices (solvers/base/utils.hpp:40-45)
        CSR& A = a->A[p];
#pragma omp parallel for schedule(dynamic)
        for (i64 r = 0; r < rows; ++r) std::sort(A.inner.begin() + A.outer[r], A.inner.begin() + A.outer[r + 1]);
    }
    mesh_run(m, groups, ValueRunner{&m, a, sep, &groups, physics});  // calc_coeffs :177-189
    for (i64 r = 0; r < rows; ++r) if (!is_inner[r]) a->A[INNER].val[a->A[INNER].outer[r]] = 1.0;
    return a;
}

// ---------------------------------------------------------------------------------------------
// Boundary conditions and right-hand side
// ---------------------------------------------------------------------------------------------
// 1D element integration data for boundary element be, quad point q: (|J|, N_i)
double boundary_jac(const Mesh& m, i64 be, int q) {  // element_data_1d::jacobi_matrix + mesh::jacobian
    double jx = 0, jy = 0;
    for (int i = 0; i < m.nb; ++i) {
        const i64 node = m.be[be * m.nb + i];
        jx = jx + m.xy[2 * node] * m.t1.qNxi[i * m.t1.nq + q];
        jy = jy + m.xy[2 * node + 1] * m.t1.qNxi[i * m.t1.nq + q];
    }
    return std::sqrt(jx * jx + jy * jy);
}

// boundary_condition_second_kind_2d (boundary_condition_second_kind_2d.hpp:12-33) with constant values;
// thermal convection also adds h*T_a (thermal_boundary_conditions_2d.hpp:54-56).
// A position-dependent value: condition(element.quad_coord(q)) for the second kind, and
// condition(node_coord) for the first kind (boundary_condition_first_kind_2d.hpp:17-26).
using ValueFn = double (*)(double, double, int, int, void*);

void boundary_point(const Mesh& m, i64 be, int q, double* p) {   // element_1d_data::quad_coord
    double x = 0.0, y = 0.0;
    for (int i = 0; i < m.nb; ++i) {
        const i64 node = m.be[be * m.nb + i];
        x += m.xy[2 * node] * m.t1.qN[i * m.t1.nq + q];
        y += m.xy[2 * node + 1] * m.t1.qN[i * m.t1.nq + q];
    }
    p[0] = x; p[1] = y;
}

void second_kind(const Mesh& m, int dof, const nlf_boundary* bc, double* f, ValueFn fn = nullptr, void* user = nullptr) {
    for (size_t b = 0; b + 1 < m.bbeg.size(); ++b)
        for (int d = 0; d < dof; ++d) {
            double g;
            const bool flux = bc[b].kind[d] == NLF_BC_FLUX;
            if (flux) g = bc[b].value[d];
            else if (dof == 1 && bc[b].kind[d] == NLF_BC_CONVECTION) g = bc[b].heat_transfer * bc[b].ambient;
            else continue;
            for (i64 be = m.bbeg[b]; be < m.bbeg[b + 1]; ++be)
                for (int i = 0; i < m.nb; ++i) {
                    const i64 node = m.be[be * m.nb + i];
                    double integral = 0.0;
                    for (int q = 0; q < m.t1.nq; ++q) {
                        double gq = g;
                        if (flux && fn) {
                            double p[2];
                            boundary_point(m, be, q, p);
                            gq = fn(p[0], p[1], int(b), d, user);
                        }
                        integral += m.t1.w[q] * m.t1.qN[i * m.t1.nq + q] * gq * boundary_jac(m, be, q);
                    }
                    f[dof * node + d] += integral;
                }
        }
}

// convection_condition_2d (convection_condition_2d.hpp:11-31): K(row, col) += h*int N_i N_j for
// col >= row.  Entries whose row or column is Dirichlet are left out (the reference zeroes them in
// Dirichlet rows through first_kind_matrix_fill_2d).
void convection(const Mesh& m, const nlf_boundary* bc, const std::vector<uint8_t>& inner, CSR& K) {
    for (size_t b = 0; b + 1 < m.bbeg.size(); ++b) {
        if (bc[b].kind[0] != NLF_BC_CONVECTION) continue;
        for (i64 be = m.bbeg[b]; be < m.bbeg[b + 1]; ++be)
            for (int i = 0; i < m.nb; ++i) {
                const i64 row = m.be[be * m.nb + i];
                for (int j = 0; j < m.nb; ++j) {
                    const i64 col = m.be[be * m.nb + j];
                    if (col < row || !inner[row] || !inner[col]) continue;
                    double integral = 0.0;
                    for (int q = 0; q < m.t1.nq; ++q)
                        integral += m.t1.w[q] * m.t1.qN[i * m.t1.nq + q] * m.t1.qN[j * m.t1.nq + q] * boundary_jac(m, be, q);
                    K.coeffRef(row, col) += bc[b].heat_transfer * integral;
                }
            }
    }
}

// Dirichlet values x_D (calc_vector :17-26) and inner mask (solvers_utils.hpp:31-40)
void first_kind_vector(const Mesh& m, int dof, const nlf_boundary* bc, double* x, uint8_t* inner, ValueFn fn = nullptr,
                       void* user = nullptr) {
    for (size_t b = 0; b + 1 < m.bbeg.size(); ++b)
        for (int d = 0; d < dof; ++d) {
            if (bc[b].kind[d] != NLF_BC_FIRST_KIND) continue;
            for (i64 be = m.bbeg[b]; be < m.bbeg[b + 1]; ++be)
                for (int i = 0; i < m.nb; ++i) {
                    const i64 node = m.be[be * m.nb + i];
                    if (x && x[dof * node + d] == 0.0)
                        x[dof * node + d] = fn ? fn(m.xy[2 * node], m.xy[2 * node + 1], int(b), d, user) : bc[b].value[d];
                    if (inner) inner[dof * node + d] = 0;
                }
        }
}

void upper_spmv(const CSR& A, const double* x, double* y);

// boundary_condition_first_kind_2d (:50-59): f -= K_bound * x_D; f[D] = x_D
void first_kind_rhs(const CSR& bound, const std::vector<uint8_t>& inner, const double* xD, double* f) {
    for (i64 r = 0; r < bound.rows; ++r) {
        double s = 0.0;
        for (i64 k = bound.outer[r]; k < bound.outer[r + 1]; ++k) s += bound.val[k] * xD[bound.inner[k]];
        f[r] -= s;
    }
    for (i64 r = 0; r < bound.rows; ++r) if (!inner[r]) f[r] = xD[r];
}

// ---------------------------------------------------------------------------------------------
// SpMV on upper symmetric CSR (independent_symmetric_matrix_vector_product.hpp:41-91)
// ---------------------------------------------------------------------------------------------
void upper_spmv(const CSR& A, const double* x, double* y) {
    const int T = omp_get_max_threads();
    std::vector<double> acc(size_t(A.rows) * T, 0.0);
    // uniform_ranges by nnz (uniform_ranges.hpp:14-40)
    std::vector<i64> cut(T + 1, A.rows);
    cut[0] = 0;
    const i64 nnz = A.outer[A.rows];
    for (int t = 1; t < T; ++t) {
        const i64 target = nnz * t / T;
        cut[t] = std::lower_bound(A.outer.begin(), A.outer.end(), target) - A.outer.begin();
        cut[t] = std::min(std::max(cut[t], cut[t - 1]), A.rows);
    }
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        double* a = &acc[size_t(t) * A.rows];
        for (i64 row = cut[t]; row < cut[t + 1]; ++row) {
            const i64 ind = A.outer[row];
            if (ind == A.outer[row + 1]) continue;
            a[row] += A.val[ind] * x[A.inner[ind]];
            for (i64 i = ind + 1; i < A.outer[row + 1]; ++i) {
                a[row] += A.val[i] * x[A.inner[i]];
                a[A.inner[i]] += A.val[i] * x[row];
            }
        }
    }
    for (i64 r = 0; r < A.rows; ++r) {
        double s = 0.0;
        for (int t = 0; t < T; ++t) s += acc[size_t(t) * A.rows + r];
        y[r] = s;
    }
}

// ---------------------------------------------------------------------------------------------
// Preconditioners: identity, Jacobi (Eigen::DiagonalPreconditioner), IC(0) with diagonal scaling
// and shift retries (our incomplete Cholesky; Eigen::IncompleteCholesky is not vendored).
// ---------------------------------------------------------------------------------------------
struct Precond {
    int kind = NLF_PRECOND_IDENTITY;
    std::vector<double> invdiag;
    CSR U;                   // upper factor of the scaled matrix, diagonal first
    std::vector<double> scale;
    double shift = 0.0;
    bool ok = true;
    void apply(const double* r, double* z, i64 n) const {
        if (kind == NLF_PRECOND_IDENTITY) { std::copy(r, r + n, z); return; }
        if (kind == NLF_PRECOND_JACOBI) { for (i64 i = 0; i < n; ++i) z[i] = invdiag[i] * r[i]; return; }
        // z = S U^{-1} U^{-T} S r
        std::vector<double> y(n);
        for (i64 i = 0; i < n; ++i) y[i] = scale[i] * r[i];
        for (i64 k = 0; k < n; ++k) {                // forward U^T y = b (column sweep)
            const i64 b = U.outer[k];
            y[k] /= U.val[b];
            for (i64 p = b + 1; p < U.outer[k + 1]; ++p) y[U.inner[p]] -= U.val[p] * y[k];
        }
        for (i64 k = n - 1; k >= 0; --k) {           // backward U z = y
            const i64 b = U.outer[k];
            double s = y[k];
            for (i64 p = b + 1; p < U.outer[k + 1]; ++p) s -= U.val[p] * y[U.inner[p]];
            y[k] = s / U.val[b];
        }
        for (i64 i = 0; i < n; ++i) z[i] = scale[i] * y[i];
    }
};

bool ic0_factor(const CSR& A, const std::vector<double>& s, double shift, CSR& U) {
    U = A;
    const i64 n = A.rows;
    for (i64 r = 0; r < n; ++r)
        for (i64 k = U.outer[r]; k < U.outer[r + 1]; ++k) U.val[k] = s[r] * A.val[k] * s[U.inner[k]];
    for (i64 r = 0; r < n; ++r) U.val[U.outer[r]] += shift;
    for (i64 k = 0; k < n; ++k) {
        const i64 b = U.outer[k], e = U.outer[k + 1];
        const double d = U.val[b];
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        const double ukk = std::sqrt(d);
        U.val[b] = ukk;
        for (i64 p = b + 1; p < e; ++p) U.val[p] /= ukk;
        for (i64 p = b + 1; p < e; ++p) {
            const i64 j = U.inner[p];
            const double ukj = U.val[p];
            // update row j at columns l >= j present in row k
            i64 qj = U.outer[j];
            const i64 ej = U.outer[j + 1];
            for (i64 q = p; q < e; ++q) {
                const i64 l = U.inner[q];
                while (qj < ej && U.inner[qj] < l) ++qj;
                if (qj < ej && U.inner[qj] == l) U.val[qj] -= ukj * U.val[q];
            }
        }
    }
    return true;
}

Precond* make_precond(int kind, const CSR* A) {
    auto* P = new Precond();
    P->kind = kind;
    if (kind == NLF_PRECOND_JACOBI) {
        P->invdiag.resize(A->rows);
        for (i64 r = 0; r < A->rows; ++r) {
            const double d = A->coeff(r, r);
            P->invdiag[r] = d != 0.0 ? 1.0 / d : 1.0;
        }
    } else if (kind == NLF_PRECOND_ILLT) {
        P->scale.resize(A->rows);
        for (i64 r = 0; r < A->rows; ++r) {
            const double d = std::abs(A->coeff(r, r));
            P->scale[r] = d > 0.0 ? 1.0 / std::sqrt(d) : 1.0;
        }
        double shift = 0.0;
        P->ok = false;
        for (int attempt = 0; attempt < 40; ++attempt) {
            if (ic0_factor(*A, P->scale, shift, P->U)) { P->ok = true; P->shift = shift; break; }
            shift = shift == 0.0 ? 1e-3 : 2 * shift;
        }
    }
    return P;
}

// conjugate_gradient::solve (conjugate_gradient.hpp:48-79): stops on sqrt(r'z)/||b|| <= tol.
struct CGResult { i64 iterations; double residual; };

CGResult cg(const CSR& A, const Precond& M, const double* b, double* x, double tol, i64 max_it) {
    const i64 n = A.rows;
    std::vector<double> r(n), z(n), p(n);
    upper_spmv(A, x, r.data());
    for (i64 i = 0; i < n; ++i) r[i] = b[i] - r[i];
    M.apply(r.data(), p.data(), n);
    double rz = 0.0, bn = 0.0;
    for (i64 i = 0; i < n; ++i) { rz += r[i] * p[i]; bn += b[i] * b[i]; }
    bn = std::sqrt(bn);
    i64 it = 0;
    double res = std::sqrt(rz) / bn;
    while (it < max_it && res > tol) {
        upper_spmv(A, p.data(), z.data());
        double pz = 0.0;
        for (i64 i = 0; i < n; ++i) pz += p[i] * z[i];
        const double nu = rz / pz;
        for (i64 i = 0; i < n; ++i) { x[i] += nu * p[i]; r[i] -= nu * z[i]; }
        M.apply(r.data(), z.data(), n);
        const double prev = rz;
        rz = 0.0;
        for (i64 i = 0; i < n; ++i) rz += r[i] * z[i];
        const double mu = rz / prev;
        for (i64 i = 0; i < n; ++i) p[i] = z[i] + mu * p[i];
        ++it;
        res = std::sqrt(rz) / bn;
    }
    return {it, res};
}

thread_local std::string g_error;


// ---------------------------------------------------------------------------------------------
// Heat flux post-processing: heat_equation_solution_2d::calc_flux (heat_equation_solution_2d.hpp:99-137)
// ---------------------------------------------------------------------------------------------
// mesh_2d::area (mesh_2d.hpp:226-235): sum over q of w_q |J_q| sum_i N_i(q), accumulated q-outer,
// i-inner.
double element_area(const Mesh& m, i64 e) {
    double area = 0.0;
    for (int q = 0; q < m.t.nq; ++q) {
        const double factor = m.t.w[q] * m.jac[e * m.t.nq + q];
        for (int i = 0; i < m.nn; ++i) area += factor * m.t.qN[i * m.t.nq + q];
    }
    return area;
}

// qnodes_to_nodes (mesh_2d_utils.hpp:82-101): area-weighted mean, over the elements of each node in
// ascending id order, of the value at the element's quadrature point nearest to the node.
void qnodes_to_nodes(const Mesh& m, const double* x, double* out) {
#pragma omp parallel for schedule(static)
    for (i64 node = 0; node < m.N; ++node) {
        double value = 0.0, node_area = 0.0;
        for (i64 k = m.ne_off[node]; k < m.ne_off[node + 1]; ++k) {
            const i64 e = m.ne_ids[k];
            const double area = element_area(m, e);
            const int i = m.local(e, node);
            value += area * x[e * m.t.nq + m.t.nearest[i]];
            node_area += area;
        }
        out[node] = value / node_area;
    }
}

// gradient_in_qnodes (mesh_2d_utils.hpp:39-63): sum_i x_i D_i(q), then divided by |J_q|; x is read
// with a stride (1 for the temperature, 2 for one component of the displacement).
void gradient_in_qnodes(const Mesh& m, const double* x, int stride, std::vector<double>& gx, std::vector<double>& gy) {
    const int nq = m.t.nq, nn = m.nn;
    gx.assign(m.E * nq, 0.0);
    gy.assign(m.E * nq, 0.0);
#pragma omp parallel for schedule(static)
    for (i64 e = 0; e < m.E; ++e)
        for (int q = 0; q < nq; ++q) {
            const i64 qs = e * nq + q;
            for (int i = 0; i < nn; ++i) {
                const double val = x[i64(stride) * m.el[e * nn + i]];
                gx[qs] += val * m.Dx(e, i, q);
                gy[qs] += val * m.Dy(e, i, q);
            }
            gx[qs] /= m.jac[qs];
            gy[qs] /= m.jac[qs];
        }
}

// qflux: [2][E*nq] quadrature-point fluxes (x block, then y block); nflux: [2][N] node fluxes.
void heat_flux(const Mesh& m, const std::vector<Group>& groups, const double* T, double* qflux, double* nflux) {
    const int nq = m.t.nq;
    const i64 Q = m.E * nq;
    std::vector<double> fx, fy;
    gradient_in_qnodes(m, T, 1, fx, fy);
    // local_flux_in_qnodes (heat_equation_solution_2d.hpp:72-97): -lambda grad T per group
    for (int g = 0; g < m.G(); ++g) {
        const nlf_thermal_material& c = groups[g].th;
        for (i64 e = m.gbeg[g]; e < m.gbeg[g + 1]; ++e)
            for (int q = 0; q < nq; ++q) {
                const i64 qs = e * nq + q;
                if (c.kind == NLF_CONDUCTIVITY_ISOTROPIC) { fx[qs] *= -c.c[0]; fy[qs] *= -c.c[0]; }
                else if (c.kind == NLF_CONDUCTIVITY_ORTHOTROPIC) { fx[qs] *= -c.c[0]; fy[qs] *= -c.c[1]; }
                else {
                    const double x = fx[qs], y = fy[qs];
                    fx[qs] = -c.c[0] * x - c.c[2] * y;
                    fy[qs] = -c.c[2] * x - c.c[1] * y;
                }
            }
    }
    // nonlocal groups: p1 q(x) + (1 - p1) sum_{eNL in N(eL)} sum_qNL w |J| phi(x, x') q(x')
    std::vector<double> gx = fx, gy = fy;
    for (int g = 0; g < m.G(); ++g) {
        const Group& gr = groups[g];
        if (!gr.nonlocal) continue;
        if (m.nb_off.empty()) throw std::logic_error("heat flux of a nonlocal group needs the neighbour lists");
#pragma omp parallel for schedule(dynamic, 64)
        for (i64 eL = m.gbeg[g]; eL < m.gbeg[g + 1]; ++eL)
            for (int qL = 0; qL < nq; ++qL) {
                const i64 qsL = eL * nq + qL;
                double ng[2] = {0.0, 0.0};
                const double* xl = m.xqp(eL, qL);
                for (i64 s = m.nb_off[eL]; s < m.nb_off[eL + 1]; ++s) {
                    const i64 eNL = m.nb_ids[s];
                    for (int qN = 0; qN < nq; ++qN) {
                        const i64 qsN = eNL * nq + qN;
                        const double iw = m.t.w[qN] * m.jac[qsN] * influence_eval(gr.inf, xl, m.xqp(eNL, qN));
                        ng[0] += iw * fx[qsN];
                        ng[1] += iw * fy[qsN];
                    }
                }
                ng[0] *= 1.0 - gr.p1;
                ng[1] *= 1.0 - gr.p1;
                gx[qsL] *= gr.p1;
                gy[qsL] *= gr.p1;
                gx[qsL] += ng[0];
                gy[qsL] += ng[1];
            }
    }
    if (qflux) {
        std::copy(gx.begin(), gx.end(), qflux);
        std::copy(gy.begin(), gy.end(), qflux + Q);
    }
    if (nflux) {
        qnodes_to_nodes(m, gx.data(), nflux);
        qnodes_to_nodes(m, gy.data(), nflux + m.N);
    }
}

// calc_stress (mechanical_parameters_2d.hpp:27-48) for the group's Hooke kind; tensor shear strain
void calc_stress(const Group& g, double e11, double e22, double e12, double* out) {
    const double* h = g.hooke;
    if (g.hkind == NLF_ELASTIC_ISOTROPIC) {
        out[0] = h[0] * e11 + h[1] * e22;
        out[1] = h[1] * e11 + h[0] * e22;
        out[2] = 2 * h[2] * e12;
    } else if (g.hkind == NLF_ELASTIC_ORTHOTROPIC) {
        out[0] = h[0] * e11 + h[1] * e22;
        out[1] = h[1] * e11 + h[3] * e22;
        out[2] = 2 * h[2] * e12;
    } else {
        out[0] = h[0] * e11 + h[1] * e22 + 2 * h[4] * e12;
        out[1] = h[1] * e11 + h[3] * e22 + 2 * h[5] * e12;
        out[2] = h[4] * e11 + h[5] * e22 + 2 * h[2] * e12;
    }
}

// nodes_to_qnodes (mesh_2d_utils.hpp:65-80): sum_i x_i N_i(q)
std::vector<double> nodes_to_qnodes(const Mesh& m, const double* x) {
    const int nq = m.t.nq;
    std::vector<double> v(size_t(m.E * nq), 0.0);
#pragma omp parallel for schedule(static)
    for (i64 e = 0; e < m.E; ++e)
        for (int q = 0; q < nq; ++q)
            for (int i = 0; i < m.nn; ++i) v[e * nq + q] += x[m.el[e * m.nn + i]] * m.t.qN[i * nq + q];
    return v;
}

// the thermal strain of a group at a point of temperature change dt, {xx, yy, xy} (substract_temperature_strains,
// mechanical_solution_2d.hpp:135-166)
void thermal_strain(const Group& g, double dt, double* out) {
    if (g.ekind == NLF_EXPANSION_ISOTROPIC) { out[0] = g.alpha[0] * dt; out[1] = out[0]; out[2] = 0.0; }
    else if (g.ekind == NLF_EXPANSION_ORTHOTROPIC) { out[0] = g.alpha[0] * dt; out[1] = g.alpha[1] * dt; out[2] = 0.0; }
    else { out[0] = g.alpha[0] * dt; out[1] = g.alpha[1] * dt; out[2] = g.alpha[2] * dt; }
}

// _temperature_condition::calc_stress (temperature_condition_2d.hpp:20-74): the stress of the thermal
// strain for each pair of Hooke and expansion kinds, written as the reference writes each branch
void thermal_stress(const Group& g, double dt, double* out) {
    const double* h = g.hooke;
    const double* a = g.alpha;
    if (g.ekind == NLF_EXPANSION_ISOTROPIC) {
        const double ts = dt * a[0];
        if (g.hkind == NLF_ELASTIC_ISOTROPIC) { const double v = (h[0] + h[1]) * ts; out[0] = v; out[1] = v; out[2] = 0.0; }
        else if (g.hkind == NLF_ELASTIC_ORTHOTROPIC) { out[0] = (h[0] + h[1]) * ts; out[1] = (h[1] + h[3]) * ts; out[2] = 0.0; }
        else { out[0] = (h[0] + h[1]) * ts; out[1] = (h[1] + h[3]) * ts; out[2] = (h[4] + h[5]) * ts; }
        return;
    }
    const double sxx = dt * a[0], syy = dt * a[1], sxy = g.ekind == NLF_EXPANSION_ANISOTROPIC ? dt * a[2] : 0.0;
    const double h22 = g.hkind == NLF_ELASTIC_ISOTROPIC ? h[0] : h[3];
    if (g.hkind != NLF_ELASTIC_ANISOTROPIC) {
        out[0] = h[0] * sxx + h[1] * syy;
        out[1] = h[1] * sxx + h22 * syy;
        out[2] = g.ekind == NLF_EXPANSION_ANISOTROPIC ? 2 * h[2] * sxy : 0.0;
    } else if (g.ekind == NLF_EXPANSION_ORTHOTROPIC) {
        out[0] = h[0] * sxx + h[1] * syy;
        out[1] = h[1] * sxx + h[3] * syy;
        out[2] = h[4] * sxx + h[5] * syy;
    } else {
        out[0] = h[0] * sxx + 2 * h[4] * sxy + h[1] * syy;
        out[1] = h[1] * sxx + 2 * h[5] * sxy + h[3] * syy;
        out[2] = h[4] * sxx + 2 * h[2] * sxy + h[5] * syy;
    }
}

// temperature_condition (temperature_condition_2d.hpp:128-170): f[2 node + d] += the thermal load.  Per
// node, the node's elements in ascending order; per element, for a nonlocal group first the sum over
// its neighbours of (1 - p1) times the nonlocal integral (:106-124), then p1 times the local one
// (:80-92).  The nonlocal integral adds wdN_x (s_xx + s_xy) and wdN_y (s_yy + s_xy) of the averaged
// thermal stress s, as the reference writes it (:120-121).
void temperature_condition(const Mesh& m, const std::vector<Group>& groups, const double* dT, double* f) {
    const int nq = m.t.nq;
    const std::vector<double> tq = nodes_to_qnodes(m, dT);
#pragma omp parallel for schedule(dynamic, 64)
    for (i64 node = 0; node < m.N; ++node) {
        double integral[2] = {0.0, 0.0};
        for (i64 k = m.ne_off[node]; k < m.ne_off[node + 1]; ++k) {
            const i64 eL = m.ne_ids[k];
            const Group& g = groups[m.group_of[eL]];
            if (g.ekind == NLF_EXPANSION_NONE) continue;
            int iL = 0;
            for (int i = 0; i < m.nn; ++i) if (m.el[eL * m.nn + i] == node) iL = i;
            if (g.nonlocal) {
                const double w1 = 1.0 - g.p1;
                for (i64 s = m.nb_off[eL]; s < m.nb_off[eL + 1]; ++s) {
                    const i64 eNL = m.nb_ids[s];
                    double in[2] = {0.0, 0.0};
                    for (int qL = 0; qL < nq; ++qL) {
                        double acc[3] = {0.0, 0.0, 0.0};
                        const double* xl = m.xqp(eL, qL);
                        for (int qN = 0; qN < nq; ++qN) {
                            const i64 qs = eNL * nq + qN;
                            const double wgt = m.t.w[qN] * influence_eval(g.inf, xl, m.xqp(eNL, qN)) * m.jac[qs];
                            double st[3];
                            thermal_stress(g, tq[qs], st);
                            for (int c = 0; c < 3; ++c) acc[c] += wgt * st[c];
                        }
                        const double wx = m.t.w[qL] * m.Dx(eL, iL, qL), wy = m.t.w[qL] * m.Dy(eL, iL, qL);
                        in[0] += wx * (acc[0] + acc[2]);
                        in[1] += wy * (acc[1] + acc[2]);
                    }
                    integral[0] += w1 * in[0];
                    integral[1] += w1 * in[1];
                }
            }
            double lo[2] = {0.0, 0.0};
            for (int q = 0; q < nq; ++q) {
                double st[3];
                thermal_stress(g, tq[eL * nq + q], st);
                const double wx = m.t.w[q] * m.Dx(eL, iL, q), wy = m.t.w[q] * m.Dy(eL, iL, q);
                lo[0] += wx * st[0] + wy * st[2];
                lo[1] += wy * st[1] + wx * st[2];
            }
            integral[0] += g.p1 * lo[0];
            integral[1] += g.p1 * lo[1];
        }
        f[2 * node] += integral[0];
        f[2 * node + 1] += integral[1];
    }
}

// mechanical_solution_2d::calc_strain_and_stress (mechanical_solution_2d.hpp:192-232).  u: displacement
// (x, y per node); dT: the temperature change per node or null (no thermal strain).  qstress [3][E*nq];
// nstrain, nstress [3][N] (xx, yy, xy).
void strain_stress(const Mesh& m, const std::vector<Group>& groups, const double* u, const double* dT, double* qstress,
                   double* nstrain, double* nstress) {
    const int nq = m.t.nq;
    const i64 Q = m.E * nq;
    // strains_in_quadratures (:150-157)
    std::vector<double> g11, g12, g21, g22;
    gradient_in_qnodes(m, u, 2, g11, g12);
    gradient_in_qnodes(m, u + 1, 2, g21, g22);
    for (i64 q = 0; q < Q; ++q) g12[q] = 0.5 * (g12[q] + g21[q]);
    double* eps[3] = {g11.data(), g22.data(), g12.data()};
    if (nstrain)
        for (int c = 0; c < 3; ++c) qnodes_to_nodes(m, eps[c], nstrain + c * m.N);
    if (dT) {                                        // substract_temperature_strains (:135-166)
        const std::vector<double> tq = nodes_to_qnodes(m, dT);
        for (i64 e = 0; e < m.E; ++e) {
            const Group& g = groups[m.group_of[e]];
            if (g.ekind == NLF_EXPANSION_NONE) continue;
            for (int q = 0; q < nq; ++q) {
                double ts[3];
                thermal_strain(g, tq[e * nq + q], ts);
                for (int c = 0; c < 3; ++c) eps[c][e * nq + q] -= ts[c];
            }
        }
    }
    std::vector<double> st(3 * Q, 0.0);
    for (int g = 0; g < m.G(); ++g) {
        const Group& gr = groups[g];
        if (gr.nonlocal && m.nb_off.empty()) throw std::logic_error("stress of a nonlocal group needs the neighbour lists");
        const double w1 = 1.0 - gr.p1;
#pragma omp parallel for schedule(dynamic, 64)
        for (i64 eL = m.gbeg[g]; eL < m.gbeg[g + 1]; ++eL)
            for (int qL = 0; qL < nq; ++qL) {
                const i64 qsL = eL * nq + qL;
                double local[3];
                calc_stress(gr, eps[0][qsL], eps[1][qsL], eps[2][qsL], local);
                double stress[3] = {gr.p1 * local[0], gr.p1 * local[1], gr.p1 * local[2]};
                if (gr.nonlocal) {
                    // calc_nonlocal_stress (:173-189): influence_weight = w |J| (1 - p1) phi
                    double ns[3] = {0.0, 0.0, 0.0};
                    const double* xl = m.xqp(eL, qL);
                    for (i64 k = m.nb_off[eL]; k < m.nb_off[eL + 1]; ++k) {
                        const i64 eNL = m.nb_ids[k];
                        for (int qN = 0; qN < nq; ++qN) {
                            const i64 qs = eNL * nq + qN;
                            const double iw = m.t.w[qN] * m.jac[qs] * (w1 * influence_eval(gr.inf, xl, m.xqp(eNL, qN)));
                            double sn[3];
                            calc_stress(gr, eps[0][qs], eps[1][qs], eps[2][qs], sn);
                            for (int c = 0; c < 3; ++c) ns[c] += iw * sn[c];
                        }
                    }
                    for (int c = 0; c < 3; ++c) stress[c] += ns[c];
                }
                for (int c = 0; c < 3; ++c) st[c * Q + qsL] += stress[c];
            }
    }
    if (qstress) std::copy(st.begin(), st.end(), qstress);
    if (nstress)
        for (int c = 0; c < 3; ++c) qnodes_to_nodes(m, st.data() + c * Q, nstress + c * m.N);
}

}  // namespace

// =============================================================================================
// C ABI of the oracle (loaded by tests through ctypes)
// =============================================================================================
extern "C" {

const char* orc_last_error() { return g_error.c_str(); }

int orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); return omp_get_max_threads(); }

void* orc_mesh_create(i64 n_nodes, const double* xy, i64 n_elems, int nn, const i32* elems, int n_groups,
                      const i64* group_begin, i64 n_bedges, int nb, const i32* bedges, int n_bgroups,
                      const i64* bgroup_begin) {
    try {
        auto* m = new Mesh();
        m->N = n_nodes; m->E = n_elems; m->nn = nn; m->nb = nb;
        m->xy.assign(xy, xy + 2 * n_nodes);
        m->el.assign(elems, elems + n_elems * nn);
        m->gbeg.assign(group_begin, group_begin + n_groups + 1);
        m->be.assign(bedges, bedges + n_bedges * nb);
        m->bbeg.assign(bgroup_begin, bgroup_begin + n_bgroups + 1);
        mesh_init(*m);
        return m;
    } catch (const std::exception& e) { g_error = e.what(); return nullptr; }
}

void orc_mesh_destroy(void* h) { delete static_cast<Mesh*>(h); }

void orc_centres(void* h, double* out) {
    const Mesh& m = *static_cast<Mesh*>(h);
    std::copy(m.centres.begin(), m.centres.end(), out);
}

void orc_quadrature(void* h, double* xq, double* jac, double* D) {
    const Mesh& m = *static_cast<Mesh*>(h);
    if (xq) std::copy(m.xq.begin(), m.xq.end(), xq);
    if (jac) std::copy(m.jac.begin(), m.jac.end(), jac);
    if (D) std::copy(m.D.begin(), m.D.end(), D);
}

void orc_basis(int nn, double x, double y, double* N, double* Nx, double* Ny) { basis(nn, x, y, N, Nx, Ny); }

void orc_tables(int nn, double* w, double* qN, double* qNxi, double* qNeta, double* N0) {
    const Tables2D t = make_tables(nn);
    std::copy(t.w.begin(), t.w.end(), w);
    std::copy(t.qN.begin(), t.qN.end(), qN);
    std::copy(t.qNxi.begin(), t.qNxi.end(), qNxi);
    std::copy(t.qNeta.begin(), t.qNeta.end(), qNeta);
    std::copy(t.N0.begin(), t.N0.end(), N0);
}

i64 orc_find_neighbours(void* h, const double* radii, int strategy, double* expanded) {
    try { return find_neighbours(*static_cast<Mesh*>(h), radii, nullptr, strategy, expanded); }
    catch (const std::exception& e) { g_error = e.what(); return -1; }
}

i64 orc_find_neighbours_ex(void* h, const double* radii, const nlf_distance* distances, int strategy, double* expanded) {
    try { return find_neighbours(*static_cast<Mesh*>(h), radii, distances, strategy, expanded); }
    catch (const std::exception& e) { g_error = e.what(); return -1; }
}

void orc_neighbours_export(void* h, i64* offsets, i32* ids) {
    const Mesh& m = *static_cast<Mesh*>(h);
    std::copy(m.nb_off.begin(), m.nb_off.end(), offsets);
    std::copy(m.nb_ids.begin(), m.nb_ids.end(), ids);
}

double orc_influence(const nlf_model* md, const double* x, const double* y) {
    return influence_eval(make_influence(*md), x, y);
}

double orc_influence_norm(const nlf_model* md) { return make_influence(*md).norm; }

// phi(x, ys[k]) for n points ys[2n] (quadrature sums over a grid in the tests); 0 on success.
int orc_influence_points(const nlf_model* md, const double* x, i64 n, const double* ys, double* out) {
    try {
        const Influence f = make_influence(*md);
#pragma omp parallel for
        for (i64 k = 0; k < n; ++k) out[k] = influence_eval(f, x, ys + 2 * k);
        return 0;
    } catch (const std::exception& e) { g_error = e.what(); return -1; }
}

// distance(x, ys[k], r) of one group's search (nlf_distance) for n points; 0 on success.
int orc_distance_points(const nlf_distance* d, const double* x, const double* r, i64 n, const double* ys, double* out) {
    try {
        const Distance dist = make_distance(d->kind, d->n, d->n_real);
        for (i64 k = 0; k < n; ++k) out[k] = dist(x, ys + 2 * k, r);
        return 0;
    } catch (const std::exception& e) { g_error = e.what(); return -1; }
}

// physics 0 = thermal (materials: nlf_thermal_material[G]), 1 = elastic (nlf_elastic_material[G]).
void* orc_assemble(void* h, int physics, const nlf_model* models, const void* materials, const uint8_t* is_inner,
                   int symmetric, int part, int faithful_reset) {
    try {
        const Mesh& m = *static_cast<Mesh*>(h);
        std::vector<Group> groups;
        // physics 2: heat capacity matrix, local theories (heat_capacity_matrix_2d.hpp:50-53), materials = rho c
        for (int g = 0; g < m.G(); ++g) {
            groups.push_back(make_group(models[g], physics,
                                        physics == 0 ? &static_cast<const nlf_thermal_material*>(materials)[g] : nullptr,
                                        physics == 1 ? &static_cast<const nlf_elastic_material*>(materials)[g] : nullptr,
                                        part == NLF_PART_LOCAL || physics == 2));
            if (physics == 2) groups.back().capacity = static_cast<const double*>(materials)[g];
        }
        const i64 rows = (physics == 1 ? 2 : 1) * m.N;
        std::vector<uint8_t> inner(is_inner, is_inner + rows);
        return assemble(m, physics, groups, inner, symmetric != 0, faithful_reset != 0);
    } catch (const std::exception& e) { g_error = e.what(); return nullptr; }
}

// The Neumann problem's portrait and integral condition (conductivity_matrix_2d.hpp:39-72, 83-104),
// symmetric case: one more row and column, N; every row r < N ends with (r, N) = sum over the node's
// elements (ascending) of integrate_basic = sum_q w_q N_i(q) |J_q|; row N holds (N, N) = 0.
void neumann_border(const Mesh& m, Assembly& a) {
    CSR& A = a.A[INNER];
    const i64 n = A.rows;
    const int nq = m.t.nq;
    std::vector<double> c(size_t(n), 0.0);
    for (i64 node = 0; node < n; ++node)
        for (i64 k = m.ne_off[node]; k < m.ne_off[node + 1]; ++k) {
            const i64 e = m.ne_ids[k];
            const int i = m.local(e, node);
            double integral = 0.0;
            for (int q = 0; q < nq; ++q) integral += m.t.w[q] * m.t.qN[i * nq + q] * m.jac[e * nq + q];
            c[size_t(node)] += integral;
        }
    CSR B;
    B.rows = B.cols = n + 1;
    B.outer.assign(size_t(n + 2), 0);
    for (i64 r = 0; r < n; ++r) {
        for (i64 p = A.outer[r]; p < A.outer[r + 1]; ++p) { B.inner.push_back(A.inner[p]); B.val.push_back(A.val[p]); }
        B.inner.push_back(n);
        B.val.push_back(c[size_t(r)]);
        B.outer[r + 1] = i64(B.inner.size());
    }
    B.inner.push_back(n);
    B.val.push_back(0.0);
    B.outer[n + 1] = i64(B.inner.size());
    A = std::move(B);
    CSR& D = a.A[BOUND];
    D.rows = D.cols = n + 1;
    D.outer.push_back(D.outer.back());
}

// nlf_assemble_ex's options: neumann (thermal only, no first-kind rows), coefficients per quadrature point
// for each group (coeffs[g] null or E_g * nq * ncomp values in the constant's layout)
void* orc_assemble_ex(void* h, int physics, const nlf_model* models, const void* materials, const uint8_t* is_inner,
                      int symmetric, int part, int faithful_reset, int neumann, const double* const* coeffs) {
    try {
        const Mesh& m = *static_cast<Mesh*>(h);
        std::vector<Group> groups;
        bool any_nonlocal = false, any_varying = false;
        for (int g = 0; g < m.G(); ++g) {
            groups.push_back(make_group(models[g], physics,
                                        physics == 0 ? &static_cast<const nlf_thermal_material*>(materials)[g] : nullptr,
                                        physics == 1 ? &static_cast<const nlf_elastic_material*>(materials)[g] : nullptr,
                                        part == NLF_PART_LOCAL || physics == 2));
            if (physics == 2) groups.back().capacity = static_cast<const double*>(materials)[g];
            Group& gr = groups.back();
            any_nonlocal |= gr.nonlocal;
            if (coeffs && coeffs[g]) {
                if (physics == 2) throw std::domain_error("the heat capacity takes no coefficients per point");
                gr.cq = coeffs[g];
                gr.begin = m.gbeg[g];
                gr.ncomp = physics == 0 ? (gr.th.kind == NLF_CONDUCTIVITY_ISOTROPIC ? 1 :
                                           gr.th.kind == NLF_CONDUCTIVITY_ORTHOTROPIC ? 2 : 3)
                                        : (gr.hkind == NLF_ELASTIC_ISOTROPIC ? 3 : gr.hkind == NLF_ELASTIC_ORTHOTROPIC ? 4 : 6);
                any_varying = true;
            }
        }
        // problem_settings::is_symmetric (problem_settings.hpp:32-34): a nonlocal problem with coefficients
        // that vary is asymmetric (BiCGSTAB in the reference)
        if (any_varying && any_nonlocal) throw std::domain_error("coefficients per point make a nonlocal problem asymmetric");
        const i64 rows = (physics == 1 ? 2 : 1) * m.N;
        std::vector<uint8_t> inner(is_inner, is_inner + rows);
        Assembly* a = assemble(m, physics, groups, inner, symmetric != 0, faithful_reset != 0);
        if (neumann) {
            if (physics != 0) { delete a; throw std::domain_error("Neumann problems are thermal"); }
            neumann_border(m, *a);
        }
        return a;
    } catch (const std::exception& e) { g_error = e.what(); return nullptr; }
}

void orc_matrix_destroy(void* a) { delete static_cast<Assembly*>(a); }

void orc_matrix_info(void* a, int part, i64* rows, i64* nnz) {
    const CSR& A = static_cast<Assembly*>(a)->A[part];
    *rows = A.rows;
    *nnz = A.outer[A.rows];
}

void orc_matrix_export(void* a, int part, i64* outer, i64* inner, double* val) {
    const CSR& A = static_cast<Assembly*>(a)->A[part];
    if (outer) std::copy(A.outer.begin(), A.outer.end(), outer);
    if (inner) std::copy(A.inner.begin(), A.inner.end(), inner);
    if (val) std::copy(A.val.begin(), A.val.end(), val);
}

// rows [row0, row1) of one part: outer[row1 - row0 + 1] from 0, inner / val may be null
void orc_matrix_export_rows(void* a, int part, i64 row0, i64 row1, i64* outer, i64* inner, double* val) {
    const CSR& A = static_cast<Assembly*>(a)->A[part];
    const i64 b = A.outer[row0];
    for (i64 r = row0; r <= row1; ++r) outer[r - row0] = A.outer[r] - b;
    const i64 e = A.outer[row1];
    if (inner) std::copy(A.inner.begin() + b, A.inner.begin() + e, inner);
    if (val) std::copy(A.val.begin() + b, A.val.begin() + e, val);
}

void orc_matrix_set_values(void* a, int part, const double* val) {
    CSR& A = static_cast<Assembly*>(a)->A[part];
    std::copy(val, val + A.outer[A.rows], A.val.begin());
}

// Thermal boundary data: f (second kind + convection RHS), x_D, inner mask; and the convection matrix term.
void orc_boundary_fn(void* h, int dof, const nlf_boundary* bc, ValueFn fn, void* user, double* f, double* xD,
                     uint8_t* inner) {
    const Mesh& m = *static_cast<Mesh*>(h);
    const i64 rows = dof * m.N;
    if (inner) std::fill(inner, inner + rows, uint8_t(1));
    if (xD) std::fill(xD, xD + rows, 0.0);
    first_kind_vector(m, dof, bc, xD, inner, fn, user);
    if (f) { std::fill(f, f + rows, 0.0); second_kind(m, dof, bc, f, fn, user); }
}

void orc_boundary(void* h, int dof, const nlf_boundary* bc, double* f, double* xD, uint8_t* inner) {
    orc_boundary_fn(h, dof, bc, nullptr, nullptr, f, xD, inner);
}

void orc_convection(void* h, void* a, const nlf_boundary* bc, const uint8_t* is_inner) {
    const Mesh& m = *static_cast<Mesh*>(h);
    Assembly* A = static_cast<Assembly*>(a);
    std::vector<uint8_t> inner(is_inner, is_inner + A->A[0].rows);
    convection(m, bc, inner, A->A[INNER]);
}

void orc_first_kind_rhs(void* a, const uint8_t* is_inner, const double* xD, double* f) {
    Assembly* A = static_cast<Assembly*>(a);
    std::vector<uint8_t> inner(is_inner, is_inner + A->A[0].rows);
    first_kind_rhs(A->A[BOUND], inner, xD, f);
}

void orc_spmv(void* a, const double* x, double* y) { upper_spmv(static_cast<Assembly*>(a)->A[INNER], x, y); }

void* orc_precond_create(int kind, void* a) {
    return make_precond(kind, a ? &static_cast<Assembly*>(a)->A[INNER] : nullptr);
}
void orc_precond_destroy(void* p) { delete static_cast<Precond*>(p); }
int orc_precond_ok(void* p, double* shift) { auto* P = static_cast<Precond*>(p); if (shift) *shift = P->shift; return P->ok; }
void orc_precond_apply(void* p, const double* r, double* z, i64 n) { static_cast<Precond*>(p)->apply(r, z, n); }
void orc_precond_factor(void* p, i64* outer, i64* inner, double* val, double* scale) {
    const Precond& P = *static_cast<Precond*>(p);
    if (outer) std::copy(P.U.outer.begin(), P.U.outer.end(), outer);
    if (inner) std::copy(P.U.inner.begin(), P.U.inner.end(), inner);
    if (val) std::copy(P.U.val.begin(), P.U.val.end(), val);
    if (scale) std::copy(P.scale.begin(), P.scale.end(), scale);
}

i64 orc_cg(void* a, void* p, const double* b, double* x, double tol, i64 max_it, double* residual) {
    const auto r = cg(static_cast<Assembly*>(a)->A[INNER], *static_cast<Precond*>(p), b, x, tol, max_it);
    if (residual) *residual = r.residual;
    return r.iterations;
}

void orc_nearest_qnodes(int nn, int* out) {
    const Tables2D t = make_tables(nn);
    std::copy(t.nearest.begin(), t.nearest.end(), out);
}

void orc_element_areas(void* h, double* out) {
    const Mesh& m = *static_cast<Mesh*>(h);
    for (i64 e = 0; e < m.E; ++e) out[e] = element_area(m, e);
}

void orc_qnodes_to_nodes(void* h, const double* x, double* out) { qnodes_to_nodes(*static_cast<Mesh*>(h), x, out); }

// Thermal flux of temperature T (N values): qflux [2][E*nq] and nflux [2][N], either may be NULL.
int orc_heat_flux(void* h, const nlf_model* models, const nlf_thermal_material* materials, const double* T,
                  double* qflux, double* nflux) {
    try {
        const Mesh& m = *static_cast<Mesh*>(h);
        std::vector<Group> groups;
        for (int g = 0; g < m.G(); ++g) groups.push_back(make_group(models[g], 0, &materials[g], nullptr, false));
        heat_flux(m, groups, T, qflux, nflux);
        return 0;
    } catch (const std::exception& e) { g_error = e.what(); return -1; }
}

// integrate_right_part (base/right_part_2d.hpp:10-35): for every node, over its elements in
// ascending order, f[dof*node + d] += sum_q w_q N_i(q) |J_q| s_d(x_q); s from fn(x, y, group, d, user).
void orc_right_part(void* h, int dof, ValueFn fn, void* user, double* f) {
    const Mesh& m = *static_cast<Mesh*>(h);
    const int nq = m.t.nq;
    for (i64 node = 0; node < m.N; ++node)
        for (i64 k = m.ne_off[node]; k < m.ne_off[node + 1]; ++k) {
            const i64 e = m.ne_ids[k];
            const int i = m.local(e, node);
            for (int d = 0; d < dof; ++d) {
                double integral = 0.0;
                for (int q = 0; q < nq; ++q) {
                    const double* x = m.xqp(e, q);
                    integral += m.t.w[q] * m.t.qN[i * nq + q] * m.jac[e * nq + q] * fn(x[0], x[1], m.group_of[e], d, user);
                }
                f[dof * node + d] += integral;
            }
        }
}

int orc_strain_stress(void* h, const nlf_model* models, const nlf_elastic_material* materials, const double* u,
                      const double* dT, double* qstress, double* nstrain, double* nstress) {
    try {
        const Mesh& m = *static_cast<Mesh*>(h);
        std::vector<Group> groups;
        for (int g = 0; g < m.G(); ++g) groups.push_back(make_group(models[g], 1, nullptr, &materials[g], false));
        strain_stress(m, groups, u, dT, qstress, nstrain, nstress);
        return 0;
    } catch (const std::exception& e) { g_error = e.what(); return -1; }
}

// the Hooke matrix of a material in its kind's index order (elastic_parameters.hpp:62-133)
int orc_hooke(const nlf_elastic_material* e, int* kind, double* h) {
    try {
        hooke_matrix(*e, *kind, h);
        return 0;
    } catch (const std::exception& ex) { g_error = ex.what(); return -1; }
}

// f[2N] += the thermal load of the temperature change dT[N] (temperature_condition)
int orc_temperature_condition(void* h, const nlf_model* models, const nlf_elastic_material* materials, const double* dT,
                              double* f) {
    try {
        const Mesh& m = *static_cast<Mesh*>(h);
        std::vector<Group> groups;
        for (int g = 0; g < m.G(); ++g) {
            groups.push_back(make_group(models[g], 1, nullptr, &materials[g], false));
            if (groups.back().nonlocal && m.nb_off.empty())
                throw std::logic_error("the thermal load of a nonlocal group needs the neighbour lists");
        }
        temperature_condition(m, groups, dT, f);
        return 0;
    } catch (const std::exception& e) { g_error = e.what(); return -1; }
}

}  // extern "C"
