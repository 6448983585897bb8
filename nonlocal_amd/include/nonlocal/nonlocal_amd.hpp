// C++ mirror of NonLocFEM's 2D solver API over the C ABI of libnlf.so (include/nlf.h).  Namespaces,
// class and function names, template parameters and argument meanings are the reference's, so that
// host code written against the reference (its tests, its NonLocFEM driver) keeps its shape:
//
//   std::stringstream stream{solid_ring_su2_data};
//   const auto mesh = std::make_shared<mesh_2d<T, I>>(stream, mesh_format::SU2);
//   const parameters_2d<T> parameters = {{"DEFAULT", {.physical = {.conductivity = T{1}}}}};
//   thermal_boundaries_conditions_2d<T> boundaries_conditions;
//   boundaries_conditions["Inner"] = std::make_unique<temperature_2d<T>>(0.0);
//   const auto solution = stationary_heat_equation_solver_2d<I>(mesh, parameters, boundaries_conditions, {});
//
// Matrices, solver vectors and neighbour lists live on the GPU behind the handles.  Differences
// from the reference interface, all forced by the boundary:
//   * T = double, indices int64_t (as the reference's CLI and tests use);
//   * influence functions are the families the GPU evaluates (solver_2d::influence::constant,
//     polynomial, exponential, fast_polynomial) instead of arbitrary callables, and a distance is one
//     of the two the reference defines (mesh::powered_distance with any n, the rotated ellipse);
//   * neighbours_t holds the lists on the GPU: find_neighbours runs there and mesh_2d::neighbours
//     adopts its result; a list is copied to the host only when read (neighbours(e));
//   * coefficients are constants or functions of the point (solution-dependent ones, and
//     radiation, need the reference's nonlinear iteration and are refused);
//   * solution vectors are std::vector<double> instead of Eigen vectors;
//   * SU2 element numbering follows the file (groups in marker order, DEFAULT last) instead of
//     the reference reader's hash order (su2_parser.hpp:78-85); element ids only change the order of
//     floating-point sums.
// Errors raise nonlocal::nlf_error (the reference throws std::domain_error / logic_error).
#pragma once

#include <nlf.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <functional>
#include <istream>
#include <limits>
#include <memory>
#include <optional>
#include <ranges>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <variant>
#include <vector>

namespace nonlocal {

struct nlf_error : std::runtime_error {
    int code;
    nlf_error(int c, const std::string& m) : std::runtime_error("nlf error " + std::to_string(c) + ": " + m), code(c) {}
};

inline void nlf_check(int code, nlf_ctx ctx = nullptr) {
    if (code != NLF_OK) throw nlf_error(code, ctx ? nlf_last_error(ctx) : "invalid argument");
}

// One library context (device + stream); the reference is CPU-only, so this has no counterpart.
class context {
    nlf_ctx h_ = nullptr;

public:
    explicit context(int device = 0, void* stream = nullptr) { nlf_check(nlf_ctx_create(device, stream, &h_)); }
    ~context() { if (h_) nlf_ctx_destroy(h_); }
    context(const context&) = delete;
    context& operator=(const context&) = delete;
    nlf_ctx handle() const { return h_; }
    void synchronize() const { nlf_check(nlf_ctx_synchronize(h_), h_); }
    static context& default_context() {
        static context c(0);
        return c;
    }
};

// Device array of doubles (the solver's vectors); host copies in and out.
class device_vector {
    double* p_ = nullptr;
    size_t n_ = 0;

public:
    device_vector() = default;
    explicit device_vector(size_t n);
    explicit device_vector(const std::vector<double>& h) : device_vector(h.size()) { upload(h); }
    ~device_vector();
    device_vector(const device_vector&) = delete;
    device_vector& operator=(const device_vector&) = delete;
    device_vector(device_vector&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    device_vector& operator=(device_vector&& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    double* data() { return p_; }
    const double* data() const { return p_; }
    size_t size() const { return n_; }
    void upload(const std::vector<double>& h);
    std::vector<double> host() const;
};

inline constexpr size_t X = 0, Y = 1;                  // axis indices of points and flux components

enum class theory_t : bool { LOCAL, NONLOCAL };        // nonlocal_constants.hpp:16-19

// coefficient_t (solvers/base/coefficient.hpp:20-36): a constant or a function of the point.
template<class T, size_t Dimension>
using spatial_dependency = std::function<T(const std::array<T, Dimension>&)>;
template<class T, size_t Dimension>
using coefficient_t = std::variant<T, spatial_dependency<T, Dimension>>;

template<class T, size_t Dimension>
bool is_constant(const coefficient_t<T, Dimension>& coefficient) noexcept {
    return std::holds_alternative<T>(coefficient);
}
template<class T, size_t Dimension, size_t N>
bool is_constant(const std::array<coefficient_t<T, Dimension>, N>& coefficient) noexcept {
    return std::all_of(coefficient.begin(), coefficient.end(), [](const auto& c) { return is_constant(c); });
}
template<class T, size_t Dimension>
T evaluate(const coefficient_t<T, Dimension>& coefficient, const std::array<T, Dimension>& point) {
    return std::holds_alternative<T>(coefficient) ? std::get<T>(coefficient) : std::get<1>(coefficient)(point);
}
template<class T, size_t Dimension, size_t N>
std::array<T, N> evaluate(const std::array<coefficient_t<T, Dimension>, N>& coefficient, const std::array<T, Dimension>& point) {
    std::array<T, N> out;
    for (size_t i = 0; i < N; ++i) out[i] = evaluate(coefficient[i], point);
    return out;
}

// evaluated_parameters (equation_parameters.hpp:13-16): a constant, or the values at a group's
// quadrature points (index (e - first element of the group) * nq + q).
template<class T>
using evaluated_parameters = std::variant<T, std::vector<T>>;
inline constexpr size_t Constant = 0, Variable = 1;

// size_t_or (metamath/types/size_t_or.hpp:7-17): an exponent held as an integer or a real;
// Infinity (metamath/utils/constants.hpp) is the largest size_t and selects the maximum norm.
namespace metamath {
namespace types {
template<class T>
using size_t_or = std::variant<size_t, T>;
template<class T>
T get_value(const size_t_or<T>& value) {
    return std::holds_alternative<size_t>(value) ? T(std::get<size_t>(value)) : std::get<T>(value);
}
}  // namespace types
namespace constants {
template<class T>
inline constexpr T Infinity = std::numeric_limits<T>::max();
}  // namespace constants
}  // namespace metamath

namespace mesh {

// powered_distance (search_function.hpp:28-73): the Lp distance sum |(x_i - y_i) / r_i|^n with n an
// integer, a real or Infinity (then the maximum norm, not powered).  operator() evaluates it on the
// host as the reference does; the GPU evaluates the same operations in the search and the influences.
template<class T>
struct powered_distance {
    metamath::types::size_t_or<T> n = size_t{2};

    explicit powered_distance() noexcept = default;
    explicit powered_distance(const metamath::types::size_t_or<T> parameter) : n{parameter} {
        if (!(metamath::types::get_value(n) > T{0})) throw std::domain_error{"the distance exponent shall be greater than 0"};
    }

    T operator()(const std::array<T, 2>& x, const std::array<T, 2>& y, const std::array<T, 2>& r) const {
        const std::array<T, 2> z = {(x[0] - y[0]) / r[0], (x[1] - y[1]) / r[1]};
        if (std::holds_alternative<size_t>(n) && std::get<size_t>(n) == metamath::constants::Infinity<size_t>)
            return std::max(std::abs(z[0]), std::abs(z[1]));
        T sum = T{0};
        for (const T v : z) {
            if (std::holds_alternative<T>(n)) sum = sum + std::pow(std::abs(v), std::get<T>(n));
            else if (std::get<size_t>(n) == 2) sum = sum + v * v;
            else {
                T base = std::abs(v), result = T{1};              // power by squaring (power.hpp:25-40)
                for (size_t e = std::get<size_t>(n); e > 0;)
                    if (e & 1) { result *= base; --e; } else { base *= base; e >>= 1; }
                sum = sum + result;
            }
        }
        return sum;
    }
};

// powered_distance_with_rotation (search_function.hpp:75-90): an ellipse whose r0 axis points along
// x from the origin and whose r1 axis is perpendicular to it; n is 2.
template<class T>
struct powered_distance_with_rotation {
    static constexpr metamath::types::size_t_or<T> n = size_t{2};

    explicit powered_distance_with_rotation(const metamath::types::size_t_or<T>& = {}) noexcept {}

    T operator()(const std::array<T, 2>& x, const std::array<T, 2>& y, const std::array<T, 2>& r) const {
        const auto square = [](const T v) { return v * v; };
        T norm_x = T{0};
        norm_x = norm_x + square(x[0]);
        norm_x = norm_x + square(x[1]);
        return (square(r[0] * (x[0] * y[1] - x[1] * y[0])) + square(r[1] * (x[0] * (x[0] - y[0]) + x[1] * (x[1] - y[1])))) /
               (square(r[0] * r[1]) * norm_x);
    }
};

// distance_function (search_function.hpp:13-14): the reference's std::function, here one of the two
// distances the GPU evaluates
template<class T>
using distance_function = std::variant<powered_distance<T>, powered_distance_with_rotation<T>>;

}  // namespace mesh

namespace detail {

// (kind, n, n_real) of the C ABI for a distance
template<class T>
nlf_distance to_nlf(const mesh::distance_function<T>& distance) {
    nlf_distance d{NLF_DISTANCE_LP, 2, 0.0};
    if (std::holds_alternative<mesh::powered_distance_with_rotation<T>>(distance)) {
        d.kind = NLF_DISTANCE_ELLIPSE_WITH_ROTATION;
        return d;
    }
    const auto& n = std::get<mesh::powered_distance<T>>(distance).n;
    if (std::holds_alternative<T>(n)) {
        d.n = 0;
        d.n_real = double(std::get<T>(n));
    } else if (std::get<size_t>(n) == metamath::constants::Infinity<size_t>) {
        d.n = NLF_DISTANCE_N_INFINITY;
    } else {
        if (std::get<size_t>(n) > size_t(std::numeric_limits<int32_t>::max()))
            throw nlf_error(NLF_ERR_ARGUMENT, "integer distance exponent out of range");
        d.n = int32_t(std::get<size_t>(n));
    }
    return d;
}

}  // namespace detail

// The influence functions the GPU evaluates (influence_functions_2d.hpp:39-148), as a descriptor;
// the reference stores a callable in model_parameters::influence.  p and q with an integral value
// act as the integer alternative of size_t_or.
template<class T>
struct influence_function_2d {
    int32_t kind = NLF_INFLUENCE_POLYNOMIAL;
    std::array<T, 2> radius = {T{0}, T{0}};
    mesh::distance_function<T> distance;
    T p = T{2}, q = T{1};
    bool nonlocal = false;                             // false: the default (zero) influence
};

// model_parameters / equation_parameters (solvers/base/equation_parameters.hpp:18-30)
template<size_t Dimension, class T>
struct model_parameters final {
    influence_function_2d<T> influence;
    T local_weight = T{1};
};

template<size_t Dimension, class T, template<class> class Physical>
struct equation_parameters final {
    model_parameters<Dimension, T> model;
    Physical<T> physical;
};

namespace solver_2d::influence {

// constant (influence_functions_2d.hpp:39-59)
template<class T, class Distance = mesh::powered_distance<T>>
struct constant : influence_function_2d<T> {
    explicit constant(const std::array<T, 2>& radius, const metamath::types::size_t_or<T>& n = size_t{2}) {
        this->kind = NLF_INFLUENCE_CONSTANT; this->radius = radius; this->distance = Distance{n};
        this->nonlocal = true;
    }
};
// polynomial (:61-92): norm (1 - d^(p/n))^q inside the unit ball of the distance
template<class T, class Distance = mesh::powered_distance<T>>
struct polynomial : influence_function_2d<T> {
    explicit polynomial(const std::array<T, 2>& radius, const metamath::types::size_t_or<T>& n = size_t{2},
                        const metamath::types::size_t_or<T>& p = size_t{2}, const metamath::types::size_t_or<T>& q = size_t{1}) {
        this->kind = NLF_INFLUENCE_POLYNOMIAL; this->radius = radius; this->distance = Distance{n};
        this->p = metamath::types::get_value(p); this->q = metamath::types::get_value(q);
        this->nonlocal = true;
    }
};
// exponential (:94-125); the default parameters give the normal distribution
template<class T, class Distance = mesh::powered_distance<T>>
struct exponential : influence_function_2d<T> {
    explicit exponential(const std::array<T, 2>& radius, const metamath::types::size_t_or<T>& n = size_t{2},
                         const metamath::types::size_t_or<T>& p = size_t{2}, const T q = T{0.5}) {
        this->kind = NLF_INFLUENCE_EXPONENTIAL; this->radius = radius; this->distance = Distance{n};
        this->p = metamath::types::get_value(p); this->q = q;
        this->nonlocal = true;
    }
};
// fast_polynomial (:127-148): the polynomial with n = 2, p = 2, q = 1
template<class T, class Distance = mesh::powered_distance<T>>
struct fast_polynomial : polynomial<T, Distance> {
    explicit fast_polynomial(const std::array<T, 2>& radius) : polynomial<T, Distance>(radius) {}
};

}  // namespace solver_2d::influence

namespace detail {

template<size_t Dimension, class T>
nlf_model to_nlf(const model_parameters<Dimension, T>& m) {
    const auto& f = m.influence;
    const bool nonlocal = f.nonlocal && m.local_weight < 0.999;
    const nlf_distance d = to_nlf(f.distance);
    nlf_model out{};
    out.local_weight = nonlocal ? double(m.local_weight) : 1.0;
    out.radius[0] = double(f.radius[0]); out.radius[1] = double(f.radius[1]);
    out.influence = f.kind;
    out.distance_n = d.n;
    out.p = double(f.p); out.q = double(f.q);
    out.distance = d.kind;
    out.distance_real = d.n_real;
    return out;
}

template<size_t Dimension, class T>
theory_t theory(const model_parameters<Dimension, T>& m) noexcept {
    return m.influence.nonlocal && m.local_weight < 0.999 ? theory_t::NONLOCAL : theory_t::LOCAL;
}

}  // namespace detail

// ---- mesh ----------------------------------------------------------------------------------------
namespace mesh {

enum class mesh_format : uint8_t { SU2 };                                 // mesh_parser.hpp:7-9
enum class diameter_expanding_strategy : uint8_t { NO, MAX, MIN, MEAN };   // find_neighbours.hpp:7

// Arrays describing a 2D mesh: nodes, elements (reference local node order), contiguous element
// groups and boundary edge groups, each group with its name.
struct mesh_2d_data {
    std::vector<std::array<double, 2>> nodes;
    int nodes_per_element = 4;                       // 4 (Q4) or 8 (Q8)
    std::vector<int32_t> elements;                   // E * nodes_per_element
    std::vector<std::string> group_names;
    std::vector<int64_t> group_begin;                // groups + 1
    int nodes_per_edge = 2;                          // 2 (Q4) or 3 (Q8)
    std::vector<int32_t> boundary_edges;             // edges * nodes_per_edge
    std::vector<std::string> boundary_names;
    std::vector<int64_t> boundary_begin;             // boundary groups + 1
};

inline const std::string Default_Group_Name = "DEFAULT";   // mesh_container_2d.hpp:16

// mesh_parser<T, I, mesh_format::SU2>::parse (su2_parser.hpp:165-195).  Sections NELEM (element
// lines "type n0 n1 ... index"), NPOIN ("x y index") and NMARK (named markers, lines "type n0 n1 ...").
// An element line counts once per section (the reference keeps them in sets, :74-88); a marker whose
// first element is two-dimensional is an element group whose elements leave the NELEM set, and the
// rest form the group DEFAULT (:140-149, :172-178); one-dimensional markers are boundary groups.  Node
// orders per VTK type (:49-72): bilinear quads (9) as read, serendipity quads (23) from corners then
// midsides to c0 m01 c1 m12 c2 m23 c3 m30, quadratic edges (21) from (a, b, mid) to (a, mid, b).
// Quadrilateral meshes only (Q4 with linear edges, Q8 with quadratic edges).
inline mesh_2d_data read_su2(std::istream& stream) {
    enum : int { LINEAR = 3, QUADRATIC = 21, TRIANGLE = 5, QUADRATIC_TRIANGLE = 22, BILINEAR = 9, SERENDIPITY = 23,
                 LAGRANGE = 28 };
    auto perm_of = [](int kind) -> std::vector<int> {
        switch (kind) {
            case LINEAR: return {0, 1};
            case QUADRATIC: return {0, 2, 1};
            case BILINEAR: return {0, 1, 2, 3};
            case SERENDIPITY: return {0, 2, 4, 6, 1, 3, 5, 7};
            case TRIANGLE: case QUADRATIC_TRIANGLE: case LAGRANGE:
                throw nlf_error(NLF_ERR_UNSUPPORTED, "SU2 element type " + std::to_string(kind) + " is not supported");
            default: throw nlf_error(NLF_ERR_ARGUMENT, "Unknown element type.");
        }
    };
    auto parse_element = [&](const std::string& line, int& kind) {
        std::istringstream in(line);
        in >> kind;
        const std::vector<int> perm = perm_of(kind);
        std::vector<int32_t> ids(perm.size());
        for (size_t k = 0; k < perm.size(); ++k) {
            int64_t v;
            if (!(in >> v)) throw nlf_error(NLF_ERR_ARGUMENT, "element line with too few node ids: " + line);
            ids[size_t(perm[k])] = int32_t(v);
        }
        return ids;
    };
    auto trim = [](std::string s) {
        while (!s.empty() && (s.back() == ' ' || s.back() == '\t' || s.back() == '\r')) s.pop_back();
        size_t b = 0;
        while (b < s.size() && (s[b] == ' ' || s[b] == '\t')) ++b;
        return s.substr(b);
    };
    auto next = [&]() {
        std::string line;
        while (std::getline(stream, line)) {
            line = trim(line);
            if (!line.empty()) return line;
        }
        throw nlf_error(NLF_ERR_ARGUMENT, "unexpected end of SU2 file");
    };
    auto keyword = [&](const std::string& name) {
        const std::string line = next();
        const size_t eq = line.find('=');
        if (eq == std::string::npos || trim(line.substr(0, eq)) != name)
            throw nlf_error(NLF_ERR_ARGUMENT, "expected " + name + "=, found " + line);
        return trim(line.substr(eq + 1));
    };
    auto unique = [](std::vector<std::string> lines) {
        std::unordered_set<std::string> seen;
        std::vector<std::string> out;
        for (auto& l : lines)
            if (seen.insert(l).second) out.push_back(std::move(l));
        return out;
    };
    keyword("NDIME");
    const int64_t n_el = std::stoll(keyword("NELEM"));
    std::vector<std::string> elements;
    for (int64_t k = 0; k < n_el; ++k) {
        std::string line = next();
        const size_t sp = line.find_last_of(" \t");
        elements.push_back(sp == std::string::npos ? line : trim(line.substr(0, sp)));   // drop the index
    }
    elements = unique(std::move(elements));
    std::istringstream npoin(keyword("NPOIN"));
    int64_t n_nodes = 0;
    npoin >> n_nodes;
    mesh_2d_data d;
    d.nodes.resize(size_t(n_nodes));
    for (auto& node : d.nodes) {
        std::istringstream in(next());
        in >> node[0] >> node[1];
    }
    const int64_t n_markers = std::stoll(keyword("NMARK"));
    std::vector<std::pair<std::string, std::vector<std::string>>> markers;
    for (int64_t m = 0; m < n_markers; ++m) {
        std::string name = keyword("MARKER_TAG");
        const int64_t count = std::stoll(keyword("MARKER_ELEMS"));
        std::vector<std::string> lines;
        for (int64_t k = 0; k < count; ++k) lines.push_back(next());
        if (name == Default_Group_Name)
            throw nlf_error(NLF_ERR_ARGUMENT, "The group name cannot be named \"DEFAULT\", as it is a reserved group name.");
        markers.emplace_back(std::move(name), unique(std::move(lines)));
    }
    auto is_2d = [&](const std::string& line) {
        int kind;
        parse_element(line, kind);
        return kind == BILINEAR || kind == SERENDIPITY;
    };
    std::vector<std::pair<std::string, std::vector<std::string>>> groups_2d, groups_1d;
    for (auto& [name, lines] : markers) {
        if (lines.empty()) continue;
        (is_2d(lines.front()) ? groups_2d : groups_1d).emplace_back(name, lines);
    }
    std::unordered_set<std::string> in_groups;
    for (const auto& g : groups_2d) in_groups.insert(g.second.begin(), g.second.end());
    std::vector<std::string> rest;
    for (const auto& l : elements)
        if (!in_groups.count(l)) rest.push_back(l);
    if (!rest.empty()) groups_2d.emplace_back(Default_Group_Name, std::move(rest));
    if (groups_2d.empty()) throw nlf_error(NLF_ERR_ARGUMENT, "the mesh has no 2D elements");
    int kind2d = -1;
    d.group_begin = {0};
    for (const auto& [name, lines] : groups_2d) {
        for (const auto& l : lines) {
            int kind;
            const auto ids = parse_element(l, kind);
            if (kind != BILINEAR && kind != SERENDIPITY) throw nlf_error(NLF_ERR_ARGUMENT, "a 2D group holds the 1D element " + l);
            if (kind2d >= 0 && kind != kind2d) throw nlf_error(NLF_ERR_UNSUPPORTED, "meshes mixing element types are not supported");
            kind2d = kind;
            d.elements.insert(d.elements.end(), ids.begin(), ids.end());
        }
        d.group_names.push_back(name);
        d.group_begin.push_back(int64_t(d.elements.size()) / (kind2d == BILINEAR ? 4 : 8));
    }
    d.nodes_per_element = kind2d == BILINEAR ? 4 : 8;
    d.nodes_per_edge = kind2d == BILINEAR ? 2 : 3;
    const int edge_kind = kind2d == BILINEAR ? LINEAR : QUADRATIC;
    d.boundary_begin = {0};
    for (const auto& [name, lines] : groups_1d) {
        for (const auto& l : lines) {
            int kind;
            const auto ids = parse_element(l, kind);
            if (kind != edge_kind) throw nlf_error(NLF_ERR_ARGUMENT, "boundary element " + l + " does not match the quads");
            d.boundary_edges.insert(d.boundary_edges.end(), ids.begin(), ids.end());
        }
        d.boundary_names.push_back(name);
        d.boundary_begin.push_back(int64_t(d.boundary_edges.size()) / d.nodes_per_edge);
    }
    return d;
}

// mesh_container_2d (mesh_container_2d.hpp): the mesh's arrays as the reference's container exposes them.
template<class T = double, class I = int64_t>
class mesh_container_2d {
    const mesh_2d_data* d_;

public:
    explicit mesh_container_2d(const mesh_2d_data& d) : d_(&d) {}
    size_t nodes_count() const noexcept { return d_->nodes.size(); }
    auto nodes() const noexcept { return std::views::iota(size_t{0}, nodes_count()); }
    std::array<T, 2> node_coord(const size_t node) const { return d_->nodes[node]; }
    size_t elements_2d_count() const noexcept { return d_->elements.size() / size_t(d_->nodes_per_element); }
    auto elements_2d() const noexcept { return std::views::iota(size_t{0}, elements_2d_count()); }
    size_t nodes_count(const size_t) const noexcept { return size_t(d_->nodes_per_element); }
    I node_number(const size_t e, const size_t i) const { return I(d_->elements[e * size_t(d_->nodes_per_element) + i]); }
    const std::vector<std::string>& groups_2d() const noexcept { return d_->group_names; }
    const std::vector<std::string>& groups_1d() const noexcept { return d_->boundary_names; }
    std::pair<size_t, size_t> elements(const std::string& group) const {
        for (size_t g = 0; g < d_->group_names.size(); ++g)
            if (d_->group_names[g] == group) return {size_t(d_->group_begin[g]), size_t(d_->group_begin[g + 1])};
        throw std::domain_error("unknown element group " + group);
    }
    const mesh_2d_data& data() const noexcept { return *d_; }
};

template<class T>
struct influence final {                          // search_function.hpp:15-19
    distance_function<T> distance;
    std::array<T, 2> radius = {T{0}, T{0}};
};

template<class T>
using influences = std::unordered_map<std::string, influence<T>>;

template<class T, class I>
class mesh_2d;

// The neighbour lists of one search, resident on the GPU (the reference holds them in a
// std::vector<std::vector<I>>): size() elements, operator[](e) copies one list to the host.
template<class I>
class neighbours_list {
    nlf_mesh h_ = nullptr;
    nlf_ctx ctx_ = nullptr;
    size_t elements_ = 0;
    uint64_t generation_ = 0;
    template<class, class> friend class mesh_2d;

public:
    neighbours_list() = default;
    neighbours_list(nlf_mesh h, nlf_ctx ctx, size_t elements, uint64_t generation)
        : h_(h), ctx_(ctx), elements_(elements), generation_(generation) {}
    size_t size() const noexcept { return elements_; }
    uint64_t generation() const noexcept { return generation_; }
    std::vector<I> operator[](const size_t e) const {
        int64_t total = 0;
        nlf_check(nlf_neighbours_count(h_, &total), ctx_);
        std::vector<int64_t> off(elements_ + 1);
        std::vector<int32_t> ids(size_t(std::max<int64_t>(total, 1)));
        nlf_check(nlf_neighbours_export(h_, off.data(), ids.data()), ctx_);
        return std::vector<I>(ids.begin() + off[e], ids.begin() + off[e + 1]);
    }
};

template<class T, class I>
using neighbours_t = std::pair<influences<T>, neighbours_list<I>>;   // search_function.hpp:25-26

// mesh_2d<T, I> (mesh_2d.hpp:18-81): geometry, quadrature precompute and neighbour lists on the GPU.
template<class T = double, class I = int64_t>
class mesh_2d {
    static_assert(std::is_same<T, double>::value, "the GPU path computes in double");
    nlf_mesh h_ = nullptr;
    context* ctx_;
    mesh_2d_data data_;
    mesh_container_2d<T, I> container_;
    influences<T> influences_;
    uint64_t searches_ = 0, adopted_ = 0;

    void create() {
        const auto& d = data_;
        std::vector<double> xy(2 * d.nodes.size());
        for (size_t i = 0; i < d.nodes.size(); ++i) { xy[2 * i] = d.nodes[i][0]; xy[2 * i + 1] = d.nodes[i][1]; }
        const int64_t E = int64_t(d.elements.size()) / d.nodes_per_element;
        const int64_t B = d.nodes_per_edge ? int64_t(d.boundary_edges.size()) / d.nodes_per_edge : 0;
        nlf_check(nlf_mesh_create(ctx_->handle(), int64_t(d.nodes.size()), xy.data(), E, d.nodes_per_element,
                                  d.elements.data(), int(d.group_begin.size()) - 1, d.group_begin.data(), B,
                                  d.nodes_per_edge, d.boundary_edges.data(), int(d.boundary_begin.size()) - 1,
                                  d.boundary_begin.data(), &h_),
                  ctx_->handle());
    }
    template<class U, class J>
    friend neighbours_t<U, J> find_neighbours(const mesh_2d<U, J>&, const influences<U>&, diameter_expanding_strategy);

public:
    explicit mesh_2d(mesh_2d_data data, context& ctx = context::default_context())
        : ctx_(&ctx), data_(std::move(data)), container_(data_) { create(); }
    // mesh_2d(stream, format) (mesh_2d.hpp:44, 88-92)
    mesh_2d(std::istream& stream, const mesh_format format, context& ctx = context::default_context())
        : ctx_(&ctx), data_(format == mesh_format::SU2 ? read_su2(stream) : mesh_2d_data{}), container_(data_) { create(); }
    // mesh_2d(path) (mesh_2d.hpp:42, 83-86): the format from the extension
    explicit mesh_2d(const std::filesystem::path& path, context& ctx = context::default_context()) : ctx_(&ctx), container_(data_) {
        if (path.extension() != ".su2") throw nlf_error(NLF_ERR_ARGUMENT, "Unknown mesh format: " + path.string());
        std::ifstream in(path);
        if (!in) throw nlf_error(NLF_ERR_ARGUMENT, "cannot open " + path.string());
        data_ = read_su2(in);
        container_ = mesh_container_2d<T, I>(data_);
        create();
    }
    ~mesh_2d() { if (h_) nlf_mesh_destroy(h_); }
    mesh_2d(const mesh_2d&) = delete;
    mesh_2d& operator=(const mesh_2d&) = delete;

    const mesh_container_2d<T, I>& container() const noexcept { return container_; }
    nlf_mesh handle() const { return h_; }
    context& ctx() const { return *ctx_; }
    const mesh_2d_data& data() const { return data_; }
    size_t nodes_count() const { return data_.nodes.size(); }
    size_t elements_count() const { return data_.elements.size() / size_t(data_.nodes_per_element); }
    size_t groups_count() const { return data_.group_names.size(); }
    int group_index(const std::string& name) const {
        for (size_t g = 0; g < data_.group_names.size(); ++g)
            if (data_.group_names[g] == name) return int(g);
        throw std::domain_error("unknown element group " + name);
    }

    // mesh_2d::neighbours(neighbours_t&&) (mesh_2d.hpp:205-212): adopts the lists of the latest search
    void neighbours(neighbours_t<T, I>&& data) {
        auto&& [infl, lists] = data;
        if (lists.size() != elements_count())
            throw std::domain_error{"The neighbor list length does not match the number of 2D mesh elements."};
        if (lists.h_ != h_ || lists.generation_ != searches_)
            throw nlf_error(NLF_ERR_STATE, "the neighbour lists are not those of this mesh's latest find_neighbours");
        influences_ = std::move(infl);
        adopted_ = searches_;
    }
    // mesh_2d::neighbours(e) (mesh_2d.hpp:214-217), copied from the GPU
    std::vector<I> neighbours(const size_t e) const {
        if (!has_neighbours()) throw nlf_error(NLF_ERR_STATE, "no neighbour lists on this mesh");
        return neighbours_list<I>(h_, ctx_->handle(), elements_count(), searches_)[e];
    }
    bool has_neighbours() const noexcept { return adopted_ != 0 && adopted_ == searches_; }
    const influences<T>& get_influences() const noexcept { return influences_; }
};

// find_neighbours (find_neighbours.hpp:119-160): the search (and the radius expansion) on the GPU;
// the lists stay there until mesh_2d::neighbours adopts them.
template<class T, class I>
neighbours_t<T, I> find_neighbours(const mesh_2d<T, I>& mesh, const influences<T>& influence,
                                   const diameter_expanding_strategy strategy = diameter_expanding_strategy::MAX) {
    const size_t G = mesh.groups_count();
    std::vector<double> r(2 * G, 0.0), expanded(2 * G, 0.0);
    std::vector<nlf_distance> distances(G, nlf_distance{NLF_DISTANCE_LP, 2, 0.0});
    for (size_t g = 0; g < G; ++g) {
        const auto it = influence.find(mesh.data().group_names[g]);
        if (it == influence.end()) continue;
        distances[g] = ::nonlocal::detail::to_nlf(it->second.distance);
        r[2 * g] = it->second.radius[0];
        r[2 * g + 1] = it->second.radius[1];
    }
    nlf_check(nlf_find_neighbours_ex(mesh.handle(), r.data(), distances.data(), int(strategy), expanded.data()),
              mesh.ctx().handle());
    auto& m = const_cast<mesh_2d<T, I>&>(mesh);       // the search's generation: the lists now on the GPU
    ++m.searches_;
    influences<T> out;
    for (size_t g = 0; g < G; ++g)
        if (r[2 * g] > 0) {
            const auto& name = mesh.data().group_names[g];
            out[name] = mesh::influence<T>{influence.at(name).distance, {expanded[2 * g], expanded[2 * g + 1]}};
        }
    return {std::move(out), neighbours_list<I>(mesh.handle(), mesh.ctx().handle(), mesh.elements_count(), m.searches_)};
}

namespace utils {

// integrate (mesh_2d_utils.hpp:135-149): sum_e sum_q w_q x_i N_i(q) |J_e(q)| in the reference's loop
// order, accumulated in long double.
template<class T, class I>
T integrate(const mesh_2d<T, I>& mesh, const std::vector<T>& x) {
    if (mesh.nodes_count() != x.size())
        throw std::logic_error{"The integral cannot be found because the vector size does not match the number of nodes."};
    int64_t N = 0, E = 0;
    int nn = 0, nq = 0;
    nlf_check(nlf_mesh_info(mesh.handle(), &N, &E, &nn, &nq), mesh.ctx().handle());
    std::vector<double> w(static_cast<size_t>(nq)), qN(static_cast<size_t>(nn * nq)), jac(static_cast<size_t>(E * nq));
    nlf_check(nlf_mesh_export_basis(mesh.handle(), w.data(), qN.data()), mesh.ctx().handle());
    nlf_check(nlf_mesh_export_quadrature(mesh.handle(), nullptr, jac.data(), nullptr, nullptr), mesh.ctx().handle());
    const auto& el = mesh.data().elements;
    long double integral = 0;
    for (int64_t e = 0; e < E; ++e)
        for (int q = 0; q < nq; ++q)
            for (int i = 0; i < nn; ++i)
                integral += w[size_t(q)] * x[size_t(el[size_t(e * nn + i)])] * qN[size_t(i * nq + q)] * jac[size_t(e * nq + q)];
    return T(integral);
}

}  // namespace utils

}  // namespace mesh

// ---- matrices -----------------------------------------------------------------------------------

// Host CSR with int64 indices (Eigen::SparseMatrix<T, RowMajor, J> layout: outer, inner, values).
struct csr_matrix {
    std::vector<int64_t> outer, inner;
    std::vector<double> values;
};

// finite_element_matrix (base/finite_element_matrix.hpp:12-21): inner (upper CSR of the
// Dirichlet-eliminated matrix) and bound parts, held on the GPU.
class finite_element_matrix {
    nlf_matrix h_ = nullptr;
    nlf_ctx ctx_ = nullptr;

    csr_matrix part(int which) const {
        int64_t rows, up, full, bound;
        nlf_check(nlf_matrix_info(h_, &rows, &up, &full, &bound), ctx_);
        const int64_t nnz = (which == 0 || which == 3) ? up : which == 1 ? bound : full;
        csr_matrix m;
        m.outer.resize(size_t(rows + 1));
        m.inner.resize(size_t(std::max<int64_t>(nnz, 1)));
        m.values.resize(size_t(std::max<int64_t>(nnz, 1)));
        nlf_check(nlf_matrix_export(h_, which, m.outer.data(), m.inner.data(), m.values.data()), ctx_);
        m.inner.resize(size_t(nnz));
        m.values.resize(size_t(nnz));
        return m;
    }

public:
    finite_element_matrix() = default;
    finite_element_matrix(nlf_matrix h, nlf_ctx ctx) : h_(h), ctx_(ctx) {}
    ~finite_element_matrix() { if (h_) nlf_matrix_destroy(h_); }
    finite_element_matrix(const finite_element_matrix&) = delete;
    finite_element_matrix& operator=(const finite_element_matrix&) = delete;
    finite_element_matrix(finite_element_matrix&& o) noexcept : h_(o.h_), ctx_(o.ctx_) { o.h_ = nullptr; }
    finite_element_matrix& operator=(finite_element_matrix&& o) noexcept {
        if (h_) nlf_matrix_destroy(h_);
        h_ = o.h_; ctx_ = o.ctx_; o.h_ = nullptr;
        return *this;
    }
    nlf_matrix handle() const { return h_; }
    nlf_ctx ctx() const { return ctx_; }
    int64_t rows() const {
        int64_t r, a, b, c;
        nlf_check(nlf_matrix_info(h_, &r, &a, &b, &c), ctx_);
        return r;
    }
    csr_matrix inner() const { return part(0); }
    csr_matrix bound() const { return part(1); }
    csr_matrix inner_full() const { return part(2); }
    // FP32 values for solve(..., NLF_MODE_MIXED): the FP64 values rounded, or an FP32 assembly
    csr_matrix inner_fp32() const { return part(3); }
    void make_fp32() { nlf_check(nlf_matrix_make_fp32(h_), ctx_); }
    void assemble_fp32() { nlf_check(nlf_matrix_assemble_fp32(h_), ctx_); }
};

namespace solver_2d {

enum class assemble_part : uint8_t { LOCAL, NONLOCAL, FULL };   // matrix_assembler_2d.hpp:15-19

// problem_settings (solver_2d/base/problem_settings.hpp:14-35)
struct problem_settings final {
    std::unordered_map<std::string, theory_t> theories;
    bool is_neumann = false;
    bool is_nonlinear_boundary = false;
    bool is_nonconstant_parameters = false;
    bool is_solution_dependent = false;
    std::vector<bool> is_inner_nodes;

    constexpr bool is_nonlinear() const noexcept { return is_nonlinear_boundary || is_solution_dependent; }
    bool is_nonlocal() const {
        return std::any_of(theories.begin(), theories.end(), [](const auto& t) { return t.second == theory_t::NONLOCAL; });
    }
    bool is_symmetric() const { return !(is_nonlocal() && is_nonconstant_parameters); }
};

namespace detail {

inline int nlf_part(const assemble_part part) {
    if (part == assemble_part::NONLOCAL)
        throw nlf_error(NLF_ERR_UNSUPPORTED, "the nonlocal part alone is not assembled (LOCAL or FULL)");
    return part == assemble_part::LOCAL ? NLF_PART_LOCAL : NLF_PART_FULL;
}

// values of a coefficient-valued parameter at the quadrature points of each element of a group
template<class T, class I, class F>
auto at_quadrature(const mesh::mesh_2d<T, I>& mesh, const std::string& group, const F& value) {
    int64_t N = 0, E = 0;
    int nn = 0, nq = 0;
    nlf_check(nlf_mesh_info(mesh.handle(), &N, &E, &nn, &nq), mesh.ctx().handle());
    std::vector<double> xq(size_t(2 * E * nq));
    nlf_check(nlf_mesh_export_quadrature(mesh.handle(), xq.data(), nullptr, nullptr, nullptr), mesh.ctx().handle());
    const auto [b, e] = mesh.container().elements(group);
    std::vector<std::invoke_result_t<F, std::array<T, 2>>> out;
    out.reserve((e - b) * size_t(nq));
    for (size_t el = b; el < e; ++el)
        for (int q = 0; q < nq; ++q)
            out.push_back(value(std::array<T, 2>{xq[2 * (el * size_t(nq) + size_t(q))], xq[2 * (el * size_t(nq) + size_t(q)) + 1]}));
    return out;
}

// A boundary value as a function of the point, one per boundary group and degree of freedom,
// through nlf_boundary_data_fn.
using point_function = std::function<double(const std::array<double, 2>&)>;
inline void boundary_data(nlf_mesh mesh, nlf_ctx ctx, int dof, const std::vector<nlf_boundary>& bcs,
                          const std::vector<std::array<point_function, 2>>& functions, std::vector<double>& f,
                          std::vector<double>& xD, std::vector<uint8_t>& inner) {
    struct closure {
        const std::vector<nlf_boundary>* bcs;
        const std::vector<std::array<point_function, 2>>* fn;
    } c{&bcs, &functions};
    const nlf_boundary_value_fn value = [](double x, double y, int group, int d, void* user) -> double {
        const auto& cl = *static_cast<const closure*>(user);
        const point_function& g = (*cl.fn)[size_t(group)][size_t(d)];
        return g ? g({x, y}) : (*cl.bcs)[size_t(group)].value[d];
    };
    nlf_check(nlf_boundary_data_fn(mesh, dof, bcs.data(), value, &c, f.data(), xD.data(), inner.data()), ctx);
}

// integrate_right_part (base/right_part_2d.hpp:10-35): f += int s N_i for DoF d (dof per node),
// s evaluated on the host at the quadrature points, integrated on the GPU
template<class T, class I>
void integrate_right_part(device_vector& f, const mesh::mesh_2d<T, I>& mesh, int dof, int d,
                          const std::function<T(const std::array<T, 2>&)>& s) {
    int64_t N = 0, E = 0;
    int nn = 0, nq = 0;
    nlf_check(nlf_mesh_info(mesh.handle(), &N, &E, &nn, &nq), mesh.ctx().handle());
    std::vector<double> xq(size_t(2 * E * nq)), values(size_t(dof * E * nq), 0.0);
    nlf_check(nlf_mesh_export_quadrature(mesh.handle(), xq.data(), nullptr, nullptr, nullptr), mesh.ctx().handle());
    for (size_t k = 0; k < size_t(E * nq); ++k) values[size_t(d) * size_t(E * nq) + k] = s({xq[2 * k], xq[2 * k + 1]});
    device_vector dv(values);
    nlf_check(nlf_right_part(mesh.handle(), dof, dv.data(), f.data()), mesh.ctx().handle());
}

}  // namespace detail

}  // namespace solver_2d

// ---- SLAE ---------------------------------------------------------------------------------------
namespace slae {

// preconditioner_base<T, I> (slae/preconditioner_base.hpp:7-15): compute() sets the factorisation
// up from a matrix; solve() applies M^-1; computation_info() reports success (never throws).
class preconditioner_base {
protected:
    nlf_precond h_ = nullptr;
    nlf_ctx ctx_ = nullptr;

public:
    virtual ~preconditioner_base() { if (h_) nlf_precond_destroy(h_); }
    virtual void compute(const finite_element_matrix& matrix) = 0;
    bool computation_info() const {
        if (!h_) return false;
        int ok = 0, fw = 0, bw = 0;
        double shift = 0, sec = 0;
        nlf_check(nlf_precond_info(h_, &ok, &shift, &fw, &bw, &sec), ctx_);
        return ok != 0;
    }
    nlf_precond handle() const { return h_; }
    std::vector<double> solve(const std::vector<double>& r) const {
        device_vector dr(r), dz(r.size());
        nlf_check(nlf_precond_apply(h_, dr.data(), dz.data()), ctx_);
        return dz.host();
    }
};

template<int KIND>
class device_preconditioner : public preconditioner_base {
public:
    device_preconditioner() = default;
    explicit device_preconditioner(const finite_element_matrix& matrix) { compute(matrix); }
    void compute(const finite_element_matrix& matrix) override {
        if (h_) nlf_precond_destroy(h_);
        h_ = nullptr;
        ctx_ = matrix.ctx();
        nlf_check(nlf_precond_create(KIND == NLF_PRECOND_IDENTITY ? nullptr : matrix.handle(), KIND, &h_), ctx_);
    }
};

using eigen_identity_preconditioner = device_preconditioner<NLF_PRECOND_IDENTITY>;   // eigen_preconditioner.hpp:50-51
using eigen_diagonal_preconditioner = device_preconditioner<NLF_PRECOND_JACOBI>;     // :53-54
using eigen_ILLT_preconditioner = device_preconditioner<NLF_PRECOND_ILLT>;           // :56-57 (IC(0) here)

// conjugate_gradient<T, I> (slae/conjugate_gradient.hpp:15-42): keeps a non-owning reference to the
// matrix and owns its preconditioner; tolerance 1e-15 and 10000 iterations by default
// (iterative_solver_base.hpp:12-13).
template<class T = double, class I = int64_t>
class conjugate_gradient {
    const finite_element_matrix& matrix_;
    std::unique_ptr<preconditioner_base> precond_;
    T tol_ = T{1e-15};
    uintmax_t max_it_ = 10000;
    mutable nlf_solve_stats stats_{};

public:
    explicit conjugate_gradient(const finite_element_matrix& matrix) : matrix_(matrix) {}

    template<class P, class... Args>
    preconditioner_base& init_preconditioner(Args&&... args) {
        precond_ = std::make_unique<P>(std::forward<Args>(args)...);
        return *precond_;
    }
    const preconditioner_base* preconditioner() const { return precond_.get(); }
    T tolerance() const { return tol_; }
    uintmax_t max_iterations() const { return max_it_; }
    void set_tolerance(T tol) { tol_ = tol; }
    void set_max_interations(uintmax_t it) { max_it_ = it; }      // reference spelling (iterative_solver_base.hpp:49-51)
    uintmax_t iterations() const { return uintmax_t(stats_.iterations); }
    T residual() const { return stats_.residual; }
    const nlf_solve_stats& statistics() const { return stats_; }

    // solve(b, x0) (:48-79) with device-resident vectors
    void solve(const device_vector& b, device_vector& x, int mode = NLF_MODE_FP64) const {
        nlf_check(nlf_cg_solve(matrix_.handle(), precond_ ? precond_->handle() : nullptr, b.data(), x.data(), tol_,
                               int64_t(max_it_), mode, &stats_),
                  matrix_.ctx());
    }
    std::vector<T> solve(const std::vector<T>& b, const std::optional<std::vector<T>>& x0 = std::nullopt) const {
        device_vector db(b), dx(b.size());
        dx.upload(x0 ? *x0 : std::vector<T>(b.size(), T{0}));
        solve(db, dx);
        return dx.host();
    }
};

}  // namespace slae

// ---- thermal ------------------------------------------------------------------------------------
namespace solver_2d::thermal {

// conductivity (thermal_parameters_2d.hpp:7-42): l, {l11, l22} or {l11, l22, l12}
template<class T>
using raw_conductivity_t = std::variant<coefficient_t<T, 2>, std::array<coefficient_t<T, 2>, 2>, std::array<coefficient_t<T, 2>, 3>>;
template<class T>
using evaluated_conductivity_t = std::variant<evaluated_parameters<T>, evaluated_parameters<std::array<T, 2>>,
                                              evaluated_parameters<std::array<T, 3>>>;

template<class T>
struct parameter_2d final {                            // thermal_parameters_2d.hpp:44-50
    raw_conductivity_t<T> conductivity = T{1};
    T capacity = T{1};
    T density = T{1};
    T relaxation_time = T{0};
};

template<class T>
using parameters_2d = std::unordered_map<std::string, equation_parameters<2, T, parameter_2d>>;
template<class T>
using evaluated_conductivity_2d = std::unordered_map<std::string, equation_parameters<2, T, evaluated_conductivity_t>>;

// evaluate_conductivity (thermal/evaluate_conductivity.hpp:16-37): constants stay constants, functions of
// the point are evaluated at the group's quadrature points
template<class T, class I>
evaluated_conductivity_2d<T> evaluate_conductivity(const mesh::mesh_2d<T, I>& mesh, const parameters_2d<T>& parameters) {
    evaluated_conductivity_2d<T> out;
    for (const auto& [group, p] : parameters) {
        if (!mesh.container().elements(group).second && !mesh.groups_count()) continue;
        auto& r = out[group];
        r.model = p.model;
        std::visit([&](const auto& c) {
            using C = std::decay_t<decltype(c)>;
            if constexpr (std::is_same_v<C, coefficient_t<T, 2>>) {
                if (is_constant(c)) r.physical = evaluated_parameters<T>{std::in_place_index<Constant>, std::get<T>(c)};
                else r.physical = evaluated_parameters<T>{std::in_place_index<Variable>, solver_2d::detail::at_quadrature(
                                                              mesh, group, [&c](const std::array<T, 2>& x) { return evaluate(c, x); })};
            } else {
                using A = std::array<T, std::tuple_size_v<C>>;
                if (is_constant(c)) r.physical = evaluated_parameters<A>{std::in_place_index<Constant>, evaluate(c, std::array<T, 2>{})};
                else r.physical = evaluated_parameters<A>{std::in_place_index<Variable>, solver_2d::detail::at_quadrature(
                                                              mesh, group, [&c](const std::array<T, 2>& x) { return evaluate(c, x); })};
            }
        }, p.physical.conductivity);
    }
    return out;
}

// conductivity_matrix_2d<T, I, J> (thermal/conductivity_matrix_2d.hpp:11-33)
template<class T = double, class I = int64_t, class J = int64_t>
class conductivity_matrix_2d {
    std::shared_ptr<mesh::mesh_2d<T, I>> mesh_;
    finite_element_matrix matrix_;

public:
    explicit conductivity_matrix_2d(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh) : mesh_(mesh) {}

    // compute (:166-193): the conductivity per group, constant or at the quadrature points
    void compute(const evaluated_conductivity_2d<T>& conductivity, const std::vector<bool>& is_inner,
                 const bool is_symmetric = true, const bool is_neumann = false, const assemble_part part = assemble_part::FULL) {
        const size_t G = mesh_->groups_count();
        std::vector<nlf_model> models(G);
        std::vector<nlf_thermal_material> mats(G);
        std::vector<std::vector<double>> values(G);
        std::vector<const double*> coefficients(G, nullptr);
        bool varying = false;
        for (size_t g = 0; g < G; ++g) {
            const auto& p = conductivity.at(mesh_->data().group_names[g]);
            models[g] = nonlocal::detail::to_nlf(p.model);
            std::visit([&](const auto& c) {
                using C = std::decay_t<decltype(c)>;
                using V = std::variant_alternative_t<0, C>;
                constexpr size_t nc = [] {
                    if constexpr (std::is_same_v<V, T>) return size_t{1}; else return std::tuple_size_v<V>;
                }();
                constexpr int kind = nc == 1 ? NLF_CONDUCTIVITY_ISOTROPIC
                                   : nc == 2 ? NLF_CONDUCTIVITY_ORTHOTROPIC : NLF_CONDUCTIVITY_ANISOTROPIC;
                mats[g] = nlf_thermal_material{kind, 0, {1.0, 1.0, 0.0}};
                auto put = [&](const V& v, double* dst) {
                    if constexpr (std::is_same_v<V, T>) dst[0] = v;
                    else for (size_t k = 0; k < v.size(); ++k) dst[k] = v[k];
                };
                if (c.index() == Constant) {
                    put(std::get<Constant>(c), mats[g].c);
                } else {
                    const auto& vs = std::get<Variable>(c);
                    values[g].resize(vs.size() * nc);
                    for (size_t k = 0; k < vs.size(); ++k) put(vs[k], values[g].data() + k * nc);
                    coefficients[g] = values[g].data();
                    varying = true;
                }
            }, p.physical);
        }
        std::vector<uint8_t> inner(is_inner.begin(), is_inner.end());
        const nlf_assemble_options opt{int32_t(is_neumann), 0, varying ? coefficients.data() : nullptr};
        nlf_matrix h = nullptr;
        nlf_check(nlf_assemble_ex(mesh_->handle(), 0, models.data(), mats.data(), inner.data(), int(is_symmetric),
                                  solver_2d::detail::nlf_part(part), &opt, &h),
                  mesh_->ctx().handle());
        matrix_ = finite_element_matrix(h, mesh_->ctx().handle());
    }
    finite_element_matrix& matrix() { return matrix_; }
    const finite_element_matrix& matrix() const { return matrix_; }
};

// Boundary conditions (thermal_boundary_conditions_2d.hpp): first kind (temperature_2d) or second kind
// (flux_2d, convection_2d, radiation_2d, combined_flux_2d), constants or functions of the point.
template<class T>
class thermal_boundary_condition_2d {
protected:
    template<class U>
    static std::function<T(const std::array<T, 2>&)> from_value(const U& value) {
        if constexpr (std::is_arithmetic_v<U>) return [v = T(value)](const std::array<T, 2>&) { return v; };
        else return value;
    }

public:
    virtual ~thermal_boundary_condition_2d() noexcept = default;
    virtual T operator()(const std::array<T, 2>& x) const = 0;
    virtual int kind() const noexcept = 0;               // NLF_BC_FIRST_KIND or NLF_BC_FLUX
    virtual T heat_transfer() const noexcept { return T{0}; }
    virtual T emissivity() const noexcept { return T{0}; }
};

template<class T>
class temperature_2d final : public thermal_boundary_condition_2d<T> {
    const std::function<T(const std::array<T, 2>&)> temperature_;

public:
    template<class U>
    explicit temperature_2d(const U& temperature) : temperature_{this->from_value(temperature)} {}
    T operator()(const std::array<T, 2>& x) const override { return temperature_(x); }
    int kind() const noexcept override { return NLF_BC_FIRST_KIND; }
};

template<class T>
class flux_2d : public thermal_boundary_condition_2d<T> {
    const std::function<T(const std::array<T, 2>&)> flux_;

public:
    template<class U>
    explicit flux_2d(const U& flux) : flux_{this->from_value(flux)} {}
    T operator()(const std::array<T, 2>& x) const override { return flux_(x); }
    int kind() const noexcept override { return NLF_BC_FLUX; }
};

// convection_2d: heat_transfer (T_ambient - T); operator() is h T_ambient (the right-hand side part)
template<class T>
class convection_2d : public thermal_boundary_condition_2d<T> {
    const std::function<T(const std::array<T, 2>&)> ambient_temperature_;
    const T heat_transfer_;

public:
    template<class U>
    explicit convection_2d(const T heat_transfer, const U& ambient_temperature)
        : ambient_temperature_{this->from_value(ambient_temperature)}, heat_transfer_{heat_transfer} {}
    T operator()(const std::array<T, 2>& x) const override { return heat_transfer_ * ambient_temperature_(x); }
    int kind() const noexcept override { return NLF_BC_FLUX; }
    T heat_transfer() const noexcept override { return heat_transfer_; }
};

// radiation_2d needs the reference's nonlinear (Newton) iteration: refused by the solvers here
template<class T>
class radiation_2d : public thermal_boundary_condition_2d<T> {
    const T emissivity_;

public:
    explicit radiation_2d(const T emissivity) : emissivity_{emissivity} {}
    T operator()(const std::array<T, 2>&) const override { return T{0}; }
    int kind() const noexcept override { return NLF_BC_FLUX; }
    T emissivity() const noexcept override { return emissivity_; }
};

// combined_flux_2d: flux + convection (+ radiation, refused when the emissivity is not 0)
template<class T>
class combined_flux_2d final : public thermal_boundary_condition_2d<T> {
    const std::function<T(const std::array<T, 2>&)> flux_, ambient_temperature_;
    const T heat_transfer_, emissivity_;

public:
    template<class Flux, class Ambient_Temperature>
    explicit combined_flux_2d(const Flux& flux, const T heat_transfer, const Ambient_Temperature& ambient_temperature,
                              const T emissivity)
        : flux_{this->from_value(flux)}, ambient_temperature_{this->from_value(ambient_temperature)},
          heat_transfer_{heat_transfer}, emissivity_{emissivity} {}
    T operator()(const std::array<T, 2>& x) const override { return flux_(x) + heat_transfer_ * ambient_temperature_(x); }
    int kind() const noexcept override { return NLF_BC_FLUX; }
    T heat_transfer() const noexcept override { return heat_transfer_; }
    T emissivity() const noexcept override { return emissivity_; }
};

template<class T>
using thermal_boundaries_conditions_2d = std::unordered_map<std::string, std::unique_ptr<thermal_boundary_condition_2d<T>>>;

namespace detail {

// The boundary records of the C ABI.  bcs drives the boundary data: first-kind groups, and every
// second-kind group as a flux whose value function is the condition's operator() (for convection
// h T_ambient(x), the reference's second-kind right-hand side).  convection carries the convection
// groups' h for the matrix term h int N_i N_j (nlf_matrix_add_convection).
template<class T, class I>
void boundaries(const mesh::mesh_2d<T, I>& mesh, const thermal_boundaries_conditions_2d<T>& conditions,
                std::vector<nlf_boundary>& bcs, std::vector<nlf_boundary>& convection,
                std::vector<std::array<solver_2d::detail::point_function, 2>>& functions) {
    const size_t GB = mesh.data().boundary_names.size();
    const nlf_boundary none{{NLF_BC_NONE, NLF_BC_NONE}, {0.0, 0.0}, 0.0, 0.0};
    bcs.assign(GB, none);
    convection.assign(GB, none);
    functions.assign(GB, {});
    for (size_t b = 0; b < GB; ++b) {
        const auto it = conditions.find(mesh.data().boundary_names[b]);
        if (it == conditions.end() || !it->second) continue;
        const auto& c = *it->second;
        if (c.emissivity() != T{0})
            throw nlf_error(NLF_ERR_UNSUPPORTED, "radiation needs the nonlinear iteration, which is not implemented");
        bcs[b].kind[0] = c.kind();
        if (c.heat_transfer() != T{0}) {
            convection[b].kind[0] = NLF_BC_CONVECTION;
            convection[b].heat_transfer = c.heat_transfer();
        }
        const thermal_boundary_condition_2d<T>* pc = &c;
        functions[b][0] = [pc](const std::array<double, 2>& x) { return (*pc)(x); };
    }
}

}  // namespace detail

// init_problem_settings (thermal/init_problem_settings.hpp:11-48), stationary or not
template<class T, class I>
problem_settings init_problem_settings(const mesh::mesh_2d<T, I>&, const parameters_2d<T>& parameters,
                                       const thermal_boundaries_conditions_2d<T>& conditions, const bool is_stationary) {
    problem_settings s;
    for (const auto& [group, p] : parameters) {
        s.theories[group] = nonlocal::detail::theory(p.model);
        s.is_nonconstant_parameters |= std::visit([](const auto& c) { return !is_constant(c); }, p.physical.conductivity);
    }
    bool all_second_kind = true;
    for (const auto& [name, c] : conditions)
        if (c) {
            all_second_kind &= c->kind() != NLF_BC_FIRST_KIND && c->heat_transfer() == T{0};
            s.is_nonlinear_boundary |= c->emissivity() != T{0};
        }
    s.is_neumann = is_stationary && all_second_kind;
    return s;
}

// heat_equation_solution_2d (thermal/heat_equation_solution_2d.hpp:10-137): the temperature and the
// heat flux; calc_flux runs nlf_heat_flux on the GPU (gradient, -lambda grad T, nonlocal averaging
// over the neighbour lists, nearest-quadrature-point node values) and keeps the result.
template<class T = double, class I = int64_t>
class heat_equation_solution_2d {
    std::shared_ptr<mesh::mesh_2d<T, I>> mesh_;
    std::vector<nlf_model> models_;
    std::vector<nlf_thermal_material> materials_;
    std::vector<T> temperature_;
    std::array<std::vector<T>, 2> flux_;

public:
    uintmax_t iterations = 0;                          // of the solve that produced the temperature
    T residual = 0, true_residual = 0;

    heat_equation_solution_2d(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh, const parameters_2d<T>& parameters,
                              std::vector<T> temperature)
        : mesh_(mesh), temperature_(std::move(temperature)) {
        if (temperature_.size() < mesh_->nodes_count())
            throw nlf_error(NLF_ERR_ARGUMENT, "The temperature vector is shorter than the number of nodes.");
        temperature_.resize(mesh_->nodes_count());       // a Neumann solution ends with its multiplier
        for (size_t g = 0; g < mesh_->groups_count(); ++g) {
            const auto& p = parameters.at(mesh_->data().group_names[g]);
            models_.push_back(nonlocal::detail::to_nlf(p.model));
            nlf_thermal_material m{NLF_CONDUCTIVITY_ISOTROPIC, 0, {1.0, 1.0, 0.0}};
            std::visit([&m](const auto& c) {
                using C = std::decay_t<decltype(c)>;
                if constexpr (std::is_same_v<C, coefficient_t<T, 2>>) {
                    if (!is_constant(c)) throw nlf_error(NLF_ERR_UNSUPPORTED, "the flux of a varying conductivity is not implemented");
                    m.c[0] = std::get<T>(c);
                } else {
                    if (!is_constant(c)) throw nlf_error(NLF_ERR_UNSUPPORTED, "the flux of a varying conductivity is not implemented");
                    m.kind = std::tuple_size_v<C> == 2 ? NLF_CONDUCTIVITY_ORTHOTROPIC : NLF_CONDUCTIVITY_ANISOTROPIC;
                    for (size_t k = 0; k < c.size(); ++k) m.c[k] = std::get<T>(c[k]);
                }
            }, p.physical.conductivity);
            materials_.push_back(m);
        }
    }

    const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh_proxy() const noexcept { return mesh_; }
    const std::vector<T>& temperature() const noexcept { return temperature_; }
    bool is_flux_calculated() const noexcept { return !flux_[0].empty() && !flux_[1].empty(); }
    const std::array<std::vector<T>, 2>& flux() const {
        if (!is_flux_calculated()) throw std::runtime_error{"Flux wasn't calculated"};
        return flux_;
    }
    // The neighbour lists of a nonlocal group must be on the mesh, as in the reference.
    const std::array<std::vector<T>, 2>& calc_flux() {
        if (is_flux_calculated()) return flux_;
        const size_t n = mesh_->nodes_count();
        device_vector dT(temperature_), dflux(2 * n);
        nlf_check(nlf_heat_flux(mesh_->handle(), models_.data(), materials_.data(), dT.data(), dflux.data(), nullptr),
                  mesh_->ctx().handle());
        const std::vector<double> h = dflux.host();
        flux_[0].assign(h.begin(), h.begin() + std::ptrdiff_t(n));
        flux_[1].assign(h.begin() + std::ptrdiff_t(n), h.end());
        return flux_;
    }
};

// stationary_equation_parameters_2d (stationary_heat_equation_solver_2d.hpp:19-26); tolerance and
// max_iterations bound the reference's nonlinear (radiation) iteration, which is refused here
template<class T>
struct stationary_equation_parameters_2d final {
    std::optional<std::function<T(const std::array<T, 2>&)>> right_part;
    std::optional<std::function<T(const std::array<T, 2>&)>> initial_distribution;
    T tolerance = std::is_same_v<T, float> ? 1e-5 : 1e-8;
    size_t max_iterations = 100;
    T energy = T{0};
};

inline constexpr double NEUMANN_PROBLEM_ERROR_THRESHOLD = 1e-10;   // nonlocal_constants.hpp:45

// stationary_heat_equation_solver_2d<Matrix_Index> (stationary_heat_equation_solver_2d.hpp:46-120):
// the right-hand side (and, for a Neumann problem, its solvability and the energy row), assembly,
// convection, first-kind conditions, CG with ILLT of the local matrix (identity when the factor
// fails), then the flux.  The neighbour lists of nonlocal groups must be on the mesh.  The system
// is solved directly (the reference's correction form from T0 is the same linear system).
template<class Matrix_Index = int64_t, class T, class I>
heat_equation_solution_2d<T, I> stationary_heat_equation_solver_2d(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh,
                                                                   const parameters_2d<T>& parameters,
                                                                   const thermal_boundaries_conditions_2d<T>& boundaries_conditions,
                                                                   const stationary_equation_parameters_2d<T>& auxiliary_data) {
    static constexpr bool Is_Stationary = true;
    const auto settings = init_problem_settings(*mesh, parameters, boundaries_conditions, Is_Stationary);
    if (settings.is_nonlinear()) throw nlf_error(NLF_ERR_UNSUPPORTED, "nonlinear (radiation) problems are not implemented");
    if (!settings.is_symmetric())
        throw nlf_error(NLF_ERR_UNSUPPORTED, "a nonlocal problem with varying coefficients is asymmetric (BiCGSTAB in the reference)");
    nlf_ctx ctx = mesh->ctx().handle();
    std::vector<nlf_boundary> bcs, convection;
    std::vector<std::array<solver_2d::detail::point_function, 2>> functions;
    detail::boundaries(*mesh, boundaries_conditions, bcs, convection, functions);
    const size_t N = mesh->nodes_count(), rows = N + size_t(settings.is_neumann);
    std::vector<double> f(N), xD(N);
    std::vector<uint8_t> inner8(N);
    solver_2d::detail::boundary_data(mesh->handle(), ctx, 1, bcs, functions, f, xD, inner8);
    std::vector<bool> is_inner(inner8.begin(), inner8.end());
    if (settings.is_neumann) is_inner.push_back(true);
    device_vector df(f);
    if (auxiliary_data.right_part) solver_2d::detail::integrate_right_part<T, I>(df, *mesh, 1, 0, *auxiliary_data.right_part);
    if (settings.is_neumann) {                         // init_right_part (:28-44)
        std::vector<double> h = df.host();
        long double sum = 0;
        for (const double v : h) sum += v;
        if (std::abs(double(sum)) > NEUMANN_PROBLEM_ERROR_THRESHOLD) throw std::domain_error{"It's unsolvable Neumann problem."};
        h.push_back(auxiliary_data.energy);
        xD.push_back(0.0);
        df = device_vector(h);
    }
    const auto conductivity = evaluate_conductivity(*mesh, parameters);
    conductivity_matrix_2d<T, I, Matrix_Index> K(mesh);
    K.compute(conductivity, is_inner, settings.is_symmetric(), settings.is_neumann);
    nlf_check(nlf_matrix_add_convection(K.matrix().handle(), convection.data()), ctx);
    device_vector dxD(xD), dx(rows);
    nlf_check(nlf_apply_first_kind(K.matrix().handle(), dxD.data(), df.data()), ctx);
    slae::conjugate_gradient<T, Matrix_Index> solver{K.matrix()};
    conductivity_matrix_2d<T, I, Matrix_Index> KL(mesh);
    KL.compute(conductivity, is_inner, settings.is_symmetric(), settings.is_neumann, assemble_part::LOCAL);
    auto& P = solver.template init_preconditioner<slae::eigen_ILLT_preconditioner>(KL.matrix());
    if (!P.computation_info()) solver.template init_preconditioner<slae::eigen_identity_preconditioner>(K.matrix());
    dx.upload(std::vector<double>(rows, 0.0));
    solver.solve(df, dx);
    heat_equation_solution_2d<T, I> solution{mesh, parameters, dx.host()};
    solution.iterations = solver.iterations();
    solution.residual = solver.residual();
    solution.true_residual = solver.statistics().true_residual;
    solution.calc_flux();                              // as the reference's solver does (:117-119)
    return solution;
}

// nonstationary_heat_equation_solver_2d<T, I, Matrix_Index> (thermal/nonstationary_heat_equation_solver_2d.hpp:
// 17-106): implicit Euler (C + tau K) T_{n+1} = C T_n + tau (f_boundary + f_source).  compute() assembles
// tau K + C on the GPU (nlf_matrix_combine); calc_step() forms the right-hand side, applies the
// first-kind conditions and runs CG (no preconditioner, as in the reference) from T_n.
template<class T = double, class I = int64_t, class Matrix_Index = int64_t>
class nonstationary_heat_equation_solver_2d final {
    std::shared_ptr<mesh::mesh_2d<T, I>> mesh_;
    conductivity_matrix_2d<T, I, Matrix_Index> conductivity_;
    finite_element_matrix capacity_;
    std::unique_ptr<slae::conjugate_gradient<T, Matrix_Index>> solver_;
    device_vector temperature_;
    std::vector<uint8_t> inner_;
    T time_step_ = 1;

    void boundary(const thermal_boundaries_conditions_2d<T>& conditions, std::vector<double>& f, std::vector<double>& xD,
                  std::vector<uint8_t>& inner, std::vector<nlf_boundary>& convection) const {
        std::vector<nlf_boundary> bcs;
        std::vector<std::array<solver_2d::detail::point_function, 2>> functions;
        detail::boundaries(*mesh_, conditions, bcs, convection, functions);
        const size_t rows = mesh_->nodes_count();
        f.assign(rows, 0.0); xD.assign(rows, 0.0); inner.assign(rows, 1);
        solver_2d::detail::boundary_data(mesh_->handle(), mesh_->ctx().handle(), 1, bcs, functions, f, xD, inner);
    }

public:
    nonstationary_heat_equation_solver_2d(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh, const T time_step)
        : mesh_(mesh), conductivity_(mesh), temperature_(mesh->nodes_count()), time_step_(time_step) {}

    std::vector<T> temperature() const { return temperature_.host(); }
    T time_step() const noexcept { return time_step_; }

    template<class Init_Dist>
    void compute(const parameters_2d<T>& parameters, const thermal_boundaries_conditions_2d<T>& conditions,
                 const Init_Dist& init_dist) {
        const size_t G = mesh_->groups_count();
        std::vector<nlf_model> models(G);
        std::vector<double> capacities(G);
        for (size_t g = 0; g < G; ++g) {
            const auto& p = parameters.at(mesh_->data().group_names[g]);
            capacities[g] = p.physical.density * p.physical.capacity;
            models[g] = nlf_model{1.0, {0.0, 0.0}, {0.0, 0.0}, NLF_INFLUENCE_POLYNOMIAL, 2, 2.0, 1.0};
        }
        std::vector<double> f, xD;
        std::vector<nlf_boundary> bcs;
        boundary(conditions, f, xD, inner_, bcs);
        const std::vector<bool> is_inner(inner_.begin(), inner_.end());
        conductivity_.compute(evaluate_conductivity(*mesh_, parameters), is_inner);
        nlf_check(nlf_matrix_add_convection(conductivity_.matrix().handle(), bcs.data()), mesh_->ctx().handle());
        nlf_matrix hc = nullptr;                       // heat_capacity_matrix_2d::calc_matrix
        nlf_check(nlf_assemble(mesh_->handle(), 2, models.data(), capacities.data(), inner_.data(), 1, NLF_PART_LOCAL, &hc),
                  mesh_->ctx().handle());
        capacity_ = finite_element_matrix(hc, mesh_->ctx().handle());
        nlf_check(nlf_matrix_combine(conductivity_.matrix().handle(), time_step_, capacity_.handle(), 1.0),
                  mesh_->ctx().handle());
        std::vector<double> T0(mesh_->nodes_count());
        for (size_t node = 0; node < T0.size(); ++node) T0[node] = init_dist(mesh_->data().nodes[node]);
        temperature_.upload(T0);
        solver_ = std::make_unique<slae::conjugate_gradient<T, Matrix_Index>>(conductivity_.matrix());
    }

    template<class Right_Part = std::nullptr_t>
    void calc_step(const thermal_boundaries_conditions_2d<T>& conditions, const Right_Part& right_part = nullptr) {
        std::vector<double> f, xD;
        std::vector<uint8_t> inner;
        std::vector<nlf_boundary> bcs;
        boundary(conditions, f, xD, inner, bcs);
        if (inner != inner_) throw nlf_error(NLF_ERR_ARGUMENT, "the first-kind boundary groups changed after compute()");
        device_vector df(f), dxD(xD), ct(mesh_->nodes_count());
        if constexpr (!std::is_same_v<Right_Part, std::nullptr_t>)
            solver_2d::detail::integrate_right_part<T, I>(df, *mesh_, 1, 0, right_part);
        nlf_check(nlf_spmv(capacity_.handle(), temperature_.data(), ct.data()), mesh_->ctx().handle());
        std::vector<double> rhs = df.host();
        const std::vector<double> c = ct.host();
        for (size_t i = 0; i < rhs.size(); ++i) rhs[i] = time_step_ * rhs[i] + c[i];
        df.upload(rhs);
        nlf_check(nlf_apply_first_kind(conductivity_.matrix().handle(), dxD.data(), df.data()), mesh_->ctx().handle());
        solver_->solve(df, temperature_);              // x0 = T_n, result T_{n+1}
    }

    uintmax_t iterations() const { return solver_ ? solver_->iterations() : 0; }
};

}  // namespace solver_2d::thermal

// ---- mechanical ---------------------------------------------------------------------------------
namespace solver_2d::mechanical {

inline constexpr size_t XX = 0, YY = 1, XY = 2;        // strain / stress components

// elastic_parameters (mechanical/elastic_parameters.hpp:62-141)
template<class T>
struct isotropic_elastic_parameters final {
    coefficient_t<T, 2> young_modulus = T{210.};
    coefficient_t<T, 2> poissons_ratio = T{0.3};
    bool is_constant() const { return nonlocal::is_constant(young_modulus) && nonlocal::is_constant(poissons_ratio); }
};
template<class T>
struct orthotropic_elastic_parameters final {
    std::array<coefficient_t<T, 2>, 2> young_modulus = {T{210.}, T{210.}};
    std::array<coefficient_t<T, 2>, 2> poissons_ratio = {T{0.3}, T{0.3}};
    coefficient_t<T, 2> shear_modulus = T{210.} / T{13};
    bool is_constant() const {
        return nonlocal::is_constant(young_modulus) && nonlocal::is_constant(poissons_ratio) && nonlocal::is_constant(shear_modulus);
    }
};
template<class T>
struct anisotropic_elastic_parameters final {
    orthotropic_elastic_parameters<T> main_parameters;
    coefficient_t<T, 2> angle = T{0};
    bool is_constant() const { return main_parameters.is_constant() && nonlocal::is_constant(angle); }
};
template<class T>
using elastic_parameters_t = std::variant<isotropic_elastic_parameters<T>, orthotropic_elastic_parameters<T>,
                                          anisotropic_elastic_parameters<T>>;

// thermal expansion (thermal_expansion_parameters.hpp:7-36): alpha, {a11, a22} or {a11, a22, a12}
template<class T>
using raw_thermal_expansion_t = std::variant<coefficient_t<T, 2>, std::array<coefficient_t<T, 2>, 2>,
                                             std::array<coefficient_t<T, 2>, 3>>;

template<class T>
struct raw_mechanical_parameters_t final {             // mechanical_parameters_2d.hpp:8-12
    elastic_parameters_t<T> elastic;
    raw_thermal_expansion_t<T> thermal_expansion;
};
template<class T>
using raw_mechanical_parameters = std::unordered_map<std::string, equation_parameters<2, T, raw_mechanical_parameters_t>>;

// evaluated_mechanical_parameters_t: the material record per group (constant) and, when an elastic
// coefficient is a function of the point, the Hooke matrix at the group's quadrature points in the
// reference's index order ({11, 12, 66}, {11, 12, 66, 22}, {11, 12, 66, 22, 16, 26})
template<class T>
struct evaluated_mechanical_parameters_t final {
    nlf_elastic_material material{NLF_ELASTIC_ISOTROPIC, NLF_EXPANSION_NONE, 210.0, 0.3, 0, 0, 0, 0, {0, 0, 0}};
    std::vector<T> hooke;                              // empty when constant
};
template<class T>
using evaluated_mechanical_parameters = std::unordered_map<std::string, equation_parameters<2, T, evaluated_mechanical_parameters_t>>;

namespace detail {

// hooke (elastic_parameters.hpp:70-123) of one point: E/(1 - nu^2) forms, the orthotropic matrix,
// the anisotropic one rotated by the angle; components in the reference's index order
template<class T>
std::vector<T> hooke_at(const elastic_parameters_t<T>& e, const std::array<T, 2>& x) {
    return std::visit([&x](const auto& p) -> std::vector<T> {
        using P = std::decay_t<decltype(p)>;
        if constexpr (std::is_same_v<P, isotropic_elastic_parameters<T>>) {
            const T E = evaluate(p.young_modulus, x), nu = evaluate(p.poissons_ratio, x);
            const T v = E / (1 - nu * nu);
            return {v, nu * v, T(0.5) * v * (1 - nu)};
        } else {
            const orthotropic_elastic_parameters<T>* o;
            T angle = 0;
            if constexpr (std::is_same_v<P, orthotropic_elastic_parameters<T>>) o = &p;
            else { o = &p.main_parameters; angle = evaluate(p.angle, x); }
            const auto E = evaluate(o->young_modulus, x), nu = evaluate(o->poissons_ratio, x);
            const T G = evaluate(o->shear_modulus, x);
            const T v = 1 / (1 - nu[0] * nu[1]);
            const T m11 = E[0] * v, m12 = E[0] * nu[1] * v, m66 = G, m22 = E[1] * v;
            if constexpr (std::is_same_v<P, orthotropic_elastic_parameters<T>>) return {m11, m12, m66, m22};
            else {
                const T a2 = angle + angle, s2 = std::sin(a2), c2 = std::cos(a2);
                const T c4 = T(0.125) * (c2 * c2 - s2 * s2);
                const T diff = m11 - m22, core = m11 + m22 - 2 * m12 - 4 * m66;
                const T plus = core * (c4 + T(0.125)), minus = core * (T(0.125) - c4);
                const T cdiff = diff * (c2 + T(0.5)), minor = core * c2, s4 = T(0.25) * s2;
                return {plus + T(0.5) * (m12 + m22 + cdiff) + m66, minus + m12, minus + m66,
                        plus + T(0.5) * (m12 + m11 - cdiff) + m66, s4 * (diff + minor), s4 * (diff - minor)};
            }
        }
    }, e);
}

}  // namespace detail

// evaluate_mechanical_parameters (mechanical/evaluate_mechanical_parameters.hpp:9-30)
template<class T, class I>
evaluated_mechanical_parameters<T> evaluate_mechanical_parameters(const mesh::mesh_2d<T, I>& mesh,
                                                                  const raw_mechanical_parameters<T>& parameters) {
    evaluated_mechanical_parameters<T> out;
    for (const auto& [group, p] : parameters) {
        auto& r = out[group];
        r.model = p.model;
        auto& m = r.physical.material;
        std::visit([&m](const auto& e) {
            using P = std::decay_t<decltype(e)>;
            auto c = [](const coefficient_t<T, 2>& v) { return is_constant(v) ? std::get<T>(v) : T{1}; };
            if constexpr (std::is_same_v<P, isotropic_elastic_parameters<T>>) {
                m = nlf_elastic_material{NLF_ELASTIC_ISOTROPIC, NLF_EXPANSION_NONE, c(e.young_modulus), c(e.poissons_ratio), 0, 0, 0, 0, {0, 0, 0}};
            } else {
                const auto& o = [&]() -> const orthotropic_elastic_parameters<T>& {
                    if constexpr (std::is_same_v<P, orthotropic_elastic_parameters<T>>) return e; else return e.main_parameters;
                }();
                m = nlf_elastic_material{NLF_ELASTIC_ORTHOTROPIC, NLF_EXPANSION_NONE, c(o.young_modulus[0]), c(o.poissons_ratio[0]),
                                         c(o.young_modulus[1]), c(o.poissons_ratio[1]), c(o.shear_modulus), 0, {0, 0, 0}};
                if constexpr (std::is_same_v<P, anisotropic_elastic_parameters<T>>) { m.kind = NLF_ELASTIC_ANISOTROPIC; m.angle = c(e.angle); }
            }
        }, p.physical.elastic);
        std::visit([&m](const auto& a) {
            using A = std::decay_t<decltype(a)>;
            auto c = [](const coefficient_t<T, 2>& v) {
                if (!is_constant(v)) throw nlf_error(NLF_ERR_UNSUPPORTED, "a thermal expansion varying in space is not implemented");
                return std::get<T>(v);
            };
            if constexpr (std::is_same_v<A, coefficient_t<T, 2>>) {
                const T v = c(a);
                if (v != T{0}) { m.expansion_kind = NLF_EXPANSION_ISOTROPIC; m.expansion[0] = v; }
            } else {
                m.expansion_kind = std::tuple_size_v<A> == 2 ? NLF_EXPANSION_ORTHOTROPIC : NLF_EXPANSION_ANISOTROPIC;
                for (size_t k = 0; k < a.size(); ++k) m.expansion[k] = c(a[k]);
            }
        }, p.physical.thermal_expansion);
        const bool constant = std::visit([](const auto& e) { return e.is_constant(); }, p.physical.elastic);
        if (!constant) {
            const auto& el = p.physical.elastic;
            const auto per_point = solver_2d::detail::at_quadrature(mesh, group, [&el](const std::array<T, 2>& x) { return detail::hooke_at<T>(el, x); });
            for (const auto& h : per_point) r.physical.hooke.insert(r.physical.hooke.end(), h.begin(), h.end());
        }
    }
    return out;
}

// Boundary conditions (mechanical_boundary_conditions_2d.hpp): displacement_2d (first kind) and
// pressure_2d (second kind) per degree of freedom, constants or functions of the point.
template<class T>
class mechanical_boundary_condition_2d {
protected:
    template<class U>
    static std::function<T(const std::array<T, 2>&)> from_value(const U& value) {
        if constexpr (std::is_arithmetic_v<U>) return [v = T(value)](const std::array<T, 2>&) { return v; };
        else return value;
    }

public:
    virtual ~mechanical_boundary_condition_2d() noexcept = default;
    virtual T operator()(const std::array<T, 2>& x) const = 0;
    virtual int kind() const noexcept = 0;
};

template<class T>
class displacement_2d final : public mechanical_boundary_condition_2d<T> {
    const std::function<T(const std::array<T, 2>&)> displacement_;

public:
    template<class U>
    explicit displacement_2d(const U& displacement) : displacement_{this->from_value(displacement)} {}
    T operator()(const std::array<T, 2>& x) const override { return displacement_(x); }
    int kind() const noexcept override { return NLF_BC_FIRST_KIND; }
};

template<class T>
class pressure_2d final : public mechanical_boundary_condition_2d<T> {
    const std::function<T(const std::array<T, 2>&)> pressure_;

public:
    template<class U>
    explicit pressure_2d(const U& pressure) : pressure_{this->from_value(pressure)} {}
    T operator()(const std::array<T, 2>& x) const override { return pressure_(x); }
    int kind() const noexcept override { return NLF_BC_FLUX; }
};

template<class T>
using mechanical_boundaries_conditions_2d = std::unordered_map<std::string, std::array<std::unique_ptr<mechanical_boundary_condition_2d<T>>, 2>>;

// init_problem_settings (mechanical/equilibrium_equation_2d.hpp:26)
template<class T, class I>
problem_settings init_problem_settings(const mesh::mesh_2d<T, I>&, const raw_mechanical_parameters<T>& parameters,
                                       const mechanical_boundaries_conditions_2d<T>&) {
    problem_settings s;
    for (const auto& [group, p] : parameters) {
        s.theories[group] = nonlocal::detail::theory(p.model);
        s.is_nonconstant_parameters |= !std::visit([](const auto& e) { return e.is_constant(); }, p.physical.elastic);
    }
    return s;
}

// stiffness_matrix<T, I, J> (mechanical/stiffness_matrix_2d.hpp:10-38), plane stress, 2 DoF per node
template<class T = double, class I = int64_t, class J = int64_t>
class stiffness_matrix {
    std::shared_ptr<mesh::mesh_2d<T, I>> mesh_;
    finite_element_matrix matrix_;

public:
    explicit stiffness_matrix(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh) : mesh_(mesh) {}

    // compute (:171-199): the inner-node mask of settings, symmetric unless nonlocal with varying parameters
    void compute(const evaluated_mechanical_parameters<T>& hooke, const problem_settings& settings,
                 const assemble_part part = assemble_part::FULL) {
        const size_t G = mesh_->groups_count();
        std::vector<nlf_model> models(G);
        std::vector<nlf_elastic_material> mats(G);
        std::vector<const double*> coefficients(G, nullptr);
        bool varying = false;
        for (size_t g = 0; g < G; ++g) {
            const auto& p = hooke.at(mesh_->data().group_names[g]);
            models[g] = nonlocal::detail::to_nlf(p.model);
            mats[g] = p.physical.material;
            if (!p.physical.hooke.empty()) { coefficients[g] = p.physical.hooke.data(); varying = true; }
        }
        std::vector<uint8_t> inner(settings.is_inner_nodes.begin(), settings.is_inner_nodes.end());
        const nlf_assemble_options opt{0, 0, varying ? coefficients.data() : nullptr};
        nlf_matrix h = nullptr;
        nlf_check(nlf_assemble_ex(mesh_->handle(), 1, models.data(), mats.data(), inner.data(), int(settings.is_symmetric()),
                                  solver_2d::detail::nlf_part(part), &opt, &h),
                  mesh_->ctx().handle());
        matrix_ = finite_element_matrix(h, mesh_->ctx().handle());
    }
    finite_element_matrix& matrix() { return matrix_; }
    const finite_element_matrix& matrix() const { return matrix_; }
};

// mechanical_solution_2d (mechanical/mechanical_solution_2d.hpp:11-232): displacement, and on
// calc_strain_and_stress the strains and stresses on the GPU (with the thermal strains of a given
// temperature change).  XX, YY, XY components; the xy strain is the tensor shear.
template<class T = double, class I = int64_t>
class mechanical_solution_2d {
    std::shared_ptr<mesh::mesh_2d<T, I>> mesh_;
    std::vector<nlf_model> models_;
    std::vector<nlf_elastic_material> materials_;
    std::vector<T> solution_, delta_temperature_;
    std::array<std::vector<T>, 2> displacement_;
    std::array<std::vector<T>, 3> strain_, stress_;

public:
    uintmax_t iterations = 0;
    T residual = 0, true_residual = 0;

    mechanical_solution_2d(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh, const evaluated_mechanical_parameters<T>& parameters,
                           std::vector<T> displacement, std::vector<T> delta_temperature = {})
        : mesh_(mesh), solution_(std::move(displacement)), delta_temperature_(std::move(delta_temperature)) {
        const size_t n = mesh_->nodes_count();
        if (solution_.size() < 2 * n)
            throw nlf_error(NLF_ERR_ARGUMENT, "The displacement vector is shorter than twice the number of nodes.");
        solution_.resize(2 * n);
        for (size_t g = 0; g < mesh_->groups_count(); ++g) {
            const auto& p = parameters.at(mesh_->data().group_names[g]);
            if (!p.physical.hooke.empty())
                throw nlf_error(NLF_ERR_UNSUPPORTED, "the stresses of varying elastic parameters are not implemented");
            models_.push_back(nonlocal::detail::to_nlf(p.model));
            materials_.push_back(p.physical.material);
        }
        for (auto& d : displacement_) d.resize(n);
        for (size_t i = 0; i < n; ++i) {
            displacement_[X][i] = solution_[2 * i];
            displacement_[Y][i] = solution_[2 * i + 1];
        }
    }

    const std::array<std::vector<T>, 2>& displacement() const noexcept { return displacement_; }
    const std::array<std::vector<T>, 3>& strain() const noexcept { return strain_; }
    const std::array<std::vector<T>, 3>& stress() const noexcept { return stress_; }
    bool is_strain_and_stress_calculated() const noexcept {
        for (size_t k = 0; k < 3; ++k)
            if (strain_[k].empty() || stress_[k].empty()) return false;
        return true;
    }
    // The neighbour lists of a nonlocal group must be on the mesh, as in the reference.
    void calc_strain_and_stress() {
        if (is_strain_and_stress_calculated()) return;
        const size_t n = mesh_->nodes_count();
        device_vector du(solution_), dstrain(3 * n), dstress(3 * n);
        std::unique_ptr<device_vector> ddt;
        if (!delta_temperature_.empty()) ddt = std::make_unique<device_vector>(delta_temperature_);
        nlf_check(nlf_strain_stress(mesh_->handle(), models_.data(), materials_.data(), du.data(), ddt ? ddt->data() : nullptr,
                                    dstrain.data(), dstress.data(), nullptr),
                  mesh_->ctx().handle());
        const std::vector<double> hs = dstrain.host(), ht = dstress.host();
        for (size_t k = 0; k < 3; ++k) {
            strain_[k].assign(hs.begin() + std::ptrdiff_t(k * n), hs.begin() + std::ptrdiff_t((k + 1) * n));
            stress_[k].assign(ht.begin() + std::ptrdiff_t(k * n), ht.begin() + std::ptrdiff_t((k + 1) * n));
        }
    }
};

// equilibrium_equation<Matrix_Index> (mechanical/equilibrium_equation_2d.hpp:20-62): boundary
// conditions, thermal load of delta_temperature (temperature_condition), body force, stiffness, CG
// with ILLT of the local stiffness when nonlocal (identity otherwise, and on a failed factor), then
// the strains and stresses.  The neighbour lists of nonlocal groups must be on the mesh.
template<class Matrix_Index = int64_t, class T, class I>
mechanical_solution_2d<T, I> equilibrium_equation(
    const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh, const raw_mechanical_parameters<T>& parameters,
    const mechanical_boundaries_conditions_2d<T>& boundaries_conditions, const std::vector<T>& delta_temperature = {},
    const std::optional<std::function<std::array<T, 2>(const std::array<T, 2>&)>>& right_part = std::nullopt) {
    auto settings = init_problem_settings(*mesh, parameters, boundaries_conditions);
    if (!settings.is_symmetric())
        throw nlf_error(NLF_ERR_UNSUPPORTED, "a nonlocal problem with varying coefficients is asymmetric (BiCGSTAB in the reference)");
    nlf_ctx ctx = mesh->ctx().handle();
    const size_t GB = mesh->data().boundary_names.size();
    std::vector<nlf_boundary> bcs(GB, nlf_boundary{{NLF_BC_NONE, NLF_BC_NONE}, {0.0, 0.0}, 0.0, 0.0});
    std::vector<std::array<solver_2d::detail::point_function, 2>> functions(GB);
    for (size_t b = 0; b < GB; ++b) {
        const auto it = boundaries_conditions.find(mesh->data().boundary_names[b]);
        if (it == boundaries_conditions.end()) continue;
        for (size_t d = 0; d < 2; ++d) {
            const auto* c = it->second[d].get();
            if (!c) continue;
            bcs[b].kind[d] = c->kind();
            functions[b][d] = [c](const std::array<double, 2>& x) { return (*c)(x); };
        }
    }
    const size_t rows = 2 * mesh->nodes_count();
    std::vector<double> f(rows), xD(rows);
    std::vector<uint8_t> inner8(rows);
    solver_2d::detail::boundary_data(mesh->handle(), ctx, 2, bcs, functions, f, xD, inner8);
    settings.is_inner_nodes.assign(inner8.begin(), inner8.end());
    const auto hooke = evaluate_mechanical_parameters(*mesh, parameters);
    const size_t G = mesh->groups_count();
    std::vector<nlf_model> models(G);
    std::vector<nlf_elastic_material> mats(G);
    for (size_t g = 0; g < G; ++g) {
        const auto& p = hooke.at(mesh->data().group_names[g]);
        models[g] = nonlocal::detail::to_nlf(p.model);
        mats[g] = p.physical.material;
    }
    device_vector df(f);
    if (!delta_temperature.empty()) {                  // temperature_condition (:36-38)
        if (delta_temperature.size() != mesh->nodes_count())
            throw nlf_error(NLF_ERR_ARGUMENT, "delta_temperature has one value per node");
        device_vector ddt(delta_temperature);
        nlf_check(nlf_temperature_condition(mesh->handle(), models.data(), mats.data(), ddt.data(), df.data()), ctx);
    }
    if (right_part)                                    // integrate_right_part<2> (:39-40)
        for (int d = 0; d < 2; ++d)
            solver_2d::detail::integrate_right_part<T, I>(df, *mesh, 2, d, [&right_part, d](const std::array<T, 2>& x) {
                return (*right_part)(x)[size_t(d)];
            });
    stiffness_matrix<T, I, Matrix_Index> K(mesh);
    K.compute(hooke, settings);
    device_vector dxD(xD), dx(rows);
    nlf_check(nlf_apply_first_kind(K.matrix().handle(), dxD.data(), df.data()), ctx);
    slae::conjugate_gradient<T, Matrix_Index> solver{K.matrix()};
    stiffness_matrix<T, I, Matrix_Index> KL(mesh);
    if (settings.is_nonlocal()) {
        KL.compute(hooke, settings, assemble_part::LOCAL);
        auto& P = solver.template init_preconditioner<slae::eigen_ILLT_preconditioner>(KL.matrix());
        if (!P.computation_info()) solver.template init_preconditioner<slae::eigen_identity_preconditioner>(K.matrix());
    } else {
        solver.template init_preconditioner<slae::eigen_identity_preconditioner>(K.matrix());
    }
    dx.upload(std::vector<double>(rows, 0.0));
    solver.solve(df, dx);
    mechanical_solution_2d<T, I> solution{mesh, hooke, dx.host(), delta_temperature};
    solution.iterations = solver.iterations();
    solution.residual = solver.residual();
    solution.true_residual = solver.statistics().true_residual;
    solution.calc_strain_and_stress();
    return solution;
}

}  // namespace solver_2d::mechanical

}  // namespace nonlocal

// ---- device_vector (HIP runtime) ----------------------------------------------------------------
#include <hip/hip_runtime_api.h>

namespace nonlocal {

inline device_vector::device_vector(size_t n) : n_(n) {
    if (n && hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(double)) != hipSuccess)
        throw nlf_error(NLF_ERR_HIP, "hipMalloc failed");
}

inline device_vector::~device_vector() {
    if (p_) (void)hipFree(p_);
}

// Copies synchronise the device first: the library's stream may be a non-blocking one.
inline void device_vector::upload(const std::vector<double>& h) {
    if (h.size() != n_) throw std::domain_error("device_vector: size mismatch");
    if (hipDeviceSynchronize() != hipSuccess) throw nlf_error(NLF_ERR_HIP, "hipDeviceSynchronize failed");
    if (n_ && hipMemcpy(p_, h.data(), n_ * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        throw nlf_error(NLF_ERR_HIP, "hipMemcpy failed");
}

inline std::vector<double> device_vector::host() const {
    std::vector<double> h(n_);
    if (hipDeviceSynchronize() != hipSuccess) throw nlf_error(NLF_ERR_HIP, "hipDeviceSynchronize failed");
    if (n_ && hipMemcpy(h.data(), p_, n_ * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        throw nlf_error(NLF_ERR_HIP, "hipMemcpy failed");
    return h;
}

}  // namespace nonlocal
