// C++ mirror of NonLocFEM's 2D solver/assembler API (namespaces and class names of the reference)
// over the C ABI of libnlf.so (include/nlf.h).  Host code written against the reference's
// mesh_2d / find_neighbours / conductivity_matrix_2d / stiffness_matrix / conjugate_gradient keeps
// its structure; matrices and solver vectors live on the GPU behind the handles.
//
// Differences from the reference interface, all forced by the boundary:
//   * T = double, matrix index J = int64_t (as the reference's CLI and tests use);
//   * meshes are built from arrays (mesh_2d_data) instead of an SU2 path, so element numbering is
//     the caller's, not the SU2 reader's hash order (su2_parser.hpp:78-85);
//   * finite_element_matrix exports the reference's upper CSR (inner) and bound CSR on demand;
//   * Eigen vectors are std::vector<double> on the host side.
// Errors raise nonlocal::nlf_error (the reference throws std::domain_error / logic_error).
#pragma once

#include <nlf.h>

#include <array>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
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
    double* data() { return p_; }
    const double* data() const { return p_; }
    size_t size() const { return n_; }
    void upload(const std::vector<double>& h);
    std::vector<double> host() const;
};

enum class diameter_expanding_strategy { NO, MAX, MIN, MEAN };   // find_neighbours.hpp:7
enum class assemble_part { LOCAL = NLF_PART_LOCAL, FULL = NLF_PART_FULL };   // matrix_assembler_2d.hpp:15-19

namespace mesh {

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

template<class T = double>
struct influence {                                   // find_neighbours.hpp:18-22 (radius part)
    std::array<T, 2> radius = {0, 0};
};

template<class T = double>
using influences = std::unordered_map<std::string, influence<T>>;

// mesh_2d<T, I> (mesh_2d.hpp:18-81): geometry, quadrature precompute and neighbour lists on the GPU.
template<class T = double, class I = int64_t>
class mesh_2d {
    static_assert(std::is_same<T, double>::value, "the GPU path computes in double");
    nlf_mesh h_ = nullptr;
    context* ctx_;
    mesh_2d_data data_;

public:
    explicit mesh_2d(mesh_2d_data data, context& ctx = context::default_context()) : ctx_(&ctx), data_(std::move(data)) {
        const auto& d = data_;
        std::vector<double> xy(2 * d.nodes.size());
        for (size_t i = 0; i < d.nodes.size(); ++i) { xy[2 * i] = d.nodes[i][0]; xy[2 * i + 1] = d.nodes[i][1]; }
        const int64_t E = int64_t(d.elements.size()) / d.nodes_per_element;
        const int64_t B = d.nodes_per_edge ? int64_t(d.boundary_edges.size()) / d.nodes_per_edge : 0;
        nlf_check(nlf_mesh_create(ctx.handle(), int64_t(d.nodes.size()), xy.data(), E, d.nodes_per_element,
                                  d.elements.data(), int(d.group_begin.size()) - 1, d.group_begin.data(), B,
                                  d.nodes_per_edge, d.boundary_edges.data(), int(d.boundary_begin.size()) - 1,
                                  d.boundary_begin.data(), &h_),
                  ctx.handle());
    }
    ~mesh_2d() { if (h_) nlf_mesh_destroy(h_); }
    mesh_2d(const mesh_2d&) = delete;
    mesh_2d& operator=(const mesh_2d&) = delete;

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
    int boundary_index(const std::string& name) const {
        for (size_t g = 0; g < data_.boundary_names.size(); ++g)
            if (data_.boundary_names[g] == name) return int(g);
        throw std::domain_error("unknown boundary group " + name);
    }

    // mesh_2d::neighbours(e) (mesh_2d.hpp:70-72), copied back from the GPU
    std::vector<std::vector<I>> neighbours() const {
        int64_t total = 0;
        nlf_check(nlf_neighbours_count(h_, &total), ctx_->handle());
        std::vector<int64_t> off(elements_count() + 1);
        std::vector<int32_t> ids(size_t(std::max<int64_t>(total, 1)));
        nlf_check(nlf_neighbours_export(h_, off.data(), ids.data()), ctx_->handle());
        std::vector<std::vector<I>> out(elements_count());
        for (size_t e = 0; e < out.size(); ++e) out[e].assign(ids.begin() + off[e], ids.begin() + off[e + 1]);
        return out;
    }
};

// find_neighbours (find_neighbours.hpp:119-160) followed by mesh_2d::neighbours(neighbours_t&&)
// (mesh_2d.hpp:205-212): the lists stay on the GPU; returns the expanded influences.
template<class T, class I>
influences<T> find_neighbours(mesh_2d<T, I>& mesh, const influences<T>& radii,
                              diameter_expanding_strategy strategy = diameter_expanding_strategy::MAX) {
    const size_t G = mesh.groups_count();
    std::vector<double> r(2 * G, 0.0), expanded(2 * G, 0.0);
    for (size_t g = 0; g < G; ++g) {
        const auto it = radii.find(mesh.data().group_names[g]);
        if (it != radii.end()) { r[2 * g] = it->second.radius[0]; r[2 * g + 1] = it->second.radius[1]; }
    }
    nlf_check(nlf_find_neighbours(mesh.handle(), r.data(), int(strategy), expanded.data()), mesh.ctx().handle());
    influences<T> out;
    for (size_t g = 0; g < G; ++g)
        if (r[2 * g] > 0) out[mesh.data().group_names[g]] = influence<T>{{expanded[2 * g], expanded[2 * g + 1]}};
    return out;
}

}  // namespace mesh

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

// Parameters of one element group: model_parameters (equation_parameters.hpp:18-24) with the
// influence read by read_influence_2d (config/read_model.hpp:144-165).
struct model_parameters {
    nlf_model model{1.0, {0.0, 0.0}, {0.0, 0.0}, NLF_INFLUENCE_POLYNOMIAL, 2, 2.0, 1.0};
    static model_parameters bell(double local_weight, double radius, double q = 2.0) {
        model_parameters m;
        m.model = nlf_model{local_weight, {radius, radius}, {0.0, 0.0}, NLF_INFLUENCE_POLYNOMIAL, 2, 2.0, q};
        return m;
    }
};

namespace thermal {

struct parameter_2d {                                  // thermal_parameters_2d.hpp:14-17, 54-55 (constant)
    model_parameters model;
    nlf_thermal_material conductivity{NLF_CONDUCTIVITY_ISOTROPIC, 0, {1.0, 0.0, 0.0}};
    double capacity = 1.0;
    double density = 1.0;
};

using parameters_2d = std::unordered_map<std::string, parameter_2d>;   // by element group name

// conductivity_matrix_2d<T, I, J> (thermal/conductivity_matrix_2d.hpp:11-33)
template<class T = double, class I = int64_t, class J = int64_t>
class conductivity_matrix_2d {
    std::shared_ptr<mesh::mesh_2d<T, I>> mesh_;
    finite_element_matrix matrix_;

public:
    explicit conductivity_matrix_2d(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh) : mesh_(mesh) {}

    // conductivity_matrix_2d::compute (:166-193)
    void compute(const parameters_2d& parameters, const std::vector<bool>& is_inner, bool is_symmetric = true,
                 bool is_neumann = false, assemble_part part = assemble_part::FULL) {
        if (is_neumann) throw nlf_error(NLF_ERR_UNSUPPORTED, "Neumann (Lagrange-multiplier) problems are not implemented");
        const size_t G = mesh_->groups_count();
        std::vector<nlf_model> models(G);
        std::vector<nlf_thermal_material> mats(G);
        for (size_t g = 0; g < G; ++g) {
            const auto& p = parameters.at(mesh_->data().group_names[g]);
            models[g] = p.model.model;
            mats[g] = p.conductivity;
        }
        std::vector<uint8_t> inner(is_inner.begin(), is_inner.end());
        nlf_matrix h = nullptr;
        nlf_check(nlf_assemble(mesh_->handle(), 0, models.data(), mats.data(), inner.data(), int(is_symmetric), int(part), &h),
                  mesh_->ctx().handle());
        matrix_ = finite_element_matrix(h, mesh_->ctx().handle());
    }
    finite_element_matrix& matrix() { return matrix_; }
};

}  // namespace thermal

namespace mechanical {

// isotropic_elastic_parameters / orthotropic_elastic_parameters / anisotropic_elastic_parameters
// (elastic_parameters.hpp:62-133) as a plain material record; expansion_kind NLF_EXPANSION_NONE
inline nlf_elastic_material isotropic_elastic_parameters(double young_modulus, double poissons_ratio) {
    return {NLF_ELASTIC_ISOTROPIC, NLF_EXPANSION_NONE, young_modulus, poissons_ratio, 0.0, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0}};
}
inline nlf_elastic_material orthotropic_elastic_parameters(std::array<double, 2> young_modulus,
                                                           std::array<double, 2> poissons_ratio, double shear_modulus) {
    return {NLF_ELASTIC_ORTHOTROPIC, NLF_EXPANSION_NONE, young_modulus[0], poissons_ratio[0], young_modulus[1],
            poissons_ratio[1], shear_modulus, 0.0, {0.0, 0.0, 0.0}};
}
inline nlf_elastic_material anisotropic_elastic_parameters(std::array<double, 2> young_modulus,
                                                           std::array<double, 2> poissons_ratio, double shear_modulus,
                                                           double angle) {
    nlf_elastic_material m = orthotropic_elastic_parameters(young_modulus, poissons_ratio, shear_modulus);
    m.kind = NLF_ELASTIC_ANISOTROPIC;
    m.angle = angle;
    return m;
}

struct parameter_2d {
    model_parameters model;
    nlf_elastic_material material{NLF_ELASTIC_ISOTROPIC, NLF_EXPANSION_NONE, 1.0, 0.3, 0.0, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0}};
};

using parameters_2d = std::unordered_map<std::string, parameter_2d>;

// stiffness_matrix<T, I, J> (mechanical/stiffness_matrix_2d.hpp:10-38), plane stress, 2 DoF per node
template<class T = double, class I = int64_t, class J = int64_t>
class stiffness_matrix {
    std::shared_ptr<mesh::mesh_2d<T, I>> mesh_;
    finite_element_matrix matrix_;

public:
    explicit stiffness_matrix(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh) : mesh_(mesh) {}
    void compute(const parameters_2d& parameters, const std::vector<bool>& is_inner, bool is_symmetric = true,
                 assemble_part part = assemble_part::FULL) {
        const size_t G = mesh_->groups_count();
        std::vector<nlf_model> models(G);
        std::vector<nlf_elastic_material> mats(G);
        for (size_t g = 0; g < G; ++g) {
            const auto& p = parameters.at(mesh_->data().group_names[g]);
            models[g] = p.model.model;
            mats[g] = p.material;
        }
        std::vector<uint8_t> inner(is_inner.begin(), is_inner.end());
        nlf_matrix h = nullptr;
        nlf_check(nlf_assemble(mesh_->handle(), 1, models.data(), mats.data(), inner.data(), int(is_symmetric), int(part), &h),
                  mesh_->ctx().handle());
        matrix_ = finite_element_matrix(h, mesh_->ctx().handle());
    }
    finite_element_matrix& matrix() { return matrix_; }
};

}  // namespace mechanical

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

struct solve_statistics {
    uintmax_t iterations = 0;
    double residual = 0, true_residual = 0;
    int fallbacks = 0;
};

// conjugate_gradient<T, I> (slae/conjugate_gradient.hpp:15-42): keeps a non-owning reference to the
// matrix and owns its preconditioner.
class conjugate_gradient {
    const finite_element_matrix& matrix_;
    std::unique_ptr<preconditioner_base> precond_;
    double tol_ = 1e-15;
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
    double tolerance() const { return tol_; }
    uintmax_t max_iterations() const { return max_it_; }
    void set_tolerance(double tol) { tol_ = tol; }
    void set_max_interations(uintmax_t it) { max_it_ = it; }      // reference spelling (iterative_solver_base.hpp:49-51)
    uintmax_t iterations() const { return uintmax_t(stats_.iterations); }
    double residual() const { return stats_.residual; }
    const nlf_solve_stats& statistics() const { return stats_; }

    // solve(b, x0) (:48-79) with device-resident vectors
    void solve(const device_vector& b, device_vector& x, int mode = NLF_MODE_FP64) const {
        nlf_check(nlf_cg_solve(matrix_.handle(), precond_ ? precond_->handle() : nullptr, b.data(), x.data(), tol_,
                               int64_t(max_it_), mode, &stats_),
                  matrix_.ctx());
    }
    std::vector<double> solve(const std::vector<double>& b, const std::vector<double>* x0 = nullptr) const {
        device_vector db(b), dx(b.size());
        dx.upload(x0 ? *x0 : std::vector<double>(b.size(), 0.0));
        solve(db, dx);
        return dx.host();
    }
};

}  // namespace slae

// A boundary value as a function of the point, like the reference's first_kind_2d / second_kind_2d.
using point_function = std::function<double(const std::array<double, 2>&)>;

namespace detail {

// Boundary data of one condition per boundary group through nlf_boundary_data_fn: functions[b][d],
// when set, give the value of group b, degree of freedom d at a point.
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

}  // namespace detail

namespace thermal {

// Boundary condition of one boundary group (thermal_boundary_conditions_2d.hpp); the temperature
// and the flux may be functions of the point.
struct boundary_condition {
    nlf_boundary c{{NLF_BC_NONE, NLF_BC_NONE}, {0.0, 0.0}, 0.0, 0.0};
    point_function value;
    static boundary_condition temperature(double t) { return {{{NLF_BC_FIRST_KIND, NLF_BC_NONE}, {t, 0.0}, 0.0, 0.0}, {}}; }
    static boundary_condition temperature(point_function t) {
        return {{{NLF_BC_FIRST_KIND, NLF_BC_NONE}, {0.0, 0.0}, 0.0, 0.0}, std::move(t)};
    }
    static boundary_condition flux(double q) { return {{{NLF_BC_FLUX, NLF_BC_NONE}, {q, 0.0}, 0.0, 0.0}, {}}; }
    static boundary_condition flux(point_function q) {
        return {{{NLF_BC_FLUX, NLF_BC_NONE}, {0.0, 0.0}, 0.0, 0.0}, std::move(q)};
    }
    static boundary_condition convection(double h, double ambient) {
        return {{{NLF_BC_CONVECTION, NLF_BC_NONE}, {0.0, 0.0}, h, ambient}, {}};
    }
};

using boundaries_conditions_2d = std::unordered_map<std::string, boundary_condition>;

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
    double residual = 0, true_residual = 0;

    heat_equation_solution_2d(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh, const parameters_2d& parameters,
                              std::vector<T> temperature)
        : mesh_(mesh), temperature_(std::move(temperature)) {
        if (temperature_.size() < mesh_->nodes_count())
            throw nlf_error(NLF_ERR_ARGUMENT, "The temperature vector is shorter than the number of nodes.");
        temperature_.resize(mesh_->nodes_count());
        for (size_t g = 0; g < mesh_->groups_count(); ++g) {
            const auto& p = parameters.at(mesh_->data().group_names[g]);
            models_.push_back(p.model.model);
            materials_.push_back(p.conductivity);
        }
    }

    const std::vector<T>& temperature() const noexcept { return temperature_; }
    bool is_flux_calculated() const noexcept { return !flux_[0].empty() && !flux_[1].empty(); }
    const std::array<std::vector<T>, 2>& flux() const {
        if (!is_flux_calculated()) throw std::runtime_error{"Flux wasn't calculated"};
        return flux_;
    }
    // The neighbour lists of a nonlocal group must be on the mesh (find_neighbours), as in the reference.
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

// stationary_heat_equation_solver_2d (thermal/stationary_heat_equation_solver_2d.hpp:46-120):
// neighbours, assembly, convection, first-kind conditions, ILLT of the local matrix (identity on
// failure), CG.  The system is solved directly (the reference's correction form with T0 = 0 is
// the same system).
template<class T = double, class I = int64_t, class J = int64_t>
heat_equation_solution_2d<T, I> stationary_heat_equation_solver_2d(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh,
                                                                   const parameters_2d& parameters,
                                                                   const boundaries_conditions_2d& conditions,
                                                                   double tolerance = 1e-15,
                                                                   uintmax_t max_iterations = 10000) {
    const size_t G = mesh->groups_count(), GB = mesh->data().boundary_names.size();
    nlf_ctx ctx = mesh->ctx().handle();
    mesh::influences<T> radii;
    bool nonlocal = false;
    for (size_t g = 0; g < G; ++g) {
        const auto& m = parameters.at(mesh->data().group_names[g]).model.model;
        if (m.local_weight < 0.999) {
            nonlocal = true;
            radii[mesh->data().group_names[g]] = mesh::influence<T>{{m.search_radius[0] > 0 ? m.search_radius[0] : m.radius[0],
                                                                    m.search_radius[1] > 0 ? m.search_radius[1] : m.radius[1]}};
        }
    }
    if (nonlocal) mesh::find_neighbours(*mesh, radii);
    std::vector<nlf_boundary> bcs(GB);
    std::vector<std::array<point_function, 2>> functions(GB);
    for (size_t b = 0; b < GB; ++b) {
        const auto it = conditions.find(mesh->data().boundary_names[b]);
        bcs[b] = it == conditions.end() ? boundary_condition{}.c : it->second.c;
        if (it != conditions.end()) functions[b][0] = it->second.value;
    }
    const size_t rows = mesh->nodes_count();
    std::vector<double> f(rows), xD(rows);
    std::vector<uint8_t> inner8(rows);
    detail::boundary_data(mesh->handle(), ctx, 1, bcs, functions, f, xD, inner8);
    const std::vector<bool> is_inner(inner8.begin(), inner8.end());
    conductivity_matrix_2d<T, I, J> K(mesh);
    K.compute(parameters, is_inner);
    nlf_check(nlf_matrix_add_convection(K.matrix().handle(), bcs.data()), ctx);
    device_vector df(f), dxD(xD), dx(rows);
    nlf_check(nlf_apply_first_kind(K.matrix().handle(), dxD.data(), df.data()), ctx);
    slae::conjugate_gradient cg(K.matrix());
    cg.set_tolerance(tolerance);
    cg.set_max_interations(max_iterations);
    conductivity_matrix_2d<T, I, J> KL(mesh);
    KL.compute(parameters, is_inner, true, false, assemble_part::LOCAL);
    auto& P = cg.init_preconditioner<slae::eigen_ILLT_preconditioner>(KL.matrix());
    if (!P.computation_info()) cg.init_preconditioner<slae::eigen_identity_preconditioner>(K.matrix());
    dx.upload(std::vector<double>(rows, 0.0));
    cg.solve(df, dx);
    heat_equation_solution_2d<T, I> solution(mesh, parameters, dx.host());
    solution.iterations = cg.iterations();
    solution.residual = cg.residual();
    solution.true_residual = cg.statistics().true_residual;
    solution.calc_flux();                              // as the reference's solver does (:117-119)
    return solution;
}

// integrate_right_part (base/right_part_2d.hpp:10-35): f += int s N_i, s evaluated on the host at the
// quadrature points, integrated on the GPU.
template<class T, class I>
void integrate_right_part(device_vector& f, const mesh::mesh_2d<T, I>& mesh, const point_function& s) {
    int64_t N = 0, E = 0;
    int nn = 0, nq = 0;
    nlf_check(nlf_mesh_info(mesh.handle(), &N, &E, &nn, &nq), mesh.ctx().handle());
    std::vector<double> xq(size_t(2 * E * nq)), values(size_t(E * nq));
    nlf_check(nlf_mesh_export_quadrature(mesh.handle(), xq.data(), nullptr, nullptr, nullptr), mesh.ctx().handle());
    for (size_t k = 0; k < values.size(); ++k) values[k] = s({xq[2 * k], xq[2 * k + 1]});
    device_vector dv(values);
    nlf_check(nlf_right_part(mesh.handle(), 1, dv.data(), f.data()), mesh.ctx().handle());
}

// nonstationary_heat_equation_solver_2d (thermal/nonstationary_heat_equation_solver_2d.hpp:17-106):
// implicit Euler (C + tau K) T_{n+1} = C T_n + tau (f_boundary + f_source).  compute() assembles
// tau K + C on the GPU (nlf_matrix_combine); calc_step() forms the right-hand side, applies the
// first-kind conditions and runs CG (no preconditioner, as in the reference) from T_n.
template<class T = double, class I = int64_t, class J = int64_t>
class nonstationary_heat_equation_solver_2d {
    std::shared_ptr<mesh::mesh_2d<T, I>> mesh_;
    conductivity_matrix_2d<T, I, J> conductivity_;
    finite_element_matrix capacity_;
    std::unique_ptr<slae::conjugate_gradient> solver_;
    device_vector temperature_;
    std::vector<uint8_t> inner_;
    T time_step_ = 1;

    void boundary(const boundaries_conditions_2d& conditions, std::vector<double>& f, std::vector<double>& xD,
                  std::vector<uint8_t>& inner) const {
        const size_t GB = mesh_->data().boundary_names.size();
        std::vector<nlf_boundary> bcs(GB, boundary_condition{}.c);
        std::vector<std::array<point_function, 2>> functions(GB);
        for (size_t b = 0; b < GB; ++b) {
            const auto it = conditions.find(mesh_->data().boundary_names[b]);
            if (it == conditions.end()) continue;
            bcs[b] = it->second.c;
            functions[b][0] = it->second.value;
        }
        const size_t rows = mesh_->nodes_count();
        f.assign(rows, 0.0); xD.assign(rows, 0.0); inner.assign(rows, 1);
        detail::boundary_data(mesh_->handle(), mesh_->ctx().handle(), 1, bcs, functions, f, xD, inner);
    }

public:
    nonstationary_heat_equation_solver_2d(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh, T time_step)
        : mesh_(mesh), conductivity_(mesh), temperature_(mesh->nodes_count()), time_step_(time_step) {}

    std::vector<T> temperature() const { return temperature_.host(); }
    T time_step() const noexcept { return time_step_; }

    template<class Init_Dist>
    void compute(const parameters_2d& parameters, const boundaries_conditions_2d& conditions, const Init_Dist& init_dist) {
        const size_t G = mesh_->groups_count();
        mesh::influences<T> radii;
        bool nonlocal = false;
        std::vector<nlf_model> models(G);
        std::vector<double> capacities(G);
        for (size_t g = 0; g < G; ++g) {
            const auto& p = parameters.at(mesh_->data().group_names[g]);
            const auto& m = p.model.model;
            capacities[g] = p.density * p.capacity;
            models[g] = nlf_model{1.0, {0.0, 0.0}, {0.0, 0.0}, NLF_INFLUENCE_POLYNOMIAL, 2, 2.0, 1.0};
            if (m.local_weight < 0.999) {
                nonlocal = true;
                radii[mesh_->data().group_names[g]] = mesh::influence<T>{{m.search_radius[0] > 0 ? m.search_radius[0] : m.radius[0],
                                                                         m.search_radius[1] > 0 ? m.search_radius[1] : m.radius[1]}};
            }
        }
        if (nonlocal) mesh::find_neighbours(*mesh_, radii);
        std::vector<double> f, xD;
        boundary(conditions, f, xD, inner_);
        const std::vector<bool> is_inner(inner_.begin(), inner_.end());
        conductivity_.compute(parameters, is_inner);
        std::vector<nlf_boundary> bcs(mesh_->data().boundary_names.size(), boundary_condition{}.c);
        for (size_t b = 0; b < bcs.size(); ++b) {
            const auto it = conditions.find(mesh_->data().boundary_names[b]);
            if (it != conditions.end()) bcs[b] = it->second.c;
        }
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
        solver_ = std::make_unique<slae::conjugate_gradient>(conductivity_.matrix());
    }

    void calc_step(const boundaries_conditions_2d& conditions, const point_function& right_part = {}) {
        std::vector<double> f, xD;
        std::vector<uint8_t> inner;
        boundary(conditions, f, xD, inner);
        if (inner != inner_) throw nlf_error(NLF_ERR_ARGUMENT, "the first-kind boundary groups changed after compute()");
        device_vector df(f), dxD(xD), ct(mesh_->nodes_count());
        if (right_part) integrate_right_part(df, *mesh_, right_part);
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

}  // namespace thermal

namespace mechanical {

// Condition of one degree of freedom of a boundary group (mechanical_boundary_conditions_2d.hpp):
// displacement_2d (first kind) or pressure_2d (second kind), a constant or a function of the point.
struct boundary_condition {
    int kind = NLF_BC_NONE;
    double value = 0.0;
    point_function function;
    static boundary_condition displacement(double u) { return {NLF_BC_FIRST_KIND, u, {}}; }
    static boundary_condition displacement(point_function u) { return {NLF_BC_FIRST_KIND, 0.0, std::move(u)}; }
    static boundary_condition pressure(double p) { return {NLF_BC_FLUX, p, {}}; }
    static boundary_condition pressure(point_function p) { return {NLF_BC_FLUX, 0.0, std::move(p)}; }
};

using boundaries_conditions_2d = std::unordered_map<std::string, std::array<boundary_condition, 2>>;

// mechanical_solution_2d (mechanical/mechanical_solution_2d.hpp:11-232) without thermal strains:
// calc_strain_and_stress runs nlf_strain_stress on the GPU and keeps the node values.
template<class T = double, class I = int64_t>
class mechanical_solution_2d {
    std::shared_ptr<mesh::mesh_2d<T, I>> mesh_;
    std::vector<nlf_model> models_;
    std::vector<nlf_elastic_material> materials_;
    std::vector<T> solution_;                           // x, y per node, as the solve returns it
    std::array<std::vector<T>, 2> displacement_;
    std::array<std::vector<T>, 3> strain_, stress_;

public:
    uintmax_t iterations = 0;
    double residual = 0, true_residual = 0;

    mechanical_solution_2d(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh, const parameters_2d& parameters,
                           std::vector<T> displacement)
        : mesh_(mesh), solution_(std::move(displacement)) {
        const size_t n = mesh_->nodes_count();
        if (solution_.size() < 2 * n)
            throw nlf_error(NLF_ERR_ARGUMENT, "The displacement vector is shorter than twice the number of nodes.");
        solution_.resize(2 * n);
        for (size_t g = 0; g < mesh_->groups_count(); ++g) {
            const auto& p = parameters.at(mesh_->data().group_names[g]);
            models_.push_back(p.model.model);
            materials_.push_back(p.material);
        }
        for (auto& d : displacement_) d.resize(n);
        for (size_t i = 0; i < n; ++i) {
            displacement_[0][i] = solution_[2 * i];
            displacement_[1][i] = solution_[2 * i + 1];
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
    // The neighbour lists of a nonlocal group must be on the mesh (find_neighbours), as in the reference.
    void calc_strain_and_stress() {
        if (is_strain_and_stress_calculated()) return;
        const size_t n = mesh_->nodes_count();
        device_vector du(solution_), dstrain(3 * n), dstress(3 * n);
        nlf_check(nlf_strain_stress(mesh_->handle(), models_.data(), materials_.data(), du.data(), nullptr,
                                    dstrain.data(), dstress.data(), nullptr),
                  mesh_->ctx().handle());
        const std::vector<double> hs = dstrain.host(), ht = dstress.host();
        for (size_t k = 0; k < 3; ++k) {
            strain_[k].assign(hs.begin() + std::ptrdiff_t(k * n), hs.begin() + std::ptrdiff_t((k + 1) * n));
            stress_[k].assign(ht.begin() + std::ptrdiff_t(k * n), ht.begin() + std::ptrdiff_t((k + 1) * n));
        }
    }
};

// equilibrium_equation (mechanical/equilibrium_equation_2d.hpp:20-62): neighbours, stiffness,
// boundary conditions, CG with ILLT of the local stiffness when nonlocal (identity otherwise, and
// on a failed factorisation), then the strains and stresses.
template<class T = double, class I = int64_t, class J = int64_t>
mechanical_solution_2d<T, I> equilibrium_equation(const std::shared_ptr<mesh::mesh_2d<T, I>>& mesh,
                                                  const parameters_2d& parameters,
                                                  const boundaries_conditions_2d& conditions,
                                                  double tolerance = 1e-15, uintmax_t max_iterations = 10000) {
    const size_t G = mesh->groups_count(), GB = mesh->data().boundary_names.size();
    nlf_ctx ctx = mesh->ctx().handle();
    mesh::influences<T> radii;
    bool nonlocal = false;
    for (size_t g = 0; g < G; ++g) {
        const auto& m = parameters.at(mesh->data().group_names[g]).model.model;
        if (m.local_weight < 0.999) {
            nonlocal = true;
            radii[mesh->data().group_names[g]] = mesh::influence<T>{{m.search_radius[0] > 0 ? m.search_radius[0] : m.radius[0],
                                                                    m.search_radius[1] > 0 ? m.search_radius[1] : m.radius[1]}};
        }
    }
    if (nonlocal) mesh::find_neighbours(*mesh, radii);
    std::vector<nlf_boundary> bcs(GB, nlf_boundary{{NLF_BC_NONE, NLF_BC_NONE}, {0.0, 0.0}, 0.0, 0.0});
    std::vector<std::array<point_function, 2>> functions(GB);
    for (size_t b = 0; b < GB; ++b) {
        const auto it = conditions.find(mesh->data().boundary_names[b]);
        if (it == conditions.end()) continue;
        for (int d = 0; d < 2; ++d) {
            bcs[b].kind[d] = it->second[size_t(d)].kind;
            bcs[b].value[d] = it->second[size_t(d)].value;
            functions[b][size_t(d)] = it->second[size_t(d)].function;
        }
    }
    const size_t rows = 2 * mesh->nodes_count();
    std::vector<double> f(rows), xD(rows);
    std::vector<uint8_t> inner8(rows);
    detail::boundary_data(mesh->handle(), ctx, 2, bcs, functions, f, xD, inner8);
    const std::vector<bool> is_inner(inner8.begin(), inner8.end());
    stiffness_matrix<T, I, J> K(mesh);
    K.compute(parameters, is_inner);
    device_vector df(f), dxD(xD), dx(rows);
    nlf_check(nlf_apply_first_kind(K.matrix().handle(), dxD.data(), df.data()), ctx);
    slae::conjugate_gradient cg(K.matrix());
    cg.set_tolerance(tolerance);
    cg.set_max_interations(max_iterations);
    stiffness_matrix<T, I, J> KL(mesh);
    if (nonlocal) {
        KL.compute(parameters, is_inner, true, assemble_part::LOCAL);
        auto& P = cg.init_preconditioner<slae::eigen_ILLT_preconditioner>(KL.matrix());
        if (!P.computation_info()) cg.init_preconditioner<slae::eigen_identity_preconditioner>(K.matrix());
    } else {
        cg.init_preconditioner<slae::eigen_identity_preconditioner>(K.matrix());
    }
    dx.upload(std::vector<double>(rows, 0.0));
    cg.solve(df, dx);
    mechanical_solution_2d<T, I> solution(mesh, parameters, dx.host());
    solution.iterations = cg.iterations();
    solution.residual = cg.residual();
    solution.true_residual = cg.statistics().true_residual;
    solution.calc_strain_and_stress();
    return solution;
}

}  // namespace mechanical

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
