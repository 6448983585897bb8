/*
 * nlf.h -- C ABI of the MI355X-native 2D nonlocal finite-element hot path.
 *
 * Every entry point returns an int status (NLF_OK = 0, negative = error) and never throws.  The
 * message of the last failure on a context is available from nlf_last_error().  All device work
 * is ordered on the context's HIP stream; "host" pointers are plain host memory owned by the
 * caller, "dev" pointers are device memory (e.g. from torch or hipMalloc).
 *
 * Each function cites the reference interface it replaces (paths relative to the NonLocFEM
 * repository, VIXER13/nonlocal).
 */
#ifndef NLF_H
#define NLF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------- */
#define NLF_OK               0
#define NLF_ERR_ARGUMENT    -1   /* invalid argument (std::domain_error / logic_error in the reference) */
#define NLF_ERR_HIP         -2   /* HIP runtime failure */
#define NLF_ERR_STATE       -3   /* call out of order (e.g. assembly before neighbours) */
#define NLF_ERR_UNSUPPORTED -4   /* combination not implemented */
#define NLF_ERR_NUMERIC     -5   /* numerical breakdown (reported, not thrown, like computation_info) */

/* ---- element and model types -------------------------------------------------------------- */
#define NLF_ELEMENT_Q4 4         /* bilinear quad, Gauss 2x2 (vtk_elements_set.hpp:62-64) */
#define NLF_ELEMENT_Q8 8         /* serendipity quad p=2/9, Gauss 3x3 (vtk_elements_set.hpp:65-67) */

#define NLF_INFLUENCE_POLYNOMIAL  0   /* influence::polynomial / fast_polynomial (influence_functions_2d.hpp:61-92,127-148) */
#define NLF_INFLUENCE_CONSTANT    1   /* influence::constant (influence_functions_2d.hpp:39-59) */
#define NLF_INFLUENCE_EXPONENTIAL 2   /* influence::exponential (influence_functions_2d.hpp:94-125) */

/* Distance of an influence (read_distance_2d, config/read_model.hpp:120-131) */
#define NLF_DISTANCE_LP                    0   /* powered_distance (mesh_2d/search_function.hpp:28-73) */
#define NLF_DISTANCE_ELLIPSE_WITH_ROTATION 1   /* powered_distance_with_rotation (search_function.hpp:75-90) */
#define NLF_DISTANCE_N_INFINITY           -1   /* distance_n of the maximum norm (metamath Infinity<size_t>) */

#define NLF_EXPAND_NO   0            /* diameter_expanding_strategy (find_neighbours.hpp:7) */
#define NLF_EXPAND_MAX  1
#define NLF_EXPAND_MIN  2
#define NLF_EXPAND_MEAN 3

#define NLF_PART_LOCAL 0             /* assemble_part (matrix_assembler_2d.hpp:15-19) */
#define NLF_PART_FULL  2

#define NLF_CONDUCTIVITY_ISOTROPIC   0   /* thermal_parameters_2d.hpp:14-17 */
#define NLF_CONDUCTIVITY_ORTHOTROPIC 1
#define NLF_CONDUCTIVITY_ANISOTROPIC 2

#define NLF_BC_NONE        0         /* insulated: no condition */
#define NLF_BC_FIRST_KIND  1         /* temperature_2d / displacement */
#define NLF_BC_FLUX        2         /* flux_2d / pressure_2d (second kind; positive = heat entering) */
#define NLF_BC_CONVECTION  3         /* convection_2d: h, T_ambient (convection_condition_2d.hpp:11-31) */

#define NLF_PRECOND_IDENTITY 0       /* eigen_identity_preconditioner */
#define NLF_PRECOND_JACOBI   1       /* eigen_diagonal_preconditioner */
#define NLF_PRECOND_ILLT     2       /* eigen_ILLT_preconditioner: here IC(0), scaled, shift retries */
#define NLF_PRECOND_LEARNED  3       /* message-passing network, flexible CG only */

#define NLF_MODE_FP64  0             /* plain FP64 CG */
#define NLF_MODE_MIXED 1             /* FP32 inner CG + FP64 iterative refinement */
#define NLF_MODE_SPLIT 2             /* values read as FP32 + 16-bit remainder (40 significand bits),
                                        FP64 vectors, FP64 true residual with reliable updates */

/* Nonlocal model of one element group: model_parameters (equation_parameters.hpp:18-24) plus the
 * influence description read by read_influence_2d (config/read_model.hpp:144-165). */
typedef struct nlf_model {
    double  local_weight;      /* p1; LOCAL theory when >= 0.999 (nonlocal_constants.hpp:21-27) */
    double  radius[2];         /* nonlocal radius of the influence function */
    double  search_radius[2];  /* neighbour-search radius before expansion; 0 => use radius */
    int32_t influence;         /* NLF_INFLUENCE_* */
    int32_t distance_n;        /* integer Lp exponent n of the distance (size_t_or<T> holding size_t):
                                  0 (= 2), 1, 2, 3, ... or NLF_DISTANCE_N_INFINITY */
    double  p, q;              /* influence parameters (polynomial: p, q; exponential: p, q) */
    int32_t distance;          /* NLF_DISTANCE_*; ELLIPSE_WITH_ROTATION takes n = 2 */
    int32_t pad_;
    double  distance_real;     /* > 0: a real Lp exponent n (size_t_or<T> holding T); distance_n is then
                                  ignored.  0: distance_n decides */
} nlf_model;

/* Distance of one element group's neighbour search: the distance member of mesh::influence
 * (search_function.hpp:16-20); kind, n and n_real as the distance fields of nlf_model. */
typedef struct nlf_distance {
    int32_t kind;              /* NLF_DISTANCE_* */
    int32_t n;                 /* integer exponent, 0 (= 2) or NLF_DISTANCE_N_INFINITY */
    double  n_real;            /* > 0: real exponent */
} nlf_distance;

typedef struct nlf_thermal_material {
    int32_t kind;              /* NLF_CONDUCTIVITY_* */
    int32_t pad_;
    double  c[3];              /* iso: c[0]; ortho: {l11, l22}; aniso: {l11, l22, l12} */
} nlf_thermal_material;

#define NLF_ELASTIC_ISOTROPIC   0      /* isotropic_elastic_parameters (elastic_parameters.hpp:62-76) */
#define NLF_ELASTIC_ORTHOTROPIC 1      /* orthotropic_elastic_parameters (elastic_parameters.hpp:78-96) */
#define NLF_ELASTIC_ANISOTROPIC 2      /* anisotropic_elastic_parameters: orthotropic rotated by an angle (:98-133) */

#define NLF_EXPANSION_NONE        0    /* no thermal expansion: the group takes no thermal load */
#define NLF_EXPANSION_ISOTROPIC   1    /* thermal_expansion_parameters.hpp:7-18: alpha */
#define NLF_EXPANSION_ORTHOTROPIC 2    /* {a_xx, a_yy} */
#define NLF_EXPANSION_ANISOTROPIC 3    /* {a_xx, a_yy, a_xy} */

/* Plane-stress material of one element group (raw_mechanical_parameters_t, mechanical_parameters_2d.hpp:8-12).
 * Isotropic: E = young, nu = poisson.  Orthotropic: E = {young, young_y}, nu = {poisson, poisson_y},
 * G = shear.  Anisotropic: the orthotropic Hooke matrix rotated by `angle` (radians). */
typedef struct nlf_elastic_material {
    int32_t kind;              /* NLF_ELASTIC_* */
    int32_t expansion_kind;    /* NLF_EXPANSION_* */
    double  young, poisson;
    double  young_y, poisson_y, shear, angle;
    double  expansion[3];      /* thermal expansion coefficients of expansion_kind */
} nlf_elastic_material;

/* One boundary-group condition; for elasticity `value` holds the two components. */
typedef struct nlf_boundary {
    int32_t kind[2];           /* NLF_BC_* per degree of freedom (thermal uses kind[0]) */
    double  value[2];          /* temperature / flux / displacement / traction */
    double  heat_transfer;     /* convection h */
    double  ambient;           /* convection T_ambient */
} nlf_boundary;

typedef struct nlf_solve_stats {
    int64_t iterations;        /* total Krylov iterations (inner ones in mixed mode) */
    int64_t outer_iterations;  /* refinement steps (mixed mode) */
    double  residual;          /* reference criterion sqrt(r'M^-1 r)/||b|| of the last (inner) solve */
    double  true_residual;     /* FP64 ||b - A x|| / ||b|| */
    int32_t preconditioner;    /* preconditioner actually used after fallbacks */
    int32_t fallbacks;         /* number of switches away from the requested preconditioner */
    double  seconds;           /* wall time of the solve on the stream */
    double  spmv_ms;           /* average SpMV kernel time (HIP events on the stream, sampled launches) */
    double  precond_ms;        /* average preconditioner-apply time (sampled) */
    int64_t spmv_samples;      /* number of sampled SpMV launches */
    int64_t spmv_bytes;        /* bytes the SpMV kernel moves per launch (values + indices + vectors) */
} nlf_solve_stats;

typedef struct nlf_ctx_s*     nlf_ctx;
typedef struct nlf_mesh_s*    nlf_mesh;
typedef struct nlf_matrix_s*  nlf_matrix;
typedef struct nlf_precond_s* nlf_precond;
typedef struct nlf_batch_s*   nlf_batch;
typedef struct nlf_trainer_s* nlf_trainer;

/* ---- context ---------------------------------------------------------------------------------
 * No reference counterpart (the reference is CPU-only); one context per host thread, `stream` may
 * be NULL (the library creates a non-blocking stream) or an existing hipStream_t. */
int         nlf_ctx_create(int device, void* stream, nlf_ctx* out);
int         nlf_ctx_destroy(nlf_ctx ctx);
const char* nlf_last_error(nlf_ctx ctx);
int         nlf_ctx_synchronize(nlf_ctx ctx);
#define NLF_OPT_ASSEMBLY_PAIRS 1     /* 1: MFMA element-pair contractions (thermal Q4), 0: node-centric */
#define NLF_OPT_ROW_PATTERNS   2     /* 1: SpMV column offsets from a dictionary of distinct rows, 0: per nonzero */
#define NLF_OPT_ASSEMBLY_THREADS 3   /* threads per node workgroup of the value pass: 64 (default), 128, 256, 512 */
#define NLF_OPT_PATTERN_THREADS  4   /* threads per node workgroup of the pattern build: 64, 128 (default), 256 */
#define NLF_OPT_PAIR_BUDGET_MIB  5   /* MiB for the MFMA pair blocks of one assembly pass; 0 (default): 8 GiB;
                                        a mesh whose blocks do not fit is assembled in node chunks */
#define NLF_OPT_SPMV_SYMMETRIC   6   /* 1 (default): the FP64 products of CG, flexible CG and nlf_spmv read the upper
                                        triangle once (the reference's symmetric product), from a slice layout built
                                        on the first solve; 0: the full (both-triangle) storage */
int         nlf_ctx_set_option(nlf_ctx ctx, int option, int value);
void*       nlf_ctx_stream(nlf_ctx ctx);

/* ---- mesh: replaces mesh_2d<T,I>(path) + mesh_2d::init (libraries/mesh/mesh_2d/mesh_2d.hpp:83-108).
 * Node ids 0..N-1, elements E x nn (nn = 4 or 8, reference local order), contiguous element groups
 * group_begin[G+1], boundary edges n_bedges x nb grouped by bgroup_begin[GB+1]. */
int nlf_mesh_create(nlf_ctx ctx, int64_t n_nodes, const double* xy, int64_t n_elems, int nn, const int32_t* elems,
                    int n_groups, const int64_t* group_begin, int64_t n_bedges, int nb, const int32_t* bedges,
                    int n_bgroups, const int64_t* bgroup_begin, nlf_mesh* out);
int nlf_mesh_destroy(nlf_mesh m);
int nlf_mesh_info(nlf_mesh m, int64_t* n_nodes, int64_t* n_elems, int* nn, int* nq);
/* quad coords [E][nq][2], |det J| [E][nq], det J * grad N [E][nn][nq][2], centres [E][2]
 * (mesh_2d::quad_coord / jacobian / derivatives, mesh_2d.hpp:141-179; centres mesh_container_2d.hpp:130-139) */
int nlf_mesh_export_quadrature(nlf_mesh m, double* xq, double* jac, double* D, double* centres);
/* the element's quadrature weights [nq] and shape values N_i(q) [nn][nq] (element_2d::weight / qN,
 * as mesh::utils::integrate uses them, mesh_2d_utils.hpp:135-149) */
int nlf_mesh_export_basis(nlf_mesh m, double* weights, double* qN);

/* ---- neighbours: replaces find_neighbours (libraries/mesh/mesh_2d/find_neighbours.hpp:119-160) and
 * mesh_2d::neighbours (mesh_2d.hpp:205-212).  search_radii[2G]; <= 0 skips the group. */
int nlf_find_neighbours(nlf_mesh m, const double* search_radii, int strategy, double* expanded_out);
/* the same search with a distance per group (distances[G]; NULL = Euclidean everywhere): the predicate
 * distance(c_eL, c_eNL, r) <= 1 of find_neighbours.hpp:152-155 with each group's distance function.
 * The radius expansion stays Euclidean (diameter_between_elements_centers, :49-72). */
int nlf_find_neighbours_ex(nlf_mesh m, const double* search_radii, const nlf_distance* distances, int strategy,
                           double* expanded_out);
int nlf_neighbours_count(nlf_mesh m, int64_t* total);
int nlf_neighbours_export(nlf_mesh m, int64_t* offsets, int32_t* ids);
int nlf_neighbours_device(nlf_mesh m, const int64_t** offsets, const int32_t** ids);

/* ---- assembly: replaces conductivity_matrix_2d::compute (thermal/conductivity_matrix_2d.hpp:166-193)
 * and stiffness_matrix::compute (mechanical/stiffness_matrix_2d.hpp:171-199); physics 0 / 1,
 * materials = nlf_thermal_material[G] or nlf_elastic_material[G], is_inner[DoF*N] host mask. */
int nlf_assemble(nlf_mesh m, int physics, const nlf_model* models, const void* materials, const uint8_t* is_inner,
                 int symmetric, int part, nlf_matrix* out);
/* Options of conductivity_matrix_2d::compute / stiffness_matrix::compute beyond nlf_assemble's.
 * neumann: is_neumann of conductivity_matrix_2d::compute (conductivity_matrix_2d.hpp:39-72, 83-104):
 *   thermal only and no first-kind row; the matrix gets the Lagrange row and column N (rows = N + 1):
 *   K(r, N) = int N_r, K(N, N) = 0.  The right-hand side then has N + 1 entries, f[N] = the energy
 *   (stationary_heat_equation_solver_2d.hpp:28-44), and CG runs on the indefinite bordered system.
 * coefficients: NULL, or per element group NULL (the material's constant) or that group's values at
 *   its quadrature points (evaluated_parameters' Variable alternative; evaluate_conductivity.hpp:16-37,
 *   evaluate_mechanical_parameters.hpp:9-30): index ((e - group_begin) * nq + q) * ncomp + k, ncomp
 *   1 / 2 / 3 for isotropic / orthotropic / anisotropic conductivity {l11, l22, l12}, 3 / 4 / 6 for the
 *   Hooke matrix of an isotropic / orthotropic / anisotropic material in the reference's index order
 *   ({11, 12, 66}, {11, 12, 66, 22}, {11, 12, 66, 22, 16, 26}).  Only for problems whose groups are all
 *   local: varying coefficients make a nonlocal problem asymmetric (problem_settings.hpp:32-34), which
 *   the reference solves with BiCGSTAB (NLF_ERR_UNSUPPORTED here). */
typedef struct nlf_assemble_options {
    int32_t neumann;
    int32_t pad_;
    const double* const* coefficients;
} nlf_assemble_options;
int nlf_assemble_ex(nlf_mesh m, int physics, const nlf_model* models, const void* materials, const uint8_t* is_inner,
                    int symmetric, int part, const nlf_assemble_options* options, nlf_matrix* out);
int nlf_matrix_destroy(nlf_matrix a);
int nlf_matrix_info(nlf_matrix a, int64_t* rows, int64_t* nnz_upper, int64_t* nnz_full, int64_t* nnz_bound);
int nlf_matrix_timings(nlf_matrix a, double* pattern_seconds, double* values_seconds);
/* finite_element_matrix::inner()/bound() export (base/finite_element_matrix.hpp:12-21), int64 indices:
 * part 0 = inner upper CSR, 1 = bound CSR, 2 = inner full storage; 3 and 4 = parts 0 and 2 of the
 * FP32 values (nlf_matrix_make_fp32 / nlf_matrix_assemble_fp32), widened to double */
int nlf_matrix_export(nlf_matrix a, int part, int64_t* outer, int64_t* inner, double* values);
/* The inner upper CSR of rows [row0, row1) only (finite_element_matrix::inner() rows, as
 * SparseMatrix::middleRows would give them): outer[row1 - row0 + 1] from 0, inner/values may be NULL. */
int nlf_matrix_export_rows(nlf_matrix a, int64_t row0, int64_t row1, int64_t* outer, int64_t* inner, double* values);
int nlf_matrix_device(nlf_matrix a, const int64_t** outer, const int32_t** cols, double** values, const int64_t** diag);
/* No reference counterpart.  Symmetric SpMV storage (NLF_OPT_SPMV_SYMMETRIC): builds it when it applies and
 * reports its size: slots (64 values each: the upper triangle padded to each 64-row slice's union of column
 * offsets) and strip doubles (the transposed sums each product writes and reads back).  NLF_ERR_STATE when it
 * does not apply (the full storage serves the products). */
int nlf_matrix_symv_info(nlf_matrix a, int64_t* slots, int64_t* strips);
/* No reference counterpart.  Row-pattern dictionary of the SpMV (NLF_OPT_ROW_PATTERNS): number of
 * distinct offset rows and their total length (both 0 when the per-nonzero offsets are used). */
int nlf_matrix_pattern_info(nlf_matrix a, int64_t* groups, int64_t* entries);
int nlf_matrix_add_pairs(nlf_matrix a, int64_t n, const int64_t* rc, const double* v);
/* convection_condition_2d (thermal/convection_condition_2d.hpp:11-31), bcs[GB] */
int nlf_matrix_add_convection(nlf_matrix a, const nlf_boundary* bcs);
/* No reference counterpart (the reference assembles in T = double only).  FP32 values for the
 * mixed-precision inner solves: make_fp32 rounds the FP64 values; assemble_fp32 repeats the value
 * pass of nlf_assemble in FP32 (FP32 element-pair contractions, MFMA for thermal Q4, and FP32
 * accumulation; Jacobian determinants, coordinates and the cutoff decisions stay FP64) and then
 * the boundary additions made since.  The FP64 values are kept for the true residual. */
int nlf_matrix_make_fp32(nlf_matrix a);
int nlf_matrix_assemble_fp32(nlf_matrix a);
/* No reference counterpart.  Values for NLF_MODE_SPLIT: val32 = FP32 rounding of the FP64 values
 * and the remainder as a 16-bit integer in units of 2^-16 ulp(val32), so that the pair carries 40
 * significand bits (relative error <= 2^-40) in 6 bytes per nonzero instead of 8. */
int nlf_matrix_make_split(nlf_matrix a);
/* No reference counterpart.  Distance of the FP32 values from the FP64 ones, computed on the
 * device over the full storage: out[4] = {max |v32 - v64|, max |v64|, ||v32 - v64||_2, ||v64||_2}. */
int nlf_matrix_compare_fp32(nlf_matrix a, double* out);

/* ---- nonstationary heat conduction and source terms.
 * nlf_assemble with physics 2 assembles the heat capacity matrix C_ij = rho c int N_i N_j
 * (heat_capacity_matrix_2d.hpp:31-62; `materials` = double[G] of density * capacity, local theory).
 * nlf_matrix_combine: K := alpha K + beta C on K's pattern, K's bound part times alpha, Dirichlet rows
 * back to identity rows -- the tau K + C of nonstationary_heat_equation_solver_2d::compute
 * (thermal/nonstationary_heat_equation_solver_2d.hpp:66-87); C's pattern must lie within K's.
 * nlf_right_part: f_dev[dof*N] += int s N_i (integrate_right_part, base/right_part_2d.hpp:10-35) with
 * the source values at the quadrature points, values_dev[(e*nq + q)*dof + d] (device arrays). */
int nlf_matrix_combine(nlf_matrix K, double alpha, nlf_matrix C, double beta);
int nlf_right_part(nlf_mesh m, int dof, const double* values_dev, double* f_dev);

/* ---- heat flux: replaces heat_equation_solution_2d::calc_flux
 * (solvers/solver_2d/thermal/heat_equation_solution_2d.hpp:99-137) with gradient_in_qnodes and
 * qnodes_to_nodes (mesh/mesh_2d/mesh_2d_utils.hpp:39-63,82-101).  T_dev[N] node temperatures;
 * flux_dev[2N] node fluxes (x block, then y block, like the reference's array of two vectors);
 * qflux_dev[2*E*nq] quadrature-point fluxes (x block, then y block; index e*nq + q).  Either output
 * may be NULL.  A nonlocal group needs nlf_find_neighbours first (NLF_ERR_STATE otherwise). */
int nlf_heat_flux(nlf_mesh m, const nlf_model* models, const nlf_thermal_material* materials, const double* T_dev,
                  double* flux_dev, double* qflux_dev);
/* ---- strains and stresses: replaces mechanical_solution_2d::calc_strain_and_stress
 * (solvers/solver_2d/mechanical/mechanical_solution_2d.hpp:192-232).  u_dev[2N] the displacement as the
 * solve returns it (x, y per node); delta_temperature_dev[N] the node temperature change or NULL: its
 * thermal strain (alpha dT of each group's expansion) is subtracted before the stresses
 * (substract_temperature_strains :135-166); strain_dev[3N] and stress_dev[3N] node values (xx, yy, xy
 * blocks; xy strain is the tensor shear (u_x,y + u_y,x)/2, xy stress 2 h66 eps_xy + h16 eps_xx + h26 eps_yy);
 * qstress_dev[3*E*nq] quadrature-point stresses after the nonlocal average.  Any output may be NULL.
 * A nonlocal group needs nlf_find_neighbours first (NLF_ERR_STATE otherwise). */
int nlf_strain_stress(nlf_mesh m, const nlf_model* models, const nlf_elastic_material* materials, const double* u_dev,
                      const double* delta_temperature_dev, double* strain_dev, double* stress_dev, double* qstress_dev);
/* ---- thermal load: replaces temperature_condition (solvers/solver_2d/mechanical/temperature_condition_2d.hpp:
 * 128-170).  f_dev[2N] += the load of the node temperature change delta_temperature_dev[N] for the groups
 * with a thermal expansion: p1 times the local integral of grad N_i . C alpha dT plus (1 - p1) times the
 * nonlocal one over the neighbour lists, in the reference's form. */
int nlf_temperature_condition(nlf_mesh m, const nlf_model* models, const nlf_elastic_material* materials,
                              const double* delta_temperature_dev, double* f_dev);

/* inner_nodes (base/solvers_utils.hpp:31-40), boundary_condition_second_kind_2d
 * (base/boundary_condition_second_kind_2d.hpp:12-33) and the first-kind values (:17-26), host arrays */
int nlf_boundary_data(nlf_mesh m, int dof, const nlf_boundary* bcs, double* f, double* xD, uint8_t* is_inner);
/* The same with position-dependent values, as the reference's conditions are functions of the point
 * (first_kind_2d / second_kind_2d, base/boundary_conditions_2d.hpp): value(x, y, group, dof, user) is
 * called on the host at each node of a first-kind group and at each Gauss point of the edges of a
 * second-kind (NLF_BC_FLUX) group, in place of bcs[group].value[dof].  Convection stays constant. */
typedef double (*nlf_boundary_value_fn)(double x, double y, int group, int dof, void* user);
int nlf_boundary_data_fn(nlf_mesh m, int dof, const nlf_boundary* bcs, nlf_boundary_value_fn value, void* user,
                         double* f, double* xD, uint8_t* is_inner);
/* boundary_condition_first_kind_2d (base/boundary_condition_first_kind_2d.hpp:50-59): f -= K_bound xD; f[D] = xD */
int nlf_apply_first_kind(nlf_matrix a, const double* xD_dev, double* f_dev);

/* ---- SpMV: replaces independent_symmetric_matrix_vector_product::matrix_vector_product
 * (slae/independent_symmetric_matrix_vector_product.hpp:55-79); device vectors */
int nlf_spmv(nlf_matrix a, const double* x_dev, double* y_dev);
int nlf_spmv_fp32(nlf_matrix a, const float* x_dev, float* y_dev);
int nlf_spmv_lpr(nlf_matrix a, int lpr, int fp32, const void* x_dev, void* y_dev);

/* ---- preconditioners: replace preconditioner_base (slae/preconditioner_base.hpp:7-15) and
 * eigen_{identity,diagonal,ILLT}_preconditioner (slae/eigen_preconditioner.hpp:50-57) */
int nlf_precond_create(nlf_matrix a, int kind, nlf_precond* out);
int nlf_precond_destroy(nlf_precond p);
int nlf_precond_info(nlf_precond p, int* ok, double* shift, int* fw_levels, int* bw_levels, double* setup_seconds);
/* No reference counterpart: the layout of the ILLT sweeps, for diagnostics.  Per sweep (forward,
 * then backward): out[0/4] levels as swept (wide levels split; for a block-diagonal matrix, whose
 * independent blocks are factored and swept side by side, the most of any block), out[1/5] compute width,
 * out[2/6] levels held in the LDS dependency window (0: none), out[3/7] kernel (0 packed,
 * 1 ELL one level at a time, 2 ELL with prefetch). */
int nlf_precond_sweep_info(nlf_precond p, int32_t out[8]);
/* No reference counterpart: the independent diagonal blocks an ILLT found in its matrix (contiguous
 * row ranges with no entry between them, such as a batch's instances, of at least 1024 rows each),
 * each factored and swept by a workgroup of its own, side by side; 1 for a connected matrix. */
int nlf_precond_blocks(nlf_precond p, int32_t* blocks);
int nlf_precond_apply(nlf_precond p, const double* r_dev, double* z_dev);
int nlf_precond_export_factor(nlf_precond p, double* values, double* scale);

/* ---- CG: replaces conjugate_gradient::solve (slae/conjugate_gradient.hpp:48-79); device vectors,
 * x holds the initial guess */
int nlf_cg_solve(nlf_matrix a, nlf_precond p, const double* b_dev, double* x_dev, double tol, int64_t max_it, int mode,
                 nlf_solve_stats* stats);

/* No reference counterpart.  For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): nlf_cg_solve in NLF_MODE_FP64 whose
 * work vectors, partial sums and solver state live in a fixed device array of the library, sized for
 * at most NLF_NOALLOC_MAX_ROWS rows (about 48 MiB, held per GPU until the process ends).  It calls no
 * hipMalloc or hipFree and records no HIP events (stats carry no kernel timings).  A larger matrix is
 * an argument error.  Every call on one GPU shares that array: calls that overlap on one GPU race.
 * Its products read the full storage (the symmetric slice storage, NLF_OPT_SPMV_SYMMETRIC, is memory it
 * does not allocate), so it matches nlf_cg_solve bitwise with that option at 0. */
#define NLF_NOALLOC_MAX_ROWS 2097152
int nlf_cg_solve_noalloc(nlf_matrix a, nlf_precond p, const double* b_dev, double* x_dev, double tol, int64_t max_it,
                         nlf_solve_stats* stats);

/* ---- learned preconditioner (no reference counterpart: an additional preconditioner_base
 * implementation, used first with ILLT / Jacobi as the fallback).  A message-passing network (3
 * layers, width 32) over the graph of the local matrix; weights are nlf_learned_num_params() floats. */
int nlf_learned_num_params(void);
int nlf_learned_init_weights(uint64_t seed, float* w_host);
int nlf_learned_create(nlf_matrix a, nlf_matrix a_local, const nlf_model* models, nlf_precond* out);
int nlf_learned_set_weights(nlf_precond p, const float* w_host);
int nlf_learned_get_weights(nlf_precond p, float* w_host);
/* the network is linear in r: set_weights compiles its symmetric part into a sparse matrix on the
 * three-hop pattern (applied as one SpMV); compiled = 0 when a row's node sets exceed the kernel's
 * capacities, the network kernels are then used */
int nlf_learned_info(nlf_precond p, int* compiled, int64_t* nnz, double* compile_seconds);
/* 0 = apply the compiled operator when available (default), 1 = always run the network kernels,
 * 2 = apply only the network's fixed Jacobi term s^2 r (s = diag^-1/2; the learned correction off,
 * the coarse part unchanged: what the trained network adds is the difference from this) */
int nlf_learned_set_apply(nlf_precond p, int mode);
/* No reference counterpart.  Two-level form of the learned preconditioner: B r = M r + P A_c^-1 P^T r
 * with P the 0/1 map of rows to about `aggregates` aggregates (a grid of cells over the node
 * coordinates, at most 4096 coarse rows and at least 16 nodes per aggregate on average) and
 * A_c = P^T a P inverted once on the device.  aggregates = 0 removes it. */
int nlf_learned_set_coarse(nlf_precond p, nlf_matrix a, int64_t aggregates);
int nlf_learned_coarse_info(nlf_precond p, int64_t* coarse_rows, double* setup_seconds);
int nlf_learned_export_graph(nlf_precond p, int64_t* outer, int32_t* cols, float* feat, float* s);
/* loss ||M(r) - e||_A^2 / ||e||_A^2 for a pair with A e = r, and its weight gradient (host) */
int nlf_learned_loss_grad(nlf_precond p, nlf_matrix a, const double* r_dev, const double* e_dev, double* loss,
                          float* grad_host);
/* training on the GPU: Adam over (residual, error) pairs of any number of problems */
int nlf_trainer_create(nlf_ctx ctx, const float* w_host, nlf_trainer* out);
int nlf_trainer_destroy(nlf_trainer t);
int nlf_trainer_step(nlf_trainer t, nlf_precond graph, nlf_matrix a, const double* r_dev, const double* e_dev, float lr,
                     double* loss);
int nlf_trainer_get_weights(nlf_trainer t, float* w_host);
/* flexible CG (conjugate_gradient.hpp:48-79 generalised to a variable preconditioner) with the
 * learned preconditioner; on stagnation continues with PCG and `fallback` */
int nlf_fcg_solve(nlf_matrix a, nlf_precond learned, nlf_precond fallback, const double* b_dev, double* x_dev, double tol,
                  int64_t max_it, int mode, nlf_solve_stats* stats);

/* ---- batched solves (no reference counterpart: the reference solves one problem per call).  Many
 * independent instances assembled into one matrix as diagonal blocks (instance i owns rows
 * [row_begin[i], row_begin[i+1])) and solved by lockstep Jacobi-PCG with per-instance scalars,
 * G iterations per captured hipGraph replay; each instance stops on its FP64 true residual. */
int nlf_batch_create(nlf_matrix a, int count, const int64_t* row_begin, nlf_batch* out);
int nlf_batch_destroy(nlf_batch b);
/* picks each instance's SpMV tile (lanes per row, rows per block) by timing; cache_path may be NULL */
int nlf_batch_autotune(nlf_batch b, const char* cache_path, double* chosen);
/* No reference counterpart.  ILLT (IC(0) of the batch's local matrix, block diagonal like the batch) for the
 * batch's solves, or NULL for Jacobi (the default); the handle must outlive the solves. */
int nlf_batch_set_preconditioner(nlf_batch b, nlf_precond p);
int nlf_batch_solve(nlf_batch b, const double* b_dev, double* x_dev, double tol, int64_t max_it, int mode, int use_graph,
                    int64_t* iterations, double* true_residual, nlf_solve_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* NLF_H */
