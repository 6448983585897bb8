// Internal declarations shared by the HIP translation units of libnlf.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nlf.h"

namespace nlf {

using i64 = int64_t;
using i32 = int32_t;
using u32 = uint32_t;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define NLF_HIP(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::nlf::Error(NLF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
    } while (0)

#define NLF_CHECK(cond, code, msg)                                                                      \
    do {                                                                                                \
        if (!(cond)) throw ::nlf::Error((code), (msg));                                                 \
    } while (0)

// Device memory of the library's buffers.  While an entry point runs on a context (CtxScope, set by
// guarded() and by the destroy functions), a block released by a buffer of that context goes back
// to the context's free list instead of hipFree, and later allocations on that context take blocks
// from there (best fit, at most 25 % larger).  All work of a context is queued on its one stream,
// so a block taken again is only used by work queued after everything that used it before its
// release.  hipMalloc of a large block costs up to a millisecond and hipFree waits for the whole
// device; a workload that assembles and solves again and again reuses its blocks.  Outside a
// context scope (or for another context's buffer) blocks are plain hipMalloc / hipFree.  The free
// lists of all contexts on a device hold at most an eighth of its memory together (past that, the
// context's oldest free blocks are returned first); a context's list is emptied when hipMalloc fails
// and when the context is destroyed.
struct CtxScope {
    nlf_ctx prev;
    explicit CtxScope(nlf_ctx c);
    ~CtxScope();
    CtxScope(const CtxScope&) = delete;
    CtxScope& operator=(const CtxScope&) = delete;
};
nlf_ctx live_ctx(nlf_ctx c);                                // c, or null once it is destroyed
void* device_alloc(size_t bytes, size_t* cap, nlf_ctx* owner);
void device_free(void* p, size_t cap, nlf_ctx owner);

// Owning device buffer.
template<class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    size_t cap = 0;              // bytes of the block
    nlf_ctx owner = nullptr;     // the context whose free list takes the block back
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), cap(o.cap), owner(o.owner) { o.forget(); }
    DevBuf& operator=(DevBuf&& o) noexcept {
        release();
        p = o.p; n = o.n; cap = o.cap; owner = o.owner;
        o.forget();
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (!count) return;
        p = static_cast<T*>(device_alloc(count * sizeof(T), &cap, &owner));
        n = count;
    }
    void release() {
        if (p) device_free(p, cap, owner);
        forget();
    }
    size_t bytes() const { return n * sizeof(T); }
private:
    void forget() { p = nullptr; n = 0; cap = 0; owner = nullptr; }
};

template<class T>
void upload(DevBuf<T>& d, const T* h, size_t count, hipStream_t s) {
    d.alloc(count);
    if (count) NLF_HIP(hipMemcpyAsync(d.p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
}

template<class T>
void download(T* h, const T* d, size_t count, hipStream_t s) {
    if (count) NLF_HIP(hipMemcpyAsync(h, d, count * sizeof(T), hipMemcpyDeviceToHost, s));
    NLF_HIP(hipStreamSynchronize(s));
}

inline unsigned blocks_for(i64 n, int threads) { return unsigned((n + threads - 1) / threads); }

// ---- FE tables (host; the device gets them by value in kernel arguments) ----------------------
constexpr int MAX_NN = 8;
constexpr int MAX_NQ = 9;

struct FeTables {
    int nn = 0, nq = 0;
    double w[MAX_NQ];
    double qN[MAX_NN * MAX_NQ], qNxi[MAX_NN * MAX_NQ], qNeta[MAX_NN * MAX_NQ];  // [i*nq + q]
    double N0[MAX_NN];
    int nearest[MAX_NN];                 // quadrature point nearest to each node (nearest_qnode)
};

struct BoundaryTables {
    int nb = 0, nq = 0;
    double w[2], qN[3 * 2], qNxi[3 * 2];
};

FeTables make_fe_tables(int nn);
BoundaryTables make_boundary_tables(int nb);

// ---- influence function (device evaluation parameters) ---------------------------------------
// Distance kinds: the Euclidean one (n = 2, the common case, with its own fast path) and the others
// of powered_distance (search_function.hpp:28-73) and powered_distance_with_rotation (:75-90).
enum : int { DIST_L2 = 0, DIST_L1 = 1, DIST_LINF = 2, DIST_LN_INT = 3, DIST_LN_REAL = 4, DIST_ROTATION = 5 };

struct DistanceParams {
    int kind = DIST_L2;
    int n = 2;                           // DIST_LN_INT exponent
    double n_real = 2.0;                 // DIST_LN_REAL exponent
    bool symmetric() const { return kind != DIST_ROTATION; }
};

// Validated distance of an nlf_distance / the distance fields of an nlf_model.
DistanceParams make_distance(int kind, int n, double n_real);

struct InfluenceParams {
    int kind = NLF_INFLUENCE_POLYNOMIAL;
    int fast = 0;                        // fast_polynomial: norm*(1-d)
    int p_int = 1, q_int = 1;            // exponent p/n and q are integers
    int p_i = 1, q_i = 1;
    double p_d = 1.0, q_d = 1.0;
    double exp_q = -0.5;
    double r[2] = {0, 0};
    double inv_r[2] = {0, 0};
    double norm = 1.0;
    DistanceParams dist;
};

InfluenceParams make_influence(const nlf_model& m);

__device__ __forceinline__ double ipow(double x, int e) {     // metamath/functions/power.hpp:25-40 (integral exponent)
    double r = 1.0;
    while (e > 0) {
        if (e & 1) { r *= x; --e; } else { x *= x; e >>= 1; }
    }
    return r;
}

__device__ __forceinline__ double exp_pow(double v, int is_int, int ei, double ed) {
    if (!is_int) return pow(v, ed);
    return ei == 1 ? v : ipow(v, ei);
}

// powered_distance(x, y, r) for the kinds other than DIST_L2, operation for operation like the
// reference (search_function.hpp:28-90; metamath distance.hpp:17-31, norm.hpp:15-43, power.hpp:9-46):
// z_i = (x_i - y_i) / r_i, then n = 1: (0 + |z0|) + |z1|; the maximum norm: the larger |z_i| (a norm,
// not powered); integer n: (0 + |z0|^n) + |z1|^n by repeated squaring; real n: the same with pow.
// Rotation (an ellipse whose r0 axis points along x from the origin and whose r1 axis is
// perpendicular to it): (r0 (x0 y1 - x1 y0))^2 + (r1 (x0 (x0 - y0) + x1 (x1 - y1)))^2 over
// (r0 r1)^2 |x|^2 -- not symmetric in x and y, and NaN at x = 0, like the reference.
__device__ __forceinline__ double general_distance(const DistanceParams& d, double r0, double r1, double ax, double ay,
                                                   double bx, double by) {
    if (d.kind == DIST_ROTATION) {
        const double u = r0 * (ax * by - ay * bx);
        const double v = r1 * (ax * (ax - bx) + ay * (ay - by));
        const double rr = r0 * r1;
        double xx = 0.0;
        xx = xx + ax * ax;
        xx = xx + ay * ay;
        return (u * u + v * v) / (rr * rr * xx);
    }
    const double z0 = fabs((ax - bx) / r0), z1 = fabs((ay - by) / r1);
    if (d.kind == DIST_LINF) return z0 < z1 ? z1 : z0;
    double s = 0.0;
    if (d.kind == DIST_L1) {
        s = s + z0;
        s = s + z1;
    } else if (d.kind == DIST_LN_INT) {
        s = s + ipow(z0, d.n);
        s = s + ipow(z1, d.n);
    } else if (d.kind == DIST_LN_REAL) {
        s = s + pow(z0, d.n_real);
        s = s + pow(z1, d.n_real);
    } else {
        s = s + z0 * z0;
        s = s + z1 * z1;
    }
    return s;
}

// phi(x, y): normalised influence of the nonlocal radius (influence_functions_2d.hpp:39-148).
__device__ __forceinline__ double influence(const InfluenceParams& f, double ax, double ay, double bx, double by) {
    double d;
    if (f.dist.kind == DIST_L2) {
        const double dx = ax - bx, dy = ay - by;
        const double zx = dx * f.inv_r[0], zy = dy * f.inv_r[1];
        d = zx * zx + zy * zy;
        if (fabs(d - 1.0) < 1e-12) {      // near the cutoff: the reference's division form decides
            const double ex = dx / f.r[0], ey = dy / f.r[1];
            d = ex * ex + ey * ey;
        }
    } else {
        d = general_distance(f.dist, f.r[0], f.r[1], ax, ay, bx, by);
    }
    if (f.fast) return d < 1.0 ? f.norm * (1.0 - d) : 0.0;
    if (f.kind == NLF_INFLUENCE_POLYNOMIAL)
        return d < 1.0 ? f.norm * exp_pow(1.0 - exp_pow(d, f.p_int, f.p_i, f.p_d), f.q_int, f.q_i, f.q_d) : 0.0;
    if (f.kind == NLF_INFLUENCE_CONSTANT) return d <= 1.0 ? f.norm : 0.0;
    return f.norm * exp(f.exp_q * exp_pow(d, f.p_int, f.p_i, f.p_d));
}

// phi specialised by influence class (influence_class).  POLY = 1 or 2: polynomial influence with p = n (d^(p/n) = d) and q = POLY
// (the fast polynomial, the bell), evaluated with the same operations as influence() -- the same
// division form near the cutoff, (1 - d) and its square as one multiply -- without the code of
// the other kinds, which keeps the loops that call it small; POLY = 0: influence() itself.
template<int POLY>
__device__ __forceinline__ double influence_poly(const InfluenceParams& f, double ax, double ay, double bx, double by) {
    if (POLY == 0) return influence(f, ax, ay, bx, by);
    const double dx = ax - bx, dy = ay - by;
    const double zx = dx * f.inv_r[0], zy = dy * f.inv_r[1];
    double d = zx * zx + zy * zy;
    if (fabs(d - 1.0) < 1e-12) {
        const double ex = dx / f.r[0], ey = dy / f.r[1];
        d = ex * ex + ey * ey;
    }
    const double s = 1.0 - d;
    return d < 1.0 ? f.norm * (POLY == 1 ? s : s * s) : 0.0;
}

// FP32 phi for the FP32 assembly: the distance and the polynomial are evaluated in FP32; within
// 4e-6 of the cutoff the FP64 evaluation decides, so that the support (and the discontinuity of
// the constant influence) is the FP64 one.
__device__ __forceinline__ float exp_pow32(float v, int is_int, int ei, double ed) {
    if (!is_int) return powf(v, float(ed));
    float r = 1.0f;
    while (ei > 0) {
        if (ei & 1) { r *= v; --ei; } else { v *= v; ei >>= 1; }
    }
    return r;
}

__device__ __forceinline__ float influence32(const InfluenceParams& f, double ax, double ay, double bx, double by) {
    if (f.dist.kind != DIST_L2) return float(influence(f, ax, ay, bx, by));
    const float zx = float(ax - bx) * float(f.inv_r[0]), zy = float(ay - by) * float(f.inv_r[1]);
    const float d = zx * zx + zy * zy;
    if (fabsf(d - 1.0f) < 4e-6f) return float(influence(f, ax, ay, bx, by));
    const float norm = float(f.norm);
    if (f.fast) return d < 1.0f ? norm * (1.0f - d) : 0.0f;
    if (f.kind == NLF_INFLUENCE_POLYNOMIAL)
        return d < 1.0f ? norm * exp_pow32(1.0f - exp_pow32(d, f.p_int, f.p_i, f.p_d), f.q_int, f.q_i, f.q_d) : 0.0f;
    if (f.kind == NLF_INFLUENCE_CONSTANT) return d <= 1.0f ? norm : 0.0f;
    return norm * expf(float(f.exp_q) * exp_pow32(d, f.p_int, f.p_i, f.p_d));
}

struct GroupParams {                     // per element group, copied to device
    double p1;                           // local weight (1 for LOCAL groups)
    double c[6];                         // thermal conductivity {l11, l22, l12}, or the Hooke matrix of any
                                         // kind in the anisotropic layout {h11, h12, h66, h22, h16, h26}
    double alpha[3];                     // thermal expansion {a_xx, a_yy, a_xy} (elasticity)
    int nonlocal;                        // theory NONLOCAL
    int ckind;                           // conductivity kind
    int ekind;                           // thermal expansion kind (NLF_EXPANSION_*)
    int varying;                         // coefficients per quadrature point (AssemblyArgs::cq) instead of c
    InfluenceParams inf;
};

GroupParams model_group_params(const nlf_model& md, bool only_local);

// Hooke matrix and thermal expansion of a plane-stress material into gp.c / gp.alpha / gp.ekind
// (elastic_parameters.hpp:62-133, thermal_expansion_parameters.hpp:7-18).  Isotropic and orthotropic
// matrices take the anisotropic layout with h22 = h11 (isotropic) and h16 = h26 = 0, so one
// integrand serves every kind.
void elastic_group_params(const nlf_elastic_material& e, GroupParams& gp);

// Influence class shared by the nonlocal groups, for influence_poly<POLY>: 1 = fast polynomial,
// 2 = bell-shaped polynomial (p = n, q = 2), 0 = anything else (or groups of different classes),
// -1 = no nonlocal group.
int influence_class(const std::vector<GroupParams>& gp);

}  // namespace nlf

// ---- opaque handle definitions -------------------------------------------------------------------
struct nlf_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string error;
    int assembly_pairs = 1;              // NLF_OPT_ASSEMBLY_PAIRS
    int row_patterns = 1;                // NLF_OPT_ROW_PATTERNS
    int assembly_threads = 64;           // NLF_OPT_ASSEMBLY_THREADS
    int pattern_threads = 128;           // NLF_OPT_PATTERN_THREADS
    int pair_budget_mib = 0;             // NLF_OPT_PAIR_BUDGET_MIB (0: automatic)
    int spmv_symmetric = 1;              // NLF_OPT_SPMV_SYMMETRIC
    // free device blocks of this context's buffers (device_alloc / device_free)
    bool alive = true;
    std::mutex blocks_mu;
    std::multimap<size_t, std::pair<uint64_t, void*>> free_blocks;                    // bytes -> (age, block)
    std::map<uint64_t, std::multimap<size_t, std::pair<uint64_t, void*>>::iterator> free_order;  // oldest first
    uint64_t free_seq = 0;
    size_t cached_bytes = 0;
};

struct nlf_mesh_s {
    nlf_ctx ctx = nullptr;
    nlf::i64 N = 0, E = 0;
    int nn = 0, nb = 0, G = 0, GB = 0;
    bool injective = true;               // for every local index j, e -> elems[e][j] is injective
    int max_degree = 0;                  // most elements at one node
    nlf::FeTables fe;
    nlf::BoundaryTables bt;
    // host copies used by boundary integrals and level analysis
    std::vector<double> xy_h;
    std::vector<nlf::i32> el_h, be_h;
    std::vector<nlf::i64> gbeg_h, bbeg_h;
    std::vector<nlf::i32> ne_off_h, ne_ids_h;   // node -> elements (ascending), as ne_off / ne_ids
    std::vector<uint32_t> el_key_h;             // Morton (Z-order) key of each element's centroid
    double bbox[4] = {0.0, 0.0, 0.0, 0.0};       // x min, x max, y min, y max of the nodes
    // device geometry and precompute
    nlf::DevBuf<double> xy;              // 2N
    nlf::DevBuf<nlf::i32> el;            // E*nn
    nlf::DevBuf<nlf::i32> group_of;      // E
    nlf::DevBuf<nlf::i32> ne_off, ne_ids; // node -> elements (ascending), int32 offsets (4*E < 2^31)
    nlf::DevBuf<nlf::i32> el_maxnode;    // max node id of each element
    nlf::DevBuf<uint8_t> el_rank;        // E*nn, non-injective meshes only: position of e in the
                                         // ascending element list of its node j
    nlf::DevBuf<double> xq, jac, D;      // 2*E*nq, E*nq, 2*E*nn*nq ([e][i][q])
    nlf::DevBuf<double> centres;         // 2E
    // neighbours
    bool has_neighbours = false;
    nlf::DevBuf<nlf::i64> nb_off;        // E+1
    nlf::DevBuf<nlf::i32> nb_ids;
    std::vector<double> expanded;        // 2G
    std::vector<int> searched;           // group searched (radius > 0)
    bool nb_symmetric = true;            // e' in N(e) <=> e in N(e') (false after a rotated-distance search)
    double last_neighbour_seconds = 0.0;
};

struct nlf_symv_s;                       // the symmetric SpMV storage (nlf_symv.inl)

struct nlf_matrix_s {
    ~nlf_matrix_s();
    nlf_ctx ctx = nullptr;
    nlf_mesh mesh = nullptr;
    int dof = 1;
    nlf::i64 rows = 0;
    // inner part stored as the full (both-triangle) pattern; the upper triangle of each row is its
    // tail starting at diag[r] (the reference's upper CSR is a view of it).
    nlf::i64 nnz_full = 0, nnz_upper = 0, nnz_bound = 0;
    nlf::DevBuf<nlf::i64> outer;         // rows+1
    nlf::DevBuf<nlf::i32> cols;
    nlf::DevBuf<nlf::i64> diag;          // position of the diagonal in each row
    nlf::DevBuf<nlf::i64> tpos;          // position of the transposed entry (filled on demand)
    nlf::DevBuf<double> val;             // full values
    nlf::DevBuf<float> val32;            // FP32 copy (mixed precision), on demand
    nlf::DevBuf<int16_t> val_lo;         // val - val32 in units of 2^-16 ulp(val32) (split mode), on demand
    nlf::DevBuf<int16_t> off16;          // col - row per nonzero when every |col - row| < 2^15 (SpMV)
    nlf::DevBuf<int16_t> pat;            // dictionary of the distinct off16 rows (SpMV), when it pays
    nlf::DevBuf<nlf::i64> pstart;        // row r's offsets are pat[pstart[r] + p], p in [outer[r], outer[r+1])
    nlf::i64 pat_size = 0, pat_groups = 0;
    bool spmv_storage_ready = false;     // off16 / pat / pstart built (ensure_spmv_storage)
    int row_patterns = 1;                // NLF_OPT_ROW_PATTERNS when the values were assembled
    nlf::DevBuf<nlf::i64> b_outer;       // bound part: rows+1
    nlf::DevBuf<nlf::i32> b_cols;
    nlf::DevBuf<double> b_val;
    nlf::DevBuf<uint8_t> inner;          // is_inner per row
    double seconds_pattern = 0.0, seconds_values = 0.0;
    // what nlf_matrix_assemble_fp32 needs to repeat the value pass in FP32
    int physics = 0;
    bool only_local = false;
    std::vector<nlf::GroupParams> groups;
    std::vector<nlf::i64> extra_rc;      // pairs added after assembly (convection, nlf_matrix_add_pairs)
    std::vector<double> extra_v;
    bool val32_assembled = false;        // val32 holds FP32-assembled values (not the rounded copy)
    bool split_ready = false;            // val32 + val_lo is the split of the current val
    bool combined = false;               // values changed by nlf_matrix_combine (no FP32 value pass)
    nlf::DevBuf<double> cq;              // coefficients per quadrature point, [(e * nq + q) * 6 + k], or empty
    bool neumann = false;                // bordered by the Lagrange row and column of a Neumann problem
    // values version: bumped by every change of val; the symmetric SpMV storage is rebuilt from val
    // when its version is stale, and on every solve once val was handed out (nlf_matrix_device)
    nlf::i64 val_version = 0, symv_refused_version = -1;
    bool values_exposed = false;
    nlf_symv_s* symv = nullptr;
};

namespace nlf {
// helpers implemented across translation units
void build_row_patterns(nlf_matrix_s* a, hipStream_t s);    // fills a->pat / a->pstart when they pay
// Same for any CSR with 16-bit offsets; returns whether the dictionary was kept.
bool build_row_patterns(i64 rows, i64 nnz, const i64* outer, const int16_t* off16, DevBuf<int16_t>& pat,
                        DevBuf<i64>& pstart, i64& entries, i64& groups, hipStream_t s);

// Column source of the SpMV: the row-pattern dictionary, else the 16-bit offsets, else int32 columns.
struct SpmvCols {
    const i32* cols;
    const int16_t* off;       // offsets (col - row); with pstart: the dictionary
    const i64* pstart;
};
i64 spmv_stream_bytes(const nlf_matrix_s* a, size_t sv, size_t sx);   // HBM bytes of one SpMV

// Aggregation coarse space for the learned preconditioner's two-level form:
// B r = M r + P A_c^-1 P^T r, P the 0/1 map of rows to aggregates (a grid of cells over the node
// coordinates, one coarse row per aggregate and DoF), A_c = P^T A P kept as its dense inverse.
struct CoarseSpace {
    i64 n = 0;                 // fine rows
    i64 nc = 0;                // coarse rows
    i64 ncp = 0;               // coarse rows padded to a multiple of 64 (identity padding)
    DevBuf<i32> agg;           // [n] coarse row of each fine row
    DevBuf<i64> mem_off;       // [nc + 1] fine rows of each coarse row (ascending)
    DevBuf<i32> mem;           // [n]
    DevBuf<double> inv;        // [ncp * ncp] A_c^-1, row-major
    DevBuf<double> rc, yc;     // [ncp] restricted residual, coarse correction
    double setup_seconds = 0.0;
};
std::unique_ptr<CoarseSpace> build_coarse(const nlf_matrix_s* a, i64 target, hipStream_t s);
void coarse_apply_add(const CoarseSpace& c, const double* r, double* z, hipStream_t s);   // z += P A_c^-1 P^T r
// Builds the full-storage SpMV's 16-bit offsets and row-pattern dictionary on first use (and waits for
// them on the matrix's stream).
void ensure_spmv_storage(const nlf_matrix_s* a);
inline SpmvCols spmv_cols(const nlf_matrix_s* a) {
    ensure_spmv_storage(a);
    if (a->pstart.p) return {a->cols.p, a->pat.p, a->pstart.p};
    return {a->cols.p, a->off16.p, nullptr};
}
void exclusive_scan_i64(i64* data, i64 n, hipStream_t s);   // in place over n+1 entries, data[n] = total
int noalloc_scratch_init(int device);   // device address of nlf_cg_solve_noalloc's work memory, per device
// (keys, values) sorted by key, stable (radix sort); n pairs
void sort_pairs_u64(const uint64_t* keys_in, uint64_t* keys_out, const i32* vals_in, i32* vals_out, i64 n, hipStream_t s);
void sort_keys_u64(const unsigned long long* keys_in, unsigned long long* keys_out, i64 n, hipStream_t s);
void set_error(nlf_ctx ctx, const std::string& msg);

// Runs f, translating exceptions into status codes and the context's last error.
template<class F>
int guarded(nlf_ctx ctx, F&& f) {
    CtxScope scope(ctx);
    try {
        f();
        return NLF_OK;
    } catch (const Error& e) {
        set_error(ctx, e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error(ctx, "host allocation failed");
        return NLF_ERR_ARGUMENT;
    } catch (const std::exception& e) {
        set_error(ctx, e.what());
        return NLF_ERR_ARGUMENT;
    }
}
}  // namespace nlf
