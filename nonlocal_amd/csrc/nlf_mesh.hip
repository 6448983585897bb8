// Context, finite-element tables, mesh upload and per-quadrature-point precompute.
//
// This translation unit is compiled with -ffp-contract=off: element centres feed the neighbour
// predicate, which must reproduce the reference's arithmetic op for op
// (mesh_container_2d.hpp:130-139, search_function.hpp:28-63).
#include "nlf_internal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <cmath>

namespace nlf {

void set_error(nlf_ctx ctx, const std::string& msg) {
    if (ctx) ctx->error = msg;
}

// ---- FE tables ---------------------------------------------------------------------------------
// Gauss rules (gaussian_quadrature.hpp:22-40); 2D points x = -1 + (g + 1)*1 and weights w = wx*wy
// as set by element_2d_integrate::set_quadrature (element_2d_integrate.hpp:69-106), q = ix*ny + iy.
static void gauss(int k, double* x, double* w) {
    if (k == 1) { x[0] = 0.0; w[0] = 2.0; }
    else if (k == 2) { const double s = std::sqrt(3.0); x[0] = -1.0 / s; x[1] = 1.0 / s; w[0] = w[1] = 1.0; }
    else { const double s = std::sqrt(3.0 / 5.0); x[0] = -s; x[1] = 0.0; x[2] = s; w[0] = 5.0 / 9.0; w[1] = 8.0 / 9.0; w[2] = 5.0 / 9.0; }
}

// Bilinear basis serendipity<T,1> (serendipity_2d.hpp:52-63).
static void basis_q4(double x, double y, double* N, double* Nx, double* Ny) {
    const double sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
    for (int i = 0; i < 4; ++i) {
        N[i] = (1 + sx[i] * x) * (1 + sy[i] * y) / 4;
        Nx[i] = sx[i] * (1 + sy[i] * y) / 4;
        Ny[i] = (1 + sx[i] * x) * sy[i] / 4;
    }
}

// Serendipity Q8 with the reference's default parameter p = 2/9 (serendipity_2d.hpp:101-112).
// Corners: (1+sx x)(1+sy y)(a(1 - sx x - sy y) + b sx sy x y)/16; midsides on the edge y = sy:
// -(1-x^2)(1+sy y)(c - sy b y)/16 (and the same with x, y swapped), a = 9p-1, b = 9p+3, c = 9p-5.
static void basis_q8(double x, double y, double* N, double* Nx, double* Ny) {
    const double p = 2.0 / 9.0, a = 9 * p - 1, b = 9 * p + 3, c = 9 * p - 5;
    const double sxc[4] = {-1, 1, 1, -1}, syc[4] = {-1, -1, 1, 1};
    for (int k = 0; k < 4; ++k) {
        const double sx = sxc[k], sy = syc[k];
        const double X = 1 + sx * x, Y = 1 + sy * y;
        const double g = a * (1 - sx * x - sy * y) + b * sx * sy * x * y;
        const double gx = -a * sx + b * sx * sy * y, gy = -a * sy + b * sx * sy * x;
        N[2 * k] = X * Y * g / 16;
        Nx[2 * k] = (sx * Y * g + X * Y * gx) / 16;
        Ny[2 * k] = (X * sy * g + X * Y * gy) / 16;
    }
    // node 1 (0,-1), node 5 (0,1): edge y = s
    for (int s = -1; s <= 1; s += 2) {
        const int id = s < 0 ? 1 : 5;
        const double f = -(1 - x * x), Y = 1 + s * y, g = c - s * b * y;
        N[id] = f * Y * g / 16;
        Nx[id] = 2 * x * Y * g / 16;
        Ny[id] = f * (s * g - Y * s * b) / 16;
    }
    // node 3 (1,0), node 7 (-1,0): edge x = s
    for (int s = -1; s <= 1; s += 2) {
        const int id = s > 0 ? 3 : 7;
        const double f = -(1 - y * y), X = 1 + s * x, g = c - s * b * x;
        N[id] = f * X * g / 16;
        Ny[id] = 2 * y * X * g / 16;
        Nx[id] = f * (s * g - X * s * b) / 16;
    }
}

FeTables make_fe_tables(int nn) {
    FeTables t{};
    t.nn = nn;
    const int k = nn == NLF_ELEMENT_Q4 ? 2 : 3;
    t.nq = k * k;
    double g[3], gw[3], px[3];
    gauss(k, g, gw);
    for (int i = 0; i < k; ++i) px[i] = -1.0 + (g[i] - (-1.0)) * 1.0;
    double N[MAX_NN], Nx[MAX_NN], Ny[MAX_NN];
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
            t.w[i * k + j] = gw[i] * 1.0 * gw[j] * 1.0;
            if (nn == NLF_ELEMENT_Q4) basis_q4(px[i], px[j], N, Nx, Ny); else basis_q8(px[i], px[j], N, Nx, Ny);
            for (int a = 0; a < nn; ++a) {
                t.qN[a * t.nq + i * k + j] = N[a];
                t.qNxi[a * t.nq + i * k + j] = Nx[a];
                t.qNeta[a * t.nq + i * k + j] = Ny[a];
            }
        }
    // nearest_qnode (element_2d_integrate.hpp:87-105): the first quadrature point at the smallest
    // distance from the node in the reference element
    static const double q4_nodes[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    static const double q8_nodes[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}};
    for (int a = 0; a < nn; ++a) {
        const double* v = nn == NLF_ELEMENT_Q4 ? q4_nodes[a] : q8_nodes[a];
        double best = 1e300;
        t.nearest[a] = 0;
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) {
                const double d = std::sqrt((v[0] - px[i]) * (v[0] - px[i]) + (v[1] - px[j]) * (v[1] - px[j]));
                if (best > d) { best = d; t.nearest[a] = i * k + j; }
            }
    }
    // centre weights N_i(0,0): 1/4 (Q4); (9p-1)/16 corners and -(9p-5)/16 midsides (Q8)
    if (nn == NLF_ELEMENT_Q4) for (int a = 0; a < 4; ++a) t.N0[a] = 0.25;
    else {
        const double p = 2.0 / 9.0;
        for (int a = 0; a < 8; a += 2) { t.N0[a] = (9 * p - 1) / 16; t.N0[a + 1] = -(9 * p - 5) / 16; }
    }
    return t;
}

// Boundary edges: linear (Gauss 1) or quadratic Lagrange with nodes -1, 0, 1 (Gauss 2)
// (vtk_elements_set.hpp:43-52).
BoundaryTables make_boundary_tables(int nb) {
    BoundaryTables t{};
    t.nb = nb;
    t.nq = nb == 2 ? 1 : 2;
    double g[3], gw[3];
    gauss(t.nq, g, gw);
    for (int q = 0; q < t.nq; ++q) {
        const double x = g[q];
        t.w[q] = gw[q];
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

// ---- precompute kernels --------------------------------------------------------------------------
struct TablesArg { FeTables t; };

// One thread per (element, quadrature point): x_q, |det J|, D_i = detJ * grad N_i
// (mesh_container_2d_utils.hpp:75-123, mesh_container_2d.hpp:141-160).
__global__ void k_precompute(i64 E, const double* __restrict__ xy, const i32* __restrict__ el, TablesArg ta,
                             double* __restrict__ xq, double* __restrict__ jac, double* __restrict__ D) {
    const FeTables& t = ta.t;
    const i64 gid = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= E * t.nq) return;
    const i64 e = gid / t.nq;
    const int q = int(gid % t.nq);
    double cx = 0, cy = 0, J00 = 0, J01 = 0, J10 = 0, J11 = 0;
    for (int i = 0; i < t.nn; ++i) {
        const i32 v = el[e * t.nn + i];
        const double x = xy[2 * v], y = xy[2 * v + 1];
        const double N = t.qN[i * t.nq + q], Nx = t.qNxi[i * t.nq + q], Ny = t.qNeta[i * t.nq + q];
        cx = cx + x * N; cy = cy + y * N;
        J00 = J00 + x * Nx; J01 = J01 + x * Ny;
        J10 = J10 + y * Nx; J11 = J11 + y * Ny;
    }
    xq[2 * gid] = cx; xq[2 * gid + 1] = cy;
    jac[gid] = fabs(J00 * J11 - J01 * J10);
    for (int i = 0; i < t.nn; ++i) {
        const double Nx = t.qNxi[i * t.nq + q], Ny = t.qNeta[i * t.nq + q];
        const i64 o = 2 * ((e * t.nn + i) * t.nq + q);
        D[o] = Nx * J11 - Ny * J10;
        D[o + 1] = -Nx * J01 + Ny * J00;
    }
}

// Element centres c_e = sum_i x_i N_i(0,0), accumulated (0 + a_0) + a_1 + ... without contraction.
__global__ void k_centres(i64 E, int nn, const double* __restrict__ xy, const i32* __restrict__ el, TablesArg ta,
                          double* __restrict__ c) {
    const i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double cx = 0.0, cy = 0.0;
    for (int i = 0; i < nn; ++i) {
        const i32 v = el[e * nn + i];
        cx = cx + xy[2 * v] * ta.t.N0[i];
        cy = cy + xy[2 * v + 1] * ta.t.N0[i];
    }
    c[2 * e] = cx;
    c[2 * e + 1] = cy;
}

__global__ void k_elem_maxnode(i64 E, int nn, const i32* __restrict__ el, i32* __restrict__ mx) {
    const i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    i32 m = el[e * nn];
    for (int i = 1; i < nn; ++i) m = max(m, el[e * nn + i]);
    mx[e] = m;
}

// ---- device blocks of a context's buffers ------------------------------------------------------
static thread_local nlf_ctx t_scope_ctx = nullptr;

CtxScope::CtxScope(nlf_ctx c) : prev(t_scope_ctx) { t_scope_ctx = c; }
CtxScope::~CtxScope() { t_scope_ctx = prev; }

nlf_ctx live_ctx(nlf_ctx c) { return c && c->alive ? c : nullptr; }

// 2 MiB granules from 1 MiB up (a workload's large buffers repeat their sizes), 512 B below
static size_t block_bytes(size_t bytes) {
    const size_t g = bytes >= (size_t(1) << 20) ? (size_t(2) << 20) : 512;
    return (bytes + g - 1) / g * g;
}

// Bytes held by the free lists of all contexts of a device, against one limit per device (an eighth of
// its memory), so that several contexts together hold no more than one would.
constexpr int BLOCK_DEVICES = 64;
static std::atomic<size_t> g_free_bytes[BLOCK_DEVICES];
static std::atomic<size_t> g_free_limit[BLOCK_DEVICES];

static bool reserve_free_bytes(int device, size_t cap) {
    if (device < 0 || device >= BLOCK_DEVICES) return false;
    const size_t limit = g_free_limit[device].load();
    size_t cur = g_free_bytes[device].load();
    while (cur + cap <= limit)
        if (g_free_bytes[device].compare_exchange_weak(cur, cur + cap)) return true;
    return false;
}
static void unreserve_free_bytes(int device, size_t cap) {
    if (device >= 0 && device < BLOCK_DEVICES) g_free_bytes[device].fetch_sub(cap);
}

// Waits for the context's stream, then returns every free block to the device; caller holds blocks_mu.
static void release_free_blocks(nlf_ctx c) {
    if (c->free_blocks.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->free_blocks) (void)hipFree(kv.second.second);
    c->free_blocks.clear();
    c->free_order.clear();
    unreserve_free_bytes(c->device, c->cached_bytes);
    c->cached_bytes = 0;
}

void* device_alloc(size_t bytes, size_t* cap, nlf_ctx* owner) {
    nlf_ctx c = t_scope_ctx;
    *owner = c;
    void* p = nullptr;
    if (!c) {
        NLF_HIP(hipMalloc(&p, bytes));
        *cap = bytes;
        return p;
    }
    const size_t want = block_bytes(bytes);
    {
        std::lock_guard<std::mutex> lock(c->blocks_mu);
        auto it = c->free_blocks.lower_bound(want);
        if (it != c->free_blocks.end() && it->first <= want + want / 4) {
            p = it->second.second;
            *cap = it->first;
            c->cached_bytes -= it->first;
            unreserve_free_bytes(c->device, it->first);
            c->free_order.erase(it->second.first);
            c->free_blocks.erase(it);
            return p;
        }
    }
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();                // the failed allocation is not the next launch's error
        {
            std::lock_guard<std::mutex> lock(c->blocks_mu);
            release_free_blocks(c);
        }
        e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(NLF_ERR_HIP, std::string("hipMalloc(") + std::to_string(want) + " bytes): " + hipGetErrorString(e));
        }
    }
    *cap = want;
    return p;
}

void device_free(void* p, size_t cap, nlf_ctx owner) {
    if (owner && owner == t_scope_ctx && owner->alive) {
        std::lock_guard<std::mutex> lock(owner->blocks_mu);
        // over the limit: the context's oldest free blocks go back to the device first, since the
        // blocks just released are the ones a repeating workload takes again
        const bool fits = owner->device >= 0 && owner->device < BLOCK_DEVICES && cap <= g_free_limit[owner->device].load();
        bool room = fits && reserve_free_bytes(owner->device, cap);
        while (fits && !room && !owner->free_order.empty()) {
            auto oldest = owner->free_order.begin()->second;
            (void)hipFree(oldest->second.second);
            owner->cached_bytes -= oldest->first;
            unreserve_free_bytes(owner->device, oldest->first);
            owner->free_order.erase(owner->free_order.begin());
            owner->free_blocks.erase(oldest);
            room = reserve_free_bytes(owner->device, cap);
        }
        if (room) {
            const uint64_t age = owner->free_seq++;
            auto it = owner->free_blocks.emplace(cap, std::make_pair(age, p));
            owner->free_order.emplace(age, it);
            owner->cached_bytes += cap;
            return;
        }
    }
    (void)hipFree(p);
}

}  // namespace nlf

using namespace nlf;

extern "C" {

// Context: one per host thread; no global mutable state (cf. the reference's OpenMP-parallel,
// non-re-entrant solver objects, iterative_solver_base.hpp:16-17).
int nlf_ctx_create(int device, void* stream, nlf_ctx* out) {
    if (!out) return NLF_ERR_ARGUMENT;
    auto* c = new nlf_ctx_s();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) { delete c; return NLF_ERR_HIP; }
    // the fixed work memory of nlf_cg_solve_noalloc; a failed look-up leaves it unset, and only that
    // entry point then reports NLF_ERR_STATE
    (void)nlf::noalloc_scratch_init(device);
    if (stream) c->stream = static_cast<hipStream_t>(stream);
    else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return NLF_ERR_HIP; }
        c->own_stream = true;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { (void)hipGetLastError(); total_b = 0; }
    if (device >= 0 && device < BLOCK_DEVICES) g_free_limit[device].store(total_b / 8);
    *out = c;
    return NLF_OK;
}

int nlf_ctx_destroy(nlf_ctx ctx) {
    if (!ctx || !ctx->alive) return NLF_OK;
    {
        std::lock_guard<std::mutex> lock(ctx->blocks_mu);
        release_free_blocks(ctx);
        ctx->alive = false;
    }
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    // The structure itself stays allocated (a few hundred bytes): objects made on the context may be
    // destroyed after it, and their destroy functions read its `alive` flag.
    return NLF_OK;
}

const char* nlf_last_error(nlf_ctx ctx) { return ctx ? ctx->error.c_str() : "null context"; }

// Context options (no reference counterpart): NLF_OPT_ASSEMBLY_PAIRS 1 = MFMA element-pair
// contractions where implemented (thermal Q4, default), 0 = node-centric contraction;
// NLF_OPT_ROW_PATTERNS 1 = shared row-pattern dictionary for the SpMV of matrices assembled
// afterwards (default), 0 = per-nonzero 16-bit offsets; NLF_OPT_ASSEMBLY_THREADS = threads per
// node workgroup of the node-centric value pass (64 default, 128, 256, 512); NLF_OPT_PATTERN_THREADS
// = threads per node workgroup of the pattern build (64, 128 default, 256); NLF_OPT_PAIR_BUDGET_MIB =
// device memory for the MFMA pair blocks of one assembly pass in MiB (0 = 8 GiB); a mesh whose blocks
// do not fit is assembled in node chunks; NLF_OPT_SPMV_SYMMETRIC 1 = FP64 products from the upper
// triangle in the slice layout (default), 0 = from the full storage.
int nlf_ctx_set_option(nlf_ctx ctx, int option, int value) {
    if (!ctx) return NLF_ERR_ARGUMENT;
    if (option == NLF_OPT_ASSEMBLY_PAIRS) ctx->assembly_pairs = value != 0;
    else if (option == NLF_OPT_ROW_PATTERNS) ctx->row_patterns = value != 0;
    else if (option == NLF_OPT_ASSEMBLY_THREADS) {
        if (value != 64 && value != 128 && value != 256 && value != 512) return NLF_ERR_ARGUMENT;
        ctx->assembly_threads = value;
    } else if (option == NLF_OPT_PATTERN_THREADS) {
        if (value != 64 && value != 128 && value != 256) return NLF_ERR_ARGUMENT;
        ctx->pattern_threads = value;
    } else if (option == NLF_OPT_PAIR_BUDGET_MIB) {
        if (value < 0) return NLF_ERR_ARGUMENT;
        ctx->pair_budget_mib = value;
    } else if (option == NLF_OPT_SPMV_SYMMETRIC) {
        ctx->spmv_symmetric = value != 0;
    }
    else return NLF_ERR_ARGUMENT;
    return NLF_OK;
}

int nlf_ctx_synchronize(nlf_ctx ctx) {
    return guarded(ctx, [&] { NLF_HIP(hipStreamSynchronize(ctx->stream)); });
}

void* nlf_ctx_stream(nlf_ctx ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

// Mesh upload + mesh_2d::init precompute (mesh_2d.hpp:96-108).  Groups are contiguous element
// ranges group_begin[G+1]; boundary groups hold 1D edges (nb = 2 linear, 3 quadratic).
int nlf_mesh_create(nlf_ctx ctx, int64_t n_nodes, const double* xy, int64_t n_elems, int nn, const int32_t* elems,
                    int n_groups, const int64_t* group_begin, int64_t n_bedges, int nb, const int32_t* bedges,
                    int n_bgroups, const int64_t* bgroup_begin, nlf_mesh* out) {
    return guarded(ctx, [&] {
        NLF_CHECK(out && xy && elems && group_begin, NLF_ERR_ARGUMENT, "null argument");
        NLF_CHECK(nn == NLF_ELEMENT_Q4 || nn == NLF_ELEMENT_Q8, NLF_ERR_UNSUPPORTED, "only Q4 and Q8 elements");
        NLF_CHECK(n_groups >= 1 && group_begin[0] == 0 && group_begin[n_groups] == n_elems, NLF_ERR_ARGUMENT,
                  "groups must be contiguous ranges covering all elements");
        NLF_CHECK(n_nodes > 0 && n_nodes < (int64_t(1) << 30) && n_elems * nn < (int64_t(1) << 31), NLF_ERR_ARGUMENT,
                  "mesh too large for 32-bit node/element ids");
        NLF_CHECK(n_bedges == 0 || nb == (nn == NLF_ELEMENT_Q4 ? 2 : 3), NLF_ERR_ARGUMENT,
                  "boundary edges must be linear for Q4 and quadratic for Q8");
        auto m = std::make_unique<nlf_mesh_s>();
        m->ctx = ctx; m->N = n_nodes; m->E = n_elems; m->nn = nn; m->nb = nb ? nb : (nn == 4 ? 2 : 3);
        m->G = n_groups; m->GB = n_bgroups;
        m->fe = make_fe_tables(nn);
        m->bt = make_boundary_tables(m->nb);
        m->xy_h.assign(xy, xy + 2 * n_nodes);
        m->el_h.assign(elems, elems + n_elems * nn);
        m->gbeg_h.assign(group_begin, group_begin + n_groups + 1);
        if (n_bgroups > 0) m->bbeg_h.assign(bgroup_begin, bgroup_begin + n_bgroups + 1);
        else m->bbeg_h.assign(1, 0);
        if (n_bedges > 0) m->be_h.assign(bedges, bedges + n_bedges * m->nb);
        for (i64 k = 0; k < n_elems * nn; ++k)
            NLF_CHECK(elems[k] >= 0 && elems[k] < n_nodes, NLF_ERR_ARGUMENT, "element node id out of range");
        for (int g = 0; g < n_groups; ++g)
            NLF_CHECK(group_begin[g] <= group_begin[g + 1], NLF_ERR_ARGUMENT, "group ranges must be ascending");
        // node -> elements in ascending element order (node_elements_2d, mesh_container_2d_utils.hpp:17-26)
        std::vector<i32> off(n_nodes + 1, 0), ids(size_t(n_elems) * nn);
        for (i64 k = 0; k < n_elems * nn; ++k) off[elems[k] + 1]++;
        for (i64 v = 0; v < n_nodes; ++v) off[v + 1] += off[v];
        std::vector<i32> fill(off.begin(), off.end() - 1);
        for (i64 e = 0; e < n_elems; ++e)
            for (int i = 0; i < nn; ++i) ids[fill[elems[e * nn + i]]++] = i32(e);
        std::vector<i32> group_of(n_elems);
        for (int g = 0; g < n_groups; ++g)
            for (i64 e = group_begin[g]; e < group_begin[g + 1]; ++e) group_of[e] = g;
        // injectivity of e -> elems[e][j] for every j (deterministic LDS accumulation in assembly)
        {
            std::vector<uint8_t> seen(n_nodes);
            for (int j = 0; j < nn && m->injective; ++j) {
                std::fill(seen.begin(), seen.end(), 0);
                for (i64 e = 0; e < n_elems; ++e) {
                    const i32 v = elems[e * nn + j];
                    if (seen[v]) { m->injective = false; break; }
                    seen[v] = 1;
                }
            }
        }
        for (i64 v = 0; v < n_nodes; ++v) m->max_degree = std::max(m->max_degree, off[v + 1] - off[v]);
        hipStream_t s = ctx->stream;
        if (!m->injective) {
            // rank of each element at each of its nodes: the assembly accumulates a node's row in
            // rounds by this rank instead of by local index, so no two contributions of one round
            // share a position on any mesh
            NLF_CHECK(m->max_degree <= 255, NLF_ERR_UNSUPPORTED, "a node shared by more than 255 elements");
            std::vector<uint8_t> rank(size_t(n_elems) * nn);
            for (i64 e = 0; e < n_elems; ++e)
                for (int i = 0; i < nn; ++i) {
                    const i32 v = elems[e * nn + i];
                    int r = 0;
                    while (ids[off[v] + r] != e) ++r;
                    rank[e * nn + i] = uint8_t(r);
                }
            upload(m->el_rank, rank.data(), rank.size(), s);
        }
        upload(m->xy, xy, size_t(2 * n_nodes), s);
        upload(m->el, elems, size_t(n_elems * nn), s);
        upload(m->ne_off, off.data(), off.size(), s);
        upload(m->ne_ids, ids.data(), ids.size(), s);
        m->ne_off_h = std::move(off);
        m->ne_ids_h = std::move(ids);
        // Morton key of each element's centroid on a 2^16 x 2^16 grid over the bounding box: the
        // pair kernel visits elements in this order, so that the elements in flight form compact
        // patches whose neighbour data stay in L2
        {
            double x0 = 1e308, x1 = -1e308, y0 = 1e308, y1 = -1e308;
            for (i64 v = 0; v < n_nodes; ++v) {
                x0 = std::min(x0, xy[2 * v]); x1 = std::max(x1, xy[2 * v]);
                y0 = std::min(y0, xy[2 * v + 1]); y1 = std::max(y1, xy[2 * v + 1]);
            }
            m->bbox[0] = x0; m->bbox[1] = x1; m->bbox[2] = y0; m->bbox[3] = y1;
            const double sx = 65535.0 / std::max(x1 - x0, 1e-300), sy = 65535.0 / std::max(y1 - y0, 1e-300);
            auto spread = [](uint32_t v) {         // bit k of v to bit 2k
                v &= 0xffffu;
                v = (v | (v << 8)) & 0x00ff00ffu;
                v = (v | (v << 4)) & 0x0f0f0f0fu;
                v = (v | (v << 2)) & 0x33333333u;
                v = (v | (v << 1)) & 0x55555555u;
                return v;
            };
            m->el_key_h.resize(size_t(n_elems));
            for (i64 e = 0; e < n_elems; ++e) {
                double cx = 0.0, cy = 0.0;
                for (int i = 0; i < nn; ++i) {
                    cx += xy[2 * i64(elems[e * nn + i])];
                    cy += xy[2 * i64(elems[e * nn + i]) + 1];
                }
                const uint32_t qx = uint32_t(std::clamp((cx / nn - x0) * sx, 0.0, 65535.0));
                const uint32_t qy = uint32_t(std::clamp((cy / nn - y0) * sy, 0.0, 65535.0));
                m->el_key_h[size_t(e)] = spread(qx) | (spread(qy) << 1);
            }
        }
        upload(m->group_of, group_of.data(), group_of.size(), s);
        const int nq = m->fe.nq;
        m->xq.alloc(size_t(2 * n_elems * nq));
        m->jac.alloc(size_t(n_elems * nq));
        m->D.alloc(size_t(2 * n_elems * nn * nq));
        m->centres.alloc(size_t(2 * n_elems));
        m->el_maxnode.alloc(size_t(n_elems));
        TablesArg ta{m->fe};
        hipLaunchKernelGGL(k_precompute, dim3(blocks_for(n_elems * nq, 256)), dim3(256), 0, s, i64(n_elems), m->xy.p,
                           m->el.p, ta, m->xq.p, m->jac.p, m->D.p);
        hipLaunchKernelGGL(k_centres, dim3(blocks_for(n_elems, 256)), dim3(256), 0, s, i64(n_elems), nn, m->xy.p,
                           m->el.p, ta, m->centres.p);
        hipLaunchKernelGGL(k_elem_maxnode, dim3(blocks_for(n_elems, 256)), dim3(256), 0, s, i64(n_elems), nn, m->el.p,
                           m->el_maxnode.p);
        NLF_HIP(hipGetLastError());
        NLF_HIP(hipStreamSynchronize(s));
        *out = m.release();
    });
}

int nlf_mesh_destroy(nlf_mesh m) {
    if (!m) return NLF_OK;
    CtxScope scope(live_ctx(m->ctx));
    delete m;
    return NLF_OK;
}

int nlf_mesh_info(nlf_mesh m, int64_t* n_nodes, int64_t* n_elems, int* nn, int* nq) {
    if (!m) return NLF_ERR_ARGUMENT;
    if (n_nodes) *n_nodes = m->N;
    if (n_elems) *n_elems = m->E;
    if (nn) *nn = m->nn;
    if (nq) *nq = m->fe.nq;
    return NLF_OK;
}

// Host export of the precompute (parity checks against the oracle's mesh_2d tables).
// The element's quadrature weights [nq] and shape values N_i at its quadrature points [nn][nq]
// (element_2d::weight / qN of the reference's mesh container), host tables.
int nlf_mesh_export_basis(nlf_mesh m, double* weights, double* qN) {
    if (!m) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        const nlf::FeTables& fe = m->fe;
        if (weights)
            for (int q = 0; q < fe.nq; ++q) weights[q] = fe.w[q];
        if (qN)
            for (int k = 0; k < fe.nn * fe.nq; ++k) qN[k] = fe.qN[k];
    });
}

int nlf_mesh_export_quadrature(nlf_mesh m, double* xq, double* jac, double* D, double* centres) {
    if (!m) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        hipStream_t s = m->ctx->stream;
        if (xq) download(xq, m->xq.p, m->xq.n, s);
        if (jac) download(jac, m->jac.p, m->jac.n, s);
        if (D) download(D, m->D.p, m->D.n, s);
        if (centres) download(centres, m->centres.p, m->centres.n, s);
    });
}

}  // extern "C"
