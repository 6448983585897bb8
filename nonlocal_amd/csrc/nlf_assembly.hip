// CSR sparsity pattern and FP64 element-pair assembly of the nonlocal conductivity / stiffness
// matrices (matrix_assembler_2d.hpp:37-189, integrator.hpp:36-91, conductivity_matrix_2d.hpp:106-193,
// stiffness_matrix_2d.hpp:44-199).
//
// Storage: the inner part is kept on the device as the FULL (both-triangle) row pattern, columns
// ascending; the reference's upper-triangular CSR (matrix_separator_base.hpp:50-58) is the tail of
// each row starting at its diagonal.  The bound part (inner rows x Dirichlet columns, and Dirichlet
// rows x other Dirichlet columns) is stored as in the reference.
//
// Traversal: like mesh_run (mesh_2d_utils.hpp:13-37) every node owns its DoF rows; one workgroup
// handles one node, visiting its elements eL and their neighbours eNL.  Values are accumulated in
// LDS in barrier-separated rounds (element slot, neighbour batch, local column index j): within a
// round two contributions never share a position because e -> node(e, j) is injective, so the sum
// order is fixed and the result is bitwise reproducible.  On a mesh where one node holds the same
// local index in two elements the rounds go by the element's rank at the node instead, with the
// same property.  No atomics touch the matrix.
#include "nlf_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <limits>
#include <map>

namespace nlf {

// Model part of a group's parameters: fix_parameters (read_model.hpp:186-192) zeroes the radii of
// LOCAL groups and sets p1 = 1 when a radius vanishes; theory_type (nonlocal_constants.hpp:21-27).
GroupParams model_group_params(const nlf_model& md, bool only_local) {
    GroupParams gp;
    std::memset(&gp, 0, sizeof(gp));
    double p1 = md.local_weight;
    double r0 = md.radius[0], r1 = md.radius[1];
    if (p1 >= 0.999) { r0 = 0; r1 = 0; }
    if (r0 < std::numeric_limits<double>::epsilon() || r1 < std::numeric_limits<double>::epsilon()) p1 = 1.0;
    gp.p1 = p1;
    gp.nonlocal = !only_local && p1 < 0.999;
    if (gp.nonlocal) {
        nlf_model m2 = md;
        m2.radius[0] = r0; m2.radius[1] = r1;
        gp.inf = make_influence(m2);
    }
    return gp;
}

void elastic_group_params(const nlf_elastic_material& e, GroupParams& gp) {
    for (int k = 0; k < 6; ++k) gp.c[k] = 0.0;
    if (e.kind == NLF_ELASTIC_ISOTROPIC) {
        const double value = e.young / (1.0 - e.poisson * e.poisson);           // :70-75
        gp.c[0] = value; gp.c[1] = e.poisson * value; gp.c[2] = 0.5 * value * (1.0 - e.poisson);
        gp.c[3] = value;
    } else {
        NLF_CHECK(e.kind == NLF_ELASTIC_ORTHOTROPIC || e.kind == NLF_ELASTIC_ANISOTROPIC, NLF_ERR_ARGUMENT,
                  "unknown elastic material kind");
        const double value = 1.0 / (1.0 - e.poisson * e.poisson_y);            // :89-95
        const double m11 = e.young * value, m12 = e.young * e.poisson_y * value, m66 = e.shear;
        const double m22 = e.young_y * value;
        if (e.kind == NLF_ELASTIC_ORTHOTROPIC) {
            gp.c[0] = m11; gp.c[1] = m12; gp.c[2] = m66; gp.c[3] = m22;
        } else {                                                                // rotate, :102-123
            const double a2 = e.angle + e.angle;
            const double s2 = std::sin(a2), c2 = std::cos(a2);
            const double c4 = 0.125 * (c2 * c2 - s2 * s2);
            const double diff = m11 - m22, core = m11 + m22 - 2 * m12 - 4 * m66;
            const double plus = core * (c4 + 0.125), minus = core * (0.125 - c4);
            const double cdiff = diff * (c2 + 0.5), minor = core * c2, s4 = 0.25 * s2;
            gp.c[0] = plus + 0.5 * (m12 + m22 + cdiff) + m66;
            gp.c[1] = minus + m12;
            gp.c[2] = minus + m66;
            gp.c[3] = plus + 0.5 * (m12 + m11 - cdiff) + m66;
            gp.c[4] = s4 * (diff + minor);
            gp.c[5] = s4 * (diff - minor);
        }
    }
    NLF_CHECK(e.expansion_kind >= NLF_EXPANSION_NONE && e.expansion_kind <= NLF_EXPANSION_ANISOTROPIC,
              NLF_ERR_ARGUMENT, "unknown thermal expansion kind");
    gp.ekind = e.expansion_kind;
    gp.alpha[0] = e.expansion[0];
    gp.alpha[1] = e.expansion_kind == NLF_EXPANSION_ISOTROPIC ? e.expansion[0] : e.expansion[1];
    gp.alpha[2] = e.expansion_kind == NLF_EXPANSION_ANISOTROPIC ? e.expansion[2] : 0.0;
}

int influence_class(const std::vector<GroupParams>& gp) {
    int poly = -1;
    for (const GroupParams& g : gp) {
        if (!g.nonlocal) continue;
        const InfluenceParams& f = g.inf;
        int cls = 0;
        if (f.dist.kind == DIST_L2 && f.kind == NLF_INFLUENCE_POLYNOMIAL &&
            (f.fast || (f.p_int && f.p_i == 1 && f.q_int && f.q_i <= 2)))
            cls = f.fast ? 1 : f.q_i;
        poly = poly < 0 ? cls : (poly == cls ? poly : 0);
    }
    return poly;
}

DistanceParams make_distance(int kind, int n, double n_real) {
    DistanceParams d;
    if (kind == NLF_DISTANCE_ELLIPSE_WITH_ROTATION) {      // n is fixed at 2 (search_function.hpp:77)
        d.kind = DIST_ROTATION;
        return d;
    }
    NLF_CHECK(kind == NLF_DISTANCE_LP, NLF_ERR_ARGUMENT, "unknown distance kind");
    // read_influence_parameter (read_model.hpp:97-114): n > 0, a number or infinity
    NLF_CHECK(n_real >= 0.0 && std::isfinite(n_real), NLF_ERR_ARGUMENT, "the distance exponent shall be greater than 0");
    if (n_real > 0.0) {
        d.kind = DIST_LN_REAL;
        d.n_real = n_real;
        return d;
    }
    NLF_CHECK(n >= 0 || n == NLF_DISTANCE_N_INFINITY, NLF_ERR_ARGUMENT, "the distance exponent shall be greater than 0");
    if (n == 0 || n == 2) d.kind = DIST_L2;
    else if (n == 1) d.kind = DIST_L1;
    else if (n == NLF_DISTANCE_N_INFINITY) d.kind = DIST_LINF;
    else { d.kind = DIST_LN_INT; d.n = n; }
    return d;
}

// The influence function of a model, selected like read_influence_2d (read_model.hpp:144-165): a
// polynomial with n = p = 2 (integer n) and q = 1 is the fast polynomial.  Normalisations follow
// influence_functions_2d.hpp:44-48 (constant), 70-75 (polynomial), 103-109 (exponential), with n as
// get_value(n) reads it (Infinity<size_t> as a double in the constant's area); the exponent p/n is
// _influence_function_2d::division (:8-17): p itself for the maximum norm, an integer when p and n are
// integers and n divides p, p / n otherwise.  A p or q with an integral value counts as an integer.
InfluenceParams make_influence(const nlf_model& md) {
    InfluenceParams f;
    f.kind = md.influence;
    f.r[0] = md.radius[0]; f.r[1] = md.radius[1];
    f.inv_r[0] = 1.0 / f.r[0]; f.inv_r[1] = 1.0 / f.r[1];
    f.dist = make_distance(md.distance, md.distance_n, md.distance_real);
    const int dk = f.dist.kind;
    const bool infinity = dk == DIST_LINF;
    const bool real_n = dk == DIST_LN_REAL;
    const long n_i = dk == DIST_L1 ? 1 : dk == DIST_LN_INT ? f.dist.n : 2;
    const double n = real_n ? f.dist.n_real : infinity ? 18446744073709551615.0 : double(n_i);
    const double p = md.p, q = md.q;
    auto integral = [](double v) { return v >= 0 && std::floor(v) == v && v < 1e9; };
    auto divide = [&](double v, int& is_int, int& vi, double& vd) {       // division(p, n)
        if (infinity) {
            is_int = integral(v);
            if (is_int) vi = int(v); else vd = v;
        } else if (!real_n && integral(v) && long(v) % n_i == 0) {
            is_int = 1; vi = int(long(v) / n_i);
        } else {
            is_int = 0; vd = v / n;
        }
    };
    if (f.kind == NLF_INFLUENCE_CONSTANT) {
        f.norm = std::tgamma(1 + 2 / n) / (4 * f.r[0] * f.r[1] * std::pow(std::tgamma(1 + 1 / n), 2));
    } else if (f.kind == NLF_INFLUENCE_POLYNOMIAL) {
        if (!real_n && n_i == 2 && !infinity && p == 2.0 && q == 1.0) {
            f.fast = 1;                                  // fast_polynomial (:127-148)
            f.norm = 2.0 / (M_PI * f.r[0] * f.r[1]);
        } else {
            divide(p, f.p_int, f.p_i, f.p_d);
            if (integral(q)) { f.q_int = 1; f.q_i = int(q); } else { f.q_int = 0; f.q_d = q; }
            f.norm = infinity ? p / (8 * f.r[0] * f.r[1] * std::beta(2 / p, q + 1))
                              : n * p / (4 * f.r[0] * f.r[1] * std::beta(1 / n, 1 / n) * std::beta(2 / p, q + 1));
        }
    } else if (f.kind == NLF_INFLUENCE_EXPONENTIAL) {
        divide(p, f.p_int, f.p_i, f.p_d);
        f.exp_q = -q;
        f.norm = infinity ? p * std::pow(q, 2 / p) / (8 * f.r[0] * f.r[1] * std::tgamma(2 / p))
                          : (n * p * std::pow(4.0, 1 / n) * std::pow(q, 2 / p)) /
                                (8 * f.r[0] * f.r[1] * std::tgamma(2 / p) * std::beta(0.5, 1 / n));
    } else {
        throw Error(NLF_ERR_ARGUMENT, "unknown influence function");
    }
    return f;
}


// ---------------------------------------------------------------------------------------------
// Pattern
// ---------------------------------------------------------------------------------------------
constexpr int PATTERN_THREADS = 256;

struct PatternArgs {
    i64 N;
    int dof, nn;
    const i32* el;
    const i32* ne_off;
    const i32* ne_ids;
    const i32* group_of;
    const int* group_nonlocal;
    const i64* nb_off;
    const i32* nb_ids;
    const uint8_t* inner;
    int max_words;
    unsigned int* saved;   // pass 0 stores each node's bitmap here (stride words), pass 1 reads it back
    int stride;
    const i32* lo_e;
    const i32* hi_e;
    // pass 0
    i64* cnt;          // [rows] inner counts
    i64* bcnt;         // [rows] bound counts
    int* overflow;
    // pass 1
    const i64* outer;
    const i64* b_outer;
    i32* cols;
    i32* b_cols;
    i64* diag;
};

template<class F>
__device__ __forceinline__ void for_each_column_element(const PatternArgs& a, i64 node, F&& f) {
    // elements whose nodes are columns of the node's rows: LOCAL eL itself, NONLOCAL all of N(eL)
    for (i32 k = a.ne_off[node]; k < a.ne_off[node + 1]; ++k) {
        const i32 eL = a.ne_ids[k];
        if (a.group_nonlocal[a.group_of[eL]]) {
            const i64 b = a.nb_off[eL], e = a.nb_off[eL + 1];
            for (i64 s = b + threadIdx.x; s < e; s += blockDim.x) f(a.nb_ids[s]);
        } else if (threadIdx.x == 0) {
            f(eL);
        }
    }
}

// exclusive block scan of up to 4 counters
// Exclusive block scan of four counters per thread (v in, exclusive prefix out) and their block
// totals: wave prefix sums by shuffles, then the wave totals (at most 16 waves) through LDS.
__device__ __forceinline__ void block_scan4(i64 v[4], i64 total[4], i64* sh /* >= 4 * waves */) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, W = int(blockDim.x + 63) >> 6;
    i64 excl[4];
    for (int k = 0; k < 4; ++k) {
        i64 x = v[k];
        for (int o = 1; o < 64; o <<= 1) {
            const i64 y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        excl[k] = x - v[k];
        if (lane == 63) sh[k * W + w] = x;
    }
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
        i64 off = 0, tot = 0;
        for (int i = 0; i < W; ++i) {
            const i64 s = sh[k * W + i];
            if (i < w) off += s;
            tot += s;
        }
        v[k] = excl[k] + off;
        total[k] = tot;
    }
    __syncthreads();
}

// Span [lo, hi] of the column node ids reachable from each element (its own nodes when LOCAL,
// the nodes of its neighbour list when NONLOCAL); one wave per element.
__global__ void k_elem_span(PatternArgs a, i64 E, const i32* __restrict__ el_minnode, const i32* __restrict__ el_maxnode,
                            i32* __restrict__ lo_out, i32* __restrict__ hi_out) {
    const i64 e = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (e >= E) return;
    const int lane = threadIdx.x & 63;
    i32 lo = el_minnode[e], hi = el_maxnode[e];
    if (a.group_nonlocal[a.group_of[e]]) {
        for (i64 s = a.nb_off[e] + lane; s < a.nb_off[e + 1]; s += 64) {
            const i32 f = a.nb_ids[s];
            lo = min(lo, el_minnode[f]);
            hi = max(hi, el_maxnode[f]);
        }
        for (int o = 32; o > 0; o >>= 1) { lo = min(lo, __shfl_xor(lo, o)); hi = max(hi, __shfl_xor(hi, o)); }
    }
    if (lane == 0) { lo_out[e] = lo; hi_out[e] = hi; }
}

__global__ void k_node_words(i64 N, const i32* __restrict__ ne_off, const i32* __restrict__ ne_ids,
                             const i32* __restrict__ lo_e, const i32* __restrict__ hi_e, unsigned int* __restrict__ mx) {
    const i64 n = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= N) return;
    i32 lo = 0x7fffffff, hi = -1;
    for (i32 k = ne_off[n]; k < ne_off[n + 1]; ++k) { lo = min(lo, lo_e[ne_ids[k]]); hi = max(hi, hi_e[ne_ids[k]]); }
    if (hi >= lo) atomicMax(mx, unsigned((hi - lo) / 32 + 1));
}

template<int PASS>
__global__ __launch_bounds__(PATTERN_THREADS) void k_pattern(PatternArgs a) {
    extern __shared__ unsigned int bitmap[];
    __shared__ i64 scan[4 * (PATTERN_THREADS + 1)];
    const i64 node = blockIdx.x;
    if (node >= a.N) return;
    i64 lo = a.N, hi = -1;
    for (i32 k = a.ne_off[node]; k < a.ne_off[node + 1]; ++k) {
        lo = min(lo, (i64)a.lo_e[a.ne_ids[k]]);
        hi = max(hi, (i64)a.hi_e[a.ne_ids[k]]);
    }
    if (hi < lo) return;                         // isolated node: no element
    const int words = int((hi - lo) / 32 + 1);
    if (words > a.max_words) {
        if (threadIdx.x == 0) atomicOr(a.overflow, 1);
        return;
    }
    if (PASS == 1 && a.saved) {
        for (int w = threadIdx.x; w < words; w += blockDim.x) bitmap[w] = a.saved[node * a.stride + w];
    } else {
        for (int w = threadIdx.x; w < words; w += blockDim.x) bitmap[w] = 0u;
        __syncthreads();
        for_each_column_element(a, node, [&](i32 e) {
            // the element's node ids in 16-byte loads (rows of el are 16-byte aligned: nn = 4 or 8)
            const int4* row = reinterpret_cast<const int4*>(a.el + i64(e) * a.nn);
            for (int q = 0; q < a.nn / 4; ++q) {
                const int4 ids = row[q];
                const i32 id[4] = {ids.x, ids.y, ids.z, ids.w};
                for (int j = 0; j < 4; ++j) {
                    const i64 v = id[j] - lo;
                    atomicOr(&bitmap[v >> 5], 1u << (v & 31));
                }
            }
        });
        __syncthreads();
        if (PASS == 0 && a.saved)
            for (int w = threadIdx.x; w < words; w += blockDim.x) a.saved[node * a.stride + w] = bitmap[w];
    }
    __syncthreads();
    // The set bits (column nodes, ascending) are dealt out evenly: thread t takes ranks
    // [t k, t k + k) of the L set bits.  A prefix of the bits per thread-owned run of words locates
    // the first rank; the walk then goes bit by bit.  (The node's columns cluster in a few runs of
    // words, so dealing out words instead would leave most threads idle.)
    __shared__ int tpre[PATTERN_THREADS + 1];
    const int per = (words + blockDim.x - 1) / blockDim.x;
    const int w0 = min(words, int(threadIdx.x) * per), w1 = min(words, w0 + per);
    {
        i64 c4[4] = {0, 0, 0, 0}, tot4[4];
        for (int w = w0; w < w1; ++w) c4[0] += __popc(bitmap[w]);
        block_scan4(c4, tot4, scan);
        tpre[threadIdx.x] = int(c4[0]);
        if (threadIdx.x == 0) tpre[blockDim.x] = int(tot4[0]);
    }
    __syncthreads();
    const int L = tpre[blockDim.x];
    const int k = (L + int(blockDim.x) - 1) / int(blockDim.x);
    const int r0 = min(L, int(threadIdx.x) * k), r1 = min(L, r0 + k);
    int wstart = 0;
    unsigned int bstart = 0u;
    if (r0 < r1) {
        int u = 0, hi_u = int(blockDim.x) - 1;                 // last thread whose prefix is <= r0
        while (u < hi_u) {
            const int mid = (u + hi_u + 1) >> 1;
            if (tpre[mid] <= r0) u = mid; else hi_u = mid - 1;
        }
        int rank = tpre[u];
        wstart = u * per;
        while (rank + __popc(bitmap[wstart]) <= r0) rank += __popc(bitmap[wstart++]);
        bstart = bitmap[wstart];
        for (int skip = r0 - rank; skip > 0; --skip) bstart &= bstart - 1;
    }
    auto walk = [&](auto&& f) {                              // column nodes of ranks [r0, r1), ascending
        int w = wstart;
        unsigned int bits = bstart;
        for (int r = r0; r < r1; ++r) {
            while (!bits) bits = bitmap[++w];
            const int b = __ffs(bits) - 1;
            bits &= bits - 1;
            f(lo + i64(w) * 32 + b);
        }
    };
    const int D = a.dof;
    i64 rowv[2], rowinner[2];
    for (int d = 0; d < D; ++d) { rowv[d] = i64(D) * node + d; rowinner[d] = a.inner[rowv[d]]; }
    i64 cnt[4] = {0, 0, 0, 0};   // inner d0, inner d1, bound d0, bound d1
    walk([&](i64 m) {
        for (int dc = 0; dc < D; ++dc) {
            const i64 c = i64(D) * m + dc;
            const bool ci = a.inner[c];
            for (int d = 0; d < D; ++d) {
                if (ci) { if (rowinner[d]) cnt[d]++; }
                else if (c != rowv[d]) cnt[2 + d]++;
            }
        }
    });
    if (PASS == 0) {
        i64 tot[4];
        block_scan4(cnt, tot, scan);
        if (threadIdx.x == 0)
            for (int d = 0; d < D; ++d) {
                a.cnt[rowv[d]] = rowinner[d] ? tot[d] : 1;      // Dirichlet rows: the diagonal only
                a.bcnt[rowv[d]] = tot[2 + d];
            }
        return;
    }
    i64 tot[4];
    block_scan4(cnt, tot, scan);
    i64 pos[4];
    for (int d = 0; d < D; ++d) { pos[d] = a.outer[rowv[d]] + cnt[d]; pos[2 + d] = a.b_outer[rowv[d]] + cnt[2 + d]; }
    walk([&](i64 m) {
        for (int dc = 0; dc < D; ++dc) {
            const i64 c = i64(D) * m + dc;
            const bool ci = a.inner[c];
            for (int d = 0; d < D; ++d) {
                if (ci) {
                    if (rowinner[d]) {
                        if (c == rowv[d]) a.diag[rowv[d]] = pos[d];
                        a.cols[pos[d]++] = i32(c);
                    }
                } else if (c != rowv[d]) {
                    a.b_cols[pos[2 + d]++] = i32(c);
                }
            }
        }
    });
    if (threadIdx.x == 0)
        for (int d = 0; d < D; ++d)
            if (!rowinner[d]) { a.cols[a.outer[rowv[d]]] = i32(rowv[d]); a.diag[rowv[d]] = a.outer[rowv[d]]; }
}

// Pattern of a matrix with no nonlocal group (assemble_part::LOCAL, or every group LOCAL): a row's
// column nodes are the nodes of the node's own elements, at most PL_MAX of them, so one thread per
// node gathers and sorts them (no bitmap).  Same classification and order as k_pattern: column nodes
// ascending, then the DoF; inner columns of inner rows, Dirichlet rows keep their diagonal only,
// bound columns of every row.  A node with more column nodes sets *overflow (k_pattern then builds).
constexpr int PL_MAX = 64;
constexpr int PL_THREADS = 64;

template<int PASS>
__global__ __launch_bounds__(PL_THREADS) void k_pattern_local(PatternArgs a) {
    __shared__ i32 buf[PL_THREADS][PL_MAX + 1];
    const i64 node = i64(blockIdx.x) * PL_THREADS + threadIdx.x;
    if (node >= a.N) return;
    i32* ids = buf[threadIdx.x];
    int n = 0;
    bool over = false;
    for (i32 k = a.ne_off[node]; k < a.ne_off[node + 1] && !over; ++k) {
        const i32 e = a.ne_ids[k];
        for (int j = 0; j < a.nn; ++j) {
            const i32 v = a.el[i64(e) * a.nn + j];
            int pos = n;                                   // insertion into the sorted unique list
            while (pos > 0 && ids[pos - 1] > v) --pos;
            if (pos > 0 && ids[pos - 1] == v) continue;
            if (n == PL_MAX) { over = true; break; }
            for (int t = n; t > pos; --t) ids[t] = ids[t - 1];
            ids[pos] = v;
            ++n;
        }
    }
    if (over) {
        atomicOr(a.overflow, 1);
        return;
    }
    const int D = a.dof;
    i64 rowv[2];
    bool rowinner[2];
    for (int d = 0; d < D; ++d) { rowv[d] = i64(D) * node + d; rowinner[d] = a.inner[rowv[d]]; }
    if (PASS == 0) {
        i64 cnt[4] = {0, 0, 0, 0};
        for (int t = 0; t < n; ++t)
            for (int dc = 0; dc < D; ++dc) {
                const i64 c = i64(D) * ids[t] + dc;
                const bool ci = a.inner[c];
                for (int d = 0; d < D; ++d) {
                    if (ci) { if (rowinner[d]) cnt[d]++; }
                    else if (c != rowv[d]) cnt[2 + d]++;
                }
            }
        for (int d = 0; d < D; ++d) {
            a.cnt[rowv[d]] = rowinner[d] ? cnt[d] : 1;
            a.bcnt[rowv[d]] = cnt[2 + d];
        }
        return;
    }
    i64 pos[4];
    for (int d = 0; d < D; ++d) { pos[d] = a.outer[rowv[d]]; pos[2 + d] = a.b_outer[rowv[d]]; }
    for (int t = 0; t < n; ++t)
        for (int dc = 0; dc < D; ++dc) {
            const i64 c = i64(D) * ids[t] + dc;
            const bool ci = a.inner[c];
            for (int d = 0; d < D; ++d) {
                if (ci) {
                    if (rowinner[d]) {
                        if (c == rowv[d]) a.diag[rowv[d]] = pos[d];
                        a.cols[pos[d]++] = i32(c);
                    }
                } else if (c != rowv[d]) {
                    a.b_cols[pos[2 + d]++] = i32(c);
                }
            }
        }
    for (int d = 0; d < D; ++d)
        if (!rowinner[d]) { a.cols[a.outer[rowv[d]]] = i32(rowv[d]); a.diag[rowv[d]] = a.outer[rowv[d]]; }
}

// ---------------------------------------------------------------------------------------------
// Values
// ---------------------------------------------------------------------------------------------
constexpr int ASM_THREADS = 256;

struct AssemblyArgs {
    i64 N;
    const i32* el;
    const i32* ne_off;
    const i32* ne_ids;
    const i32* group_of;
    const GroupParams* groups;
    const i64* nb_off;
    const i32* nb_ids;
    const i32* el_maxnode;
    const uint8_t* el_dirichlet;
    // per neighbour-list entry: the neighbour's largest node id, bit 31 set when it has a Dirichlet
    // node (the pair path's activity test read with the list, not gathered through it), or null
    const i32* nb_pack;
    const uint8_t* inner;
    const double* xq;
    const double* jac;
    const double* D;
    double w[MAX_NQ];
    const i64* outer;
    const i32* cols;
    const i64* diag;
    const i64* b_outer;
    const i32* b_cols;
    double* val;
    double* b_val;
    float* val32;        // FP32 value pass: inner values only (the bound part stays FP64)
    int local_only;      // assemble_part::LOCAL: every group treated as LOCAL
    const double* pairs; // thermal Q4: nonlocal pair blocks from k_pairs_mfma (null = compute in place)
    const float* pairs32;
    i64 plane;           // pairs[i * plane + (eoff[eL] + t) * 4 + j], t = position in eL's neighbour list
    const i64* eoff;
    i64 node0;           // nodes [node0, node0 + grid) of this launch
    int mass;            // heat capacity matrix: rho c int N_i N_j (groups' c[0] = rho c)
    const double* cq;        // coefficients per quadrature point [(e * NQ + q) * 6 + k] of groups with varying
    const uint8_t* el_rank;  // non-injective meshes: rank of element e in its node j's element list
    int rounds;              // and the number of ranks (most elements at one node)
    double qN[MAX_NN * MAX_NQ];
};

template<class T>
__device__ __forceinline__ T phi_of(const InfluenceParams& f, double ax, double ay, double bx, double by) {
    if constexpr (sizeof(T) == sizeof(double)) return influence(f, ax, ay, bx, by);
    else return influence32(f, ax, ay, bx, by);
}

__device__ __forceinline__ int lower_bound_i32(const i32* a, int n, i32 key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Local integrand entry (i, j) of element e: K^L_ij (thermal) or its 2x2 block (elastic), scaled by
// p1.  conductivity_matrix_2d.hpp:106-130, stiffness_matrix_2d.hpp:44-80.  T = float is the FP32
// value pass: w_q / |det J| is formed in FP64, the products and sums in T.  The sum over the
// quadrature points runs in q order, whichever caller asks for which j.
template<int DOF, int NN, int NQ, class T>
__device__ void local_entry(const AssemblyArgs& a, const GroupParams& g, i64 e, int i, int j, T (&v)[DOF * DOF]) {
    for (int k = 0; k < DOF * DOF; ++k) v[k] = T(0);
    if (DOF == 1 && a.mass) {
        // heat_capacity_matrix_2d::integrate_basic_pair (heat_capacity_matrix_2d.hpp:31-37), times
        // density * capacity (:55-56)
        for (int q = 0; q < NQ; ++q) {
            const T f = T(a.w[q] * a.qN[i * NQ + q]);
            const T jq = T(a.jac[e * NQ + q]);
            v[0] += f * T(a.qN[j * NQ + q]) * jq;
        }
        v[0] *= T(g.c[0]);
        return;
    }
    for (int q = 0; q < NQ; ++q) {
        // the coefficient at the quadrature point: the group's constant, or its value per point
        // (evaluated_parameters' Variable alternative, conductivity_matrix_2d.hpp:117-118)
        double cp[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) cp[k] = g.varying ? a.cq[(e * NQ + q) * 6 + k] : g.c[k];
        const T c0 = T(cp[0]), c1 = T(cp[1]), c2 = T(cp[2]);
        const T f = T(a.w[q] / a.jac[e * NQ + q]);
        const T ix = T(a.D[2 * ((e * NN + i) * NQ + q)]), iy = T(a.D[2 * ((e * NN + i) * NQ + q) + 1]);
        const T jx = T(a.D[2 * ((e * NN + j) * NQ + q)]), jy = T(a.D[2 * ((e * NN + j) * NQ + q) + 1]);
        if (DOF == 1) {
            T value;
            if (g.ckind == NLF_CONDUCTIVITY_ISOTROPIC) value = c0 * (ix * jx + iy * jy);
            else if (g.ckind == NLF_CONDUCTIVITY_ORTHOTROPIC) value = c0 * ix * jx + c1 * iy * jy;
            else value = ix * (c0 * jx + c2 * jy) + iy * (c2 * jx + c1 * jy);
            v[0] += f * value;
        } else {
            // B_i^T C B_j with C = [[h11, h12, h16], [h12, h22, h26], [h16, h26, h66]] (the
            // anisotropic integrand, stiffness_matrix_2d.hpp:66-75; isotropic and orthotropic
            // matrices have h16 = h26 = 0)
            const T h11 = c0, h12 = c1, h66 = c2, h22 = T(cp[3]), h16 = T(cp[4]), h26 = T(cp[5]);
            const T wx = f * ix, wy = f * iy;
            const T s1 = h66 * wy + h16 * wx, s2 = h66 * wx + h26 * wy;
            v[0] += (h11 * wx + h16 * wy) * jx + s1 * jy;
            v[1] += (h12 * wx + h26 * wy) * jy + s1 * jx;
            v[2] += (h12 * wy + h16 * wx) * jx + s2 * jy;
            v[DOF * DOF - 1] += (h22 * wy + h26 * wx) * jy + s2 * jx;
        }
    }
    const T p1 = T(g.p1);
    for (int k = 0; k < DOF * DOF; ++k) v[k] *= p1;
}

// The local row i of element e: local_entry for every j.
template<int DOF, int NN, int NQ, class T>
__device__ void local_row(const AssemblyArgs& a, const GroupParams& g, i64 e, int i, T (&v)[NN][DOF * DOF]) {
    for (int j = 0; j < NN; ++j) local_entry<DOF, NN, NQ, T>(a, g, e, i, j, v[j]);
}

// Nonlocal integrand row iL of pair (eL, eNL):
//   u(qNL) = sum_qL phi(x_qL, x_qNL) w_qL D_i(qL)            (B_e^T Phi contraction)
//   v_j    = (1 - p1) sum_qNL w_qNL u(qNL) . C D_j(qNL)       (... x C B_e')
// conductivity_matrix_2d.hpp:132-164, stiffness_matrix_2d.hpp:82-123 (C taken at the nonlocal point).
// T = float: coordinates stay FP64 (their differences are rounded once), phi and the contractions
// run in FP32.
template<int DOF, int NN, int NQ, class T>
__device__ void nonlocal_row(const AssemblyArgs& a, const GroupParams& g, const double* xl /*2*NQ*/,
                             const T* hl /*2*NQ: w_qL D_i(qL)*/, i64 eNL, T (&v)[NN][DOF * DOF]) {
    for (int j = 0; j < NN; ++j)
        for (int k = 0; k < DOF * DOF; ++k) v[j][k] = T(0);
    const T c0 = T(g.c[0]), c1 = T(g.c[1]), c2 = T(g.c[2]);
    for (int qn = 0; qn < NQ; ++qn) {
        const double bx = a.xq[2 * (eNL * NQ + qn)], by = a.xq[2 * (eNL * NQ + qn) + 1];
        T ux = T(0), uy = T(0);
        for (int ql = 0; ql < NQ; ++ql) {
            const T phi = phi_of<T>(g.inf, xl[2 * ql], xl[2 * ql + 1], bx, by);
            ux += phi * hl[2 * ql];
            uy += phi * hl[2 * ql + 1];
        }
        ux *= T(a.w[qn]);
        uy *= T(a.w[qn]);
        for (int j = 0; j < NN; ++j) {
            const T jx = T(a.D[2 * ((eNL * NN + j) * NQ + qn)]), jy = T(a.D[2 * ((eNL * NN + j) * NQ + qn) + 1]);
            if (DOF == 1) {
                if (g.ckind == NLF_CONDUCTIVITY_ISOTROPIC) v[j][0] += c0 * (ux * jx + uy * jy);
                else if (g.ckind == NLF_CONDUCTIVITY_ORTHOTROPIC) v[j][0] += c0 * ux * jx + c1 * uy * jy;
                else v[j][0] += ux * (c0 * jx + c2 * jy) + uy * (c2 * jx + c1 * jy);
            } else {                             // stiffness_matrix_2d.hpp:112-117, any Hooke kind
                const T h11 = c0, h12 = c1, h66 = c2, h22 = T(g.c[3]), h16 = T(g.c[4]), h26 = T(g.c[5]);
                v[j][0] += (h11 * ux + h16 * uy) * jx + (h16 * ux + h66 * uy) * jy;
                v[j][1] += (h16 * ux + h66 * uy) * jx + (h12 * ux + h26 * uy) * jy;
                v[j][2] += (h16 * ux + h12 * uy) * jx + (h66 * ux + h26 * uy) * jy;
                v[j][DOF * DOF - 1] += (h66 * ux + h26 * uy) * jx + (h26 * ux + h22 * uy) * jy;
            }
        }
    }
    const T s = T(1.0 - g.p1);
    for (int j = 0; j < NN; ++j)
        for (int k = 0; k < DOF * DOF; ++k) v[j][k] *= s;
}

// ---- MFMA element-pair contractions (thermal Q4) -------------------------------------------------
// For a pair (eL, eNL) of one group the nonlocal block is
//   K_ij = (1 - p1) sum_k sum_qn sum_ql Phi[ql][qn] HL_k[i][ql] HN_k[j][qn]
// with HL_k[i][q] = w_q D_i(q)_k and HN_k[j][q] = w_q (C D_j(q))_k (conductivity_matrix_2d.hpp:132-164,
// C taken at the nonlocal point).  V_FMA_F64_4X4X4F64 computes four independent 4x4x4 products per
// wave; lane l = 16 h + 4 g + c belongs to block g, and (measured on gfx950) holds
//   A_g[c][h], B_g[h][c], D_g[h][c].
// One wave handles one element eL and four neighbours per step (block g = neighbour slot):
//   V_k = Phi^T HL_k^T           A = Phi^T  (lane: Phi[ql = h][qn = c], one influence value per lane)
//                                B = HL_k^T (lane: HL_k[i = c][ql = h], fixed per eL)
//   K  += V_k^T-as-A . HN_k^T    D of the first product is exactly the A operand of the second
//                                (A[i = c][qn = h] = V[qn = h][i = c]); B lane: HN_k[j = c][qn = h]
// so a pair costs four MFMA instructions per four pairs and no lane permutes.  Blocks are written
// to pairs[i * plane + p * 4 + j]; the node-centric assembly reads row i of them.
template<class T>
struct PairArgs {
    i64 count;                 // elements of this chunk: list[0 .. count)
    const i32* list;           // null: element li itself
    const i64* eoff;           // chunk slot of each listed element's first neighbour pair
    const i64* nb_off;
    const i32* nb_ids;
    const i32* group_of;
    const GroupParams* groups;
    const double* xq;          // [e][q][2]
    const T* HL;               // [e][k][i][q], 32 per element
    const T* HN;
    const i32* el_min;
    const i32* el_max;
    const uint8_t* el_dir;
    T* pairs;
    i64 plane;
};

struct Weights4 {
    double w[4];
};

template<class T>
__global__ void k_pair_prep(i64 E, const double* __restrict__ D, Weights4 w4, const i32* __restrict__ group_of,
                            const GroupParams* __restrict__ groups, T* __restrict__ HL, T* __restrict__ HN) {
    const i64 t = i64(blockIdx.x) * blockDim.x + threadIdx.x;      // (e, i, q)
    if (t >= E * 16) return;
    const i64 e = t / 16;
    const int i = int(t % 16) / 4, q = int(t % 4);
    const GroupParams g = groups[group_of[e]];
    const double w = w4.w[q];
    const double dx = D[2 * ((e * 4 + i) * 4 + q)], dy = D[2 * ((e * 4 + i) * 4 + q) + 1];
    double cx, cy;
    if (g.ckind == NLF_CONDUCTIVITY_ISOTROPIC) { cx = g.c[0] * dx; cy = g.c[0] * dy; }
    else if (g.ckind == NLF_CONDUCTIVITY_ORTHOTROPIC) { cx = g.c[0] * dx; cy = g.c[1] * dy; }
    else { cx = g.c[0] * dx + g.c[2] * dy; cy = g.c[2] * dx + g.c[1] * dy; }
    HL[e * 32 + i * 4 + q] = T(w * dx);
    HL[e * 32 + 16 + i * 4 + q] = T(w * dy);
    HN[e * 32 + i * 4 + q] = T(w * cx);
    HN[e * 32 + 16 + i * 4 + q] = T(w * cy);
}

// A wave first reads 64 neighbour ids and their "needed" flags in one round, then takes the runs
// of four consecutive neighbours that hold a needed one, four runs at a time (four MFMA sets of four
// blocks), so each round of gathers serves sixteen pairs.  Every block of a run is computed and
// stored: with an element's first slot a multiple of four (padded offsets) a run fills whole
// 128-byte lines of each row plane, so no line is written in part.  A pair's block does not depend
// on which set or block slot it lands in.  POLY: influence_poly.
template<int POLY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_pairs_mfma(PairArgs<double> a) {
    constexpr int SETS = 4;
    const int l = threadIdx.x & 63;
    const int g = (l >> 2) & 3, c = l & 3, h = l >> 4;
    const i64 waves = i64(gridDim.x) * (blockDim.x >> 6);
    for (i64 li = i64(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); li < a.count; li += waves) {
        const i64 eL = a.list ? a.list[li] : li;
        const i64 slot = a.eoff[eL];
        const GroupParams gp = a.groups[a.group_of[eL]];          // one group per wave
        if (!gp.nonlocal) continue;
        const double hl0 = a.HL[eL * 32 + c * 4 + h];             // HL_0[i = c][ql = h]
        const double hl1 = a.HL[eL * 32 + 16 + c * 4 + h];
        const double xlx = a.xq[2 * (eL * 4 + h)], xly = a.xq[2 * (eL * 4 + h) + 1];
        const i64 nb_b = a.nb_off[eL], cnt = a.nb_off[eL + 1] - nb_b;
        const i32 emin = a.el_min[eL];
        const double s = 1.0 - gp.p1;
        // the ids of the next batch of 64 are loaded while the current batch is contracted
        i32 id_next = l < cnt ? a.nb_ids[nb_b + l] : 0;
        for (i64 b64 = 0; b64 < cnt; b64 += 64) {
            const i32 id = id_next;
            // the pair reaches the upper triangle of some row of eL, or a Dirichlet column
            const bool nd = b64 + l < cnt && (a.el_max[id] >= emin || a.el_dir[id]);
            if (b64 + 64 + l < cnt) id_next = a.nb_ids[nb_b + b64 + 64 + l];
            unsigned long long runs = __ballot(nd);
            runs |= runs >> 1;
            runs |= runs >> 2;
            runs &= 0x1111111111111111ull;                       // bit 4q: run q holds a needed pair
            while (runs) {
                int pos[SETS];                                    // this block's neighbour in each set
                for (int u = 0; u < SETS; ++u) {
                    pos[u] = runs ? int(__builtin_ctzll(runs)) + g : -1;
                    runs &= runs - 1;
                    if (pos[u] >= 0 && b64 + pos[u] >= cnt) pos[u] = -1;
                }
                double xnx[SETS], xny[SETS], hn0[SETS], hn1[SETS];
#pragma unroll
                for (int u = 0; u < SETS; ++u) {
                    const i64 eNL = __shfl(id, pos[u] < 0 ? 0 : pos[u]);
                    xnx[u] = xny[u] = hn0[u] = hn1[u] = 0.0;
                    if (pos[u] >= 0) {
                        xnx[u] = a.xq[2 * (eNL * 4 + c)];
                        xny[u] = a.xq[2 * (eNL * 4 + c) + 1];
                        hn0[u] = a.HN[eNL * 32 + c * 4 + h];      // HN_0[j = c][qn = h]
                        hn1[u] = a.HN[eNL * 32 + 16 + c * 4 + h];
                    }
                }
#pragma unroll
                for (int u = 0; u < SETS; ++u) {
                    const double phi = pos[u] >= 0 ? influence_poly<POLY>(gp.inf, xlx, xly, xnx[u], xny[u]) : 0.0;
                    const double v0 = __builtin_amdgcn_mfma_f64_4x4x4f64(phi, hl0, 0.0, 0, 0, 0);
                    const double v1 = __builtin_amdgcn_mfma_f64_4x4x4f64(phi, hl1, 0.0, 0, 0, 0);
                    double k = __builtin_amdgcn_mfma_f64_4x4x4f64(v0, hn0[u], 0.0, 0, 0, 0);
                    k = __builtin_amdgcn_mfma_f64_4x4x4f64(v1, hn1[u], k, 0, 0, 0);
                    if (pos[u] >= 0) a.pairs[i64(h) * a.plane + (slot + b64 + pos[u]) * 4 + c] = s * k;  // K[i = h][j = c]
                }
            }
        }
    }
}

// FP32 pair blocks with V_MFMA_F32_4X4X1_16B_F32: sixteen independent rank-1 4x4 updates per
// wave; lane l = 4 b + c belongs to block b and (measured on gfx950) holds A_b[c][0], B_b[0][c]
// and D_b[i][c] in accumulator register i.  One wave handles one element eL and sixteen
// neighbours per step (block b = neighbour slot):
//   V_k[qn][i] = sum_ql Phi[ql][qn] HL_k[i][ql]   four rank-1 steps over ql:
//                A lane: Phi[ql][qn = c], B lane: HL_k[i = c][ql]; D: lane 4b + i, register qn
//   K[i][j]   += sum_k sum_qn V_k[qn][i] HN_k[j][qn]   eight rank-1 steps over (k, qn):
//                A lane 4b + i: register qn of V_k (already in place), B lane: HN_k[j = c][qn]
// Sixteen instructions per sixteen pairs, every product of the blocks used; FP32 accumulation.
__global__ __launch_bounds__(256) void k_pairs_mfma32(PairArgs<float> a) {
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    const int l = threadIdx.x & 63;
    const int b = l >> 2, c = l & 3;
    const i64 waves = i64(gridDim.x) * (blockDim.x >> 6);
    for (i64 li = i64(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); li < a.count; li += waves) {
        const i64 eL = a.list ? a.list[li] : li;
        const i64 slot = a.eoff[eL];
        const GroupParams gp = a.groups[a.group_of[eL]];          // one group per wave
        if (!gp.nonlocal) continue;
        const f32x4 hl0 = *reinterpret_cast<const f32x4*>(a.HL + eL * 32 + c * 4);       // HL_0[i = c][ql]
        const f32x4 hl1 = *reinterpret_cast<const f32x4*>(a.HL + eL * 32 + 16 + c * 4);
        double xl[8];
        for (int q = 0; q < 8; ++q) xl[q] = a.xq[eL * 8 + q];
        const i64 nb_b = a.nb_off[eL], cnt = a.nb_off[eL + 1] - nb_b;
        const i32 emin = a.el_min[eL];
        const float s = float(1.0 - gp.p1);
        for (i64 base = 0; base < cnt; base += 16) {
            const i64 t = base + b;
            bool need = false;
            i64 eNL = eL;
            if (t < cnt) {
                eNL = a.nb_ids[nb_b + t];
                need = a.el_max[eNL] >= emin || a.el_dir[eNL];
            }
            float phi[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            f32x4 hn0 = {0.0f, 0.0f, 0.0f, 0.0f}, hn1 = {0.0f, 0.0f, 0.0f, 0.0f};
            if (need) {
                const double bx = a.xq[2 * (eNL * 4 + c)], by = a.xq[2 * (eNL * 4 + c) + 1];
                for (int ql = 0; ql < 4; ++ql) phi[ql] = influence32(gp.inf, xl[2 * ql], xl[2 * ql + 1], bx, by);
                hn0 = *reinterpret_cast<const f32x4*>(a.HN + eNL * 32 + c * 4);            // HN_0[j = c][qn]
                hn1 = *reinterpret_cast<const f32x4*>(a.HN + eNL * 32 + 16 + c * 4);
            }
            const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
            f32x4 v0 = zero, v1 = zero;
            for (int ql = 0; ql < 4; ++ql) {
                v0 = __builtin_amdgcn_mfma_f32_4x4x1f32(phi[ql], hl0[ql], v0, 0, 0, 0);
                v1 = __builtin_amdgcn_mfma_f32_4x4x1f32(phi[ql], hl1[ql], v1, 0, 0, 0);
            }
            f32x4 k = zero;
            for (int qn = 0; qn < 4; ++qn) {
                k = __builtin_amdgcn_mfma_f32_4x4x1f32(v0[qn], hn0[qn], k, 0, 0, 0);
                k = __builtin_amdgcn_mfma_f32_4x4x1f32(v1[qn], hn1[qn], k, 0, 0, 0);
            }
            if (need)
                for (int i = 0; i < 4; ++i) a.pairs[i64(i) * a.plane + (slot + t) * 4 + c] = s * k[i];  // K[i][j]
        }
    }
}

// T = double: the FP64 value pass (inner and bound parts).  T = float: the FP32 value pass with
// FP32 accumulation into val32 (inner part only).
// Activity test of the pair path from a per-entry pack (1) or gathered through the neighbour ids (0):
// 70.8 against 71.8 ms per 1000 x 1000 assembly, the pack's build (1.3 ms) included.
#ifndef NLF_ASM_NB_PACK
#define NLF_ASM_NB_PACK 1
#endif

// Neighbours per thread and step where the rows come from the pair blocks (2: measured against 3
// and 4, which were slower).
#ifndef NLF_ASM_PAIR_ITEMS
#define NLF_ASM_PAIR_ITEMS 2
#endif

template<int DOF, int NN, int NQ, bool INJECTIVE, class T, int THREADS = ASM_THREADS, bool PAIRS = false>
__global__ __launch_bounds__(THREADS) void k_assemble(AssemblyArgs a) {
    constexpr bool F64 = sizeof(T) == sizeof(double);
    extern __shared__ unsigned char smem[];
    __shared__ double xl[2 * NQ];
    __shared__ T hl[2 * NQ];
    __shared__ int row_len[DOF], brow_len[DOF];
    __shared__ int s_off[2 * DOF + 1];
    __shared__ int s_woff[DOF + 1];
    const i64 node = a.node0 + blockIdx.x;
    if (node >= a.N) return;
    i64 rows[DOF];
    for (int d = 0; d < DOF; ++d) rows[d] = i64(DOF) * node + d;
    if (threadIdx.x == 0) {
        int off = 0;
        for (int d = 0; d < DOF; ++d) {
            // upper part only: [diag, end) of the full row
            row_len[d] = int(a.outer[rows[d] + 1] - a.diag[rows[d]]);
            brow_len[d] = int(a.b_outer[rows[d] + 1] - a.b_outer[rows[d]]);
        }
        for (int d = 0; d < DOF; ++d) { s_off[d] = off; off += row_len[d]; }
        for (int d = 0; d < DOF; ++d) { s_off[DOF + d] = off; off += brow_len[d]; }
        s_off[2 * DOF] = off;
        // words of the bitmap of each upper row over its column span [row, last column]
        int woff = 0;
        for (int d = 0; d < DOF; ++d) {
            s_woff[d] = woff;
            const i64 span = a.cols[a.outer[rows[d] + 1] - 1] - rows[d];
            woff += int(span / 32) + 1;
        }
        s_woff[DOF] = woff;
    }
    __syncthreads();
    const int total = s_off[2 * DOF];
    T* acc = reinterpret_cast<T*>(smem);
    i32* cl = reinterpret_cast<i32*>(acc + total);
    // position of an upper column c of row d: bit (c - row) of the row's bitmap, and the row
    // position = the first position in its word (wpos) + the set bits below it in the word
    unsigned int* bm = reinterpret_cast<unsigned int*>(cl + total);
    uint16_t* wpos = reinterpret_cast<uint16_t*>(bm + s_woff[DOF]);
    for (int w = threadIdx.x; w < s_woff[DOF]; w += blockDim.x) bm[w] = 0u;
    for (int d = 0; d < DOF; ++d) {
        for (int k = threadIdx.x; k < row_len[d]; k += blockDim.x) {
            cl[s_off[d] + k] = a.cols[a.diag[rows[d]] + k];
            acc[s_off[d] + k] = T(0);
        }
        for (int k = threadIdx.x; k < brow_len[d]; k += blockDim.x) {
            cl[s_off[DOF + d] + k] = a.b_cols[a.b_outer[rows[d]] + k];
            acc[s_off[DOF + d] + k] = T(0);
        }
    }
    bool rinner[DOF];
    for (int d = 0; d < DOF; ++d) rinner[d] = a.inner[rows[d]];
    __syncthreads();
    for (int d = 0; d < DOF; ++d)
        for (int k = threadIdx.x; k < row_len[d]; k += blockDim.x) {
            const int off = int(cl[s_off[d] + k] - rows[d]);
            atomicOr(&bm[s_woff[d] + (off >> 5)], 1u << (off & 31));
            if (k == 0 || ((cl[s_off[d] + k - 1] - rows[d]) >> 5) != (off >> 5)) wpos[s_woff[d] + (off >> 5)] = uint16_t(k);
        }
    __syncthreads();
    for (i32 k = a.ne_off[node]; k < a.ne_off[node + 1]; ++k) {
        const i32 eL = a.ne_ids[k];
        int iL = 0;
        for (int i = 0; i < NN; ++i) if (a.el[i64(eL) * NN + i] == node) iL = i;
        const GroupParams g = a.groups[a.group_of[eL]];
        const bool nonlocal = g.nonlocal && !a.local_only;
        if (threadIdx.x < NQ) {
            const int q = threadIdx.x;
            xl[2 * q] = a.xq[2 * (i64(eL) * NQ + q)];
            xl[2 * q + 1] = a.xq[2 * (i64(eL) * NQ + q) + 1];
            hl[2 * q] = T(a.w[q] * a.D[2 * ((i64(eL) * NN + iL) * NQ + q)]);
            hl[2 * q + 1] = T(a.w[q] * a.D[2 * ((i64(eL) * NN + iL) * NQ + q) + 1]);
        }
        __syncthreads();
        const i64 nb_b = nonlocal ? a.nb_off[eL] : 0;
        const i64 cnt = nonlocal ? a.nb_off[eL + 1] - nb_b : 1;
        // ITEMS neighbours per thread and step: item u covers [base + u blockDim, base + (u + 1) blockDim),
        // their loads are issued before either is used, and they are added item after item, so the
        // order of the additions is that of ITEMS consecutive single-item steps
        // (PAIRS: the nonlocal rows come from the pair blocks of k_pairs_mfma, thermal Q4 only)
        constexpr int ITEMS = PAIRS ? NLF_ASM_PAIR_ITEMS : 1;
        const T* pairs = nullptr;
        if constexpr (F64) pairs = a.pairs; else pairs = a.pairs32;
        const bool from_pairs = PAIRS && nonlocal;
        const i64 eoff = from_pairs ? a.eoff[eL] : 0;
        for (i64 base = 0; base < cnt; base += ITEMS * i64(blockDim.x)) {
            T v[ITEMS][NN][DOF * DOF];
            int pos[ITEMS][NN][DOF * DOF];
            bool active[ITEMS];
            i64 eNL[ITEMS];
            i32 mj[ITEMS][NN];
            i32 pk[ITEMS];
            const bool packed = nonlocal && a.nb_pack != nullptr;
#pragma unroll
            for (int u = 0; u < ITEMS; ++u) {
                const i64 t = base + u * i64(blockDim.x) + threadIdx.x;
                eNL[u] = -1;
                pk[u] = 0;
                if (t < cnt) {
                    eNL[u] = nonlocal ? a.nb_ids[nb_b + t] : eL;
                    if (packed) pk[u] = a.nb_pack[nb_b + t];
                }
            }
#pragma unroll
            for (int u = 0; u < ITEMS; ++u) {
                // pairs whose columns all lie left of the node's rows and have no Dirichlet column
                // contribute nothing to the upper triangle nor to the bound part
                if (packed) active[u] = eNL[u] >= 0 && ((pk[u] & 0x7fffffff) >= node || pk[u] < 0);
                else active[u] = eNL[u] >= 0 && (a.el_maxnode[eNL[u]] >= node || a.el_dirichlet[eNL[u]]);
            }
#pragma unroll
            for (int u = 0; u < ITEMS; ++u) {
                if (!active[u]) continue;
                const i64 t = base + u * i64(blockDim.x) + threadIdx.x;
                if (from_pairs) {
                    const T* pr = pairs + i64(iL) * a.plane + (eoff + t) * 4;
                    for (int j = 0; j < NN; ++j) v[u][j][0] = pr[j];
                }
                if constexpr (NN % 4 == 0) {
                    const int4* row = reinterpret_cast<const int4*>(a.el + eNL[u] * NN);
                    for (int q = 0; q < NN / 4; ++q) {
                        const int4 ids = row[q];
                        mj[u][4 * q] = ids.x; mj[u][4 * q + 1] = ids.y; mj[u][4 * q + 2] = ids.z; mj[u][4 * q + 3] = ids.w;
                    }
                } else {
                    for (int j = 0; j < NN; ++j) mj[u][j] = a.el[eNL[u] * NN + j];
                }
            }
#pragma unroll
            for (int u = 0; u < ITEMS; ++u) {
                if (!active[u]) continue;
                if (nonlocal) {
                    if constexpr (!PAIRS) nonlocal_row<DOF, NN, NQ, T>(a, g, xl, hl, eNL[u], v[u]);
                    if (eNL[u] == eL) {
                        T lv[NN][DOF * DOF];
                        local_row<DOF, NN, NQ, T>(a, g, eL, iL, lv);
                        for (int j = 0; j < NN; ++j)
                            for (int q = 0; q < DOF * DOF; ++q) v[u][j][q] += lv[j][q];
                    }
                } else {
                    local_row<DOF, NN, NQ, T>(a, g, eL, iL, v[u]);
                }
                for (int j = 0; j < NN; ++j) {
                    const i64 m = mj[u][j];
                    for (int d = 0; d < DOF; ++d)
                        for (int dc = 0; dc < DOF; ++dc) {
                            const i64 c = i64(DOF) * m + dc;
                            int p = -1;
                            // INNER: inner row, inner column c >= row -- exactly the columns in the
                            // row's upper part, so the bitmap decides membership
                            const i64 off = c - rows[d];
                            if (rinner[d] && off >= 0 && off < i64(s_woff[d + 1] - s_woff[d]) * 32) {
                                const unsigned int word = bm[s_woff[d] + int(off >> 5)];
                                const unsigned int bit = 1u << (off & 31);
                                if (word & bit) p = s_off[d] + wpos[s_woff[d] + int(off >> 5)] + __popc(word & (bit - 1u));
                            }
                            // BOUND: Dirichlet column c != row (the bound part holds only those)
                            if (p < 0 && brow_len[d] > 0 && c != rows[d]) {
                                const int q = lower_bound_i32(cl + s_off[DOF + d], brow_len[d], i32(c));
                                if (q < brow_len[d] && cl[s_off[DOF + d] + q] == i32(c)) p = s_off[DOF + d] + q;
                            }
                            pos[u][j][d * DOF + dc] = p;
                        }
                }
            }
            if constexpr (INJECTIVE) {
                // rounds by local index j: e -> el[e][j] is injective, so no two threads of a round
                // add to one position
                for (int u = 0; u < ITEMS; ++u)
                    for (int j = 0; j < NN; ++j) {
                        __syncthreads();
                        if (active[u])
                            for (int q = 0; q < DOF * DOF; ++q) {
                                const int p = pos[u][j][q];
                                if (p >= 0) acc[p] += v[u][j][q];
                            }
                    }
            } else {
                // rounds by the element's rank at the column node: two different elements never
                // hold the same rank at one node, so a round again has at most one addition per
                // position, and the order of additions is fixed by (step, item, rank)
                for (int u = 0; u < ITEMS; ++u) {
                    int rk[NN];
                    if (active[u])
                        for (int j = 0; j < NN; ++j) rk[j] = a.el_rank[eNL[u] * NN + j];
                    for (int rho = 0; rho < a.rounds; ++rho) {
                        __syncthreads();
                        if (active[u])
                            for (int j = 0; j < NN; ++j) {
                                if (rk[j] != rho) continue;
                                for (int q = 0; q < DOF * DOF; ++q) {
                                    const int p = pos[u][j][q];
                                    if (p >= 0) acc[p] += v[u][j][q];
                                }
                            }
                    }
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    for (int d = 0; d < DOF; ++d) {
        for (int k = threadIdx.x; k < row_len[d]; k += blockDim.x) {
            const T value = rinner[d] ? acc[s_off[d] + k] : (k == 0 ? T(1) : T(0));
            if constexpr (F64) a.val[a.diag[rows[d]] + k] = value;
            else a.val32[a.diag[rows[d]] + k] = value;
        }
        if constexpr (F64)
            for (int k = threadIdx.x; k < brow_len[d]; k += blockDim.x)
                a.b_val[a.b_outer[rows[d]] + k] = acc[s_off[DOF + d] + k];
    }
}

// Local-only assembly (assemble_part::LOCAL, or no nonlocal group): one wave per node instead of
// one workgroup.  A node's rows have at most 64 upper and 64 bound entries here (checked by the
// host), so lane k owns entry k of each.  The lanes first compute one (element, local column)
// entry each -- a node's elements times nn of them, 64 at a time -- and the owners then add them,
// broadcast one after another in element, then column order (the node kernel's order), so no
// atomics and no barriers are needed.
template<int DOF, int NN, int NQ, class T>
__global__ __launch_bounds__(256) void k_assemble_local(AssemblyArgs a) {
    constexpr bool F64 = sizeof(T) == sizeof(double);
    const i64 node = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (node >= a.N) return;
    i64 rows[DOF], ub[DOF], bb[DOF];
    int ulen[DOF], blen[DOF];
    i32 cu[DOF], cb[DOF];
    T au[DOF], ab[DOF];
    bool rinner[DOF];
    for (int d = 0; d < DOF; ++d) {
        rows[d] = i64(DOF) * node + d;
        ub[d] = a.diag[rows[d]];
        ulen[d] = int(a.outer[rows[d] + 1] - ub[d]);
        bb[d] = a.b_outer[rows[d]];
        blen[d] = int(a.b_outer[rows[d] + 1] - bb[d]);
        cu[d] = lane < ulen[d] ? a.cols[ub[d] + lane] : -1;
        cb[d] = lane < blen[d] ? a.b_cols[bb[d] + lane] : -1;
        au[d] = T(0);
        ab[d] = T(0);
        rinner[d] = a.inner[rows[d]];
    }
    const i32 k0 = a.ne_off[node];
    const int P = (a.ne_off[node + 1] - k0) * NN;
    for (int p0 = 0; p0 < P; p0 += 64) {
        T vb[DOF * DOF];
        i32 mcol = -1;
        if (p0 + lane < P) {
            const i32 eL = a.ne_ids[k0 + (p0 + lane) / NN];
            const int j = (p0 + lane) % NN;
            int iL = 0;
            for (int i = 0; i < NN; ++i) if (a.el[i64(eL) * NN + i] == node) iL = i;
            local_entry<DOF, NN, NQ, T>(a, a.groups[a.group_of[eL]], eL, iL, j, vb);
            mcol = a.el[i64(eL) * NN + j];
        } else {
            for (int k = 0; k < DOF * DOF; ++k) vb[k] = T(0);
        }
        const int cnt = P - p0 < 64 ? P - p0 : 64;
        for (int q = 0; q < cnt; ++q) {
            const i64 m = __shfl(mcol, q);
            T v[DOF * DOF];
            for (int k = 0; k < DOF * DOF; ++k) v[k] = __shfl(vb[k], q);
            for (int d = 0; d < DOF; ++d)
                for (int dc = 0; dc < DOF; ++dc) {
                    const i64 c = i64(DOF) * m + dc;
                    if (a.inner[c]) {                  // INNER: inner row, c >= row (upper storage)
                        if (rinner[d] && c >= rows[d] && cu[d] == i32(c)) au[d] += v[d * DOF + dc];
                    } else if (c != rows[d]) {         // BOUND
                        if (cb[d] == i32(c)) ab[d] += v[d * DOF + dc];
                    }
                }
        }
    }
    for (int d = 0; d < DOF; ++d) {
        if (lane < ulen[d]) {
            const T value = rinner[d] ? au[d] : (lane == 0 ? T(1) : T(0));
            if constexpr (F64) a.val[ub[d] + lane] = value;
            else a.val32[ub[d] + lane] = value;
        }
        if constexpr (F64)
            if (lane < blen[d]) a.b_val[bb[d] + lane] = ab[d];
    }
}

// Lower triangle of the full storage: a(r, c) = a(c, r) for c < r.  Where rows r and c have
// mirror-image patterns (the interior of a structured mesh: the j-th entry left of r's diagonal
// and the j-th entry right of c's are each other's transposes) the guess diag[c] + j is checked
// with one load; otherwise r is found in row c by binary search.
template<class T>
__global__ void k_mirror(i64 rows, const i64* __restrict__ outer, const i32* __restrict__ cols,
                         const i64* __restrict__ diag, T* __restrict__ val) {
    const i64 r = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (r >= rows) return;
    const int lane = threadIdx.x & 63;
    const i64 dr = diag[r];
    for (i64 p = outer[r] + lane; p < dr; p += 64) {
        const i32 c = cols[p];
        const i64 b = diag[c], e = outer[c + 1];
        const i64 guess = b + (dr - p);
        if (guess < e && cols[guess] == i32(r)) {
            val[p] = val[guess];
            continue;
        }
        const int q = lower_bound_i32(cols + b, int(e - b), i32(r));
        if (q < e - b && cols[b + q] == i32(r)) val[p] = val[b + q];
    }
}

// ---- symmetric pattern from the upper rows (neighbour lists that are not a symmetric relation) ----
// With the rotated-ellipse distance e' in N(e) does not imply e in N(e'), so the columns a row reaches
// left of its diagonal need not be the transpose of the upper rows.  The reference keeps only the
// upper triangle of each row's own traversal (matrix_separator_base.hpp:50-58) and reads the matrix as
// symmetric; here the lower part of every row is rebuilt as the transpose of the upper rows: the
// (col, row) keys of the entries right of the diagonal are sorted, which lists each row's lower
// columns in ascending order.
__global__ void k_lower_keys(i64 rows, const i64* __restrict__ outer, const i32* __restrict__ cols,
                             const i64* __restrict__ diag, unsigned long long* __restrict__ keys,
                             i64* __restrict__ lower_cnt) {
    const i64 r = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (r >= rows) return;
    const int lane = threadIdx.x & 63;
    const i64 dr = diag[r];
    for (i64 p = outer[r] + lane; p < outer[r + 1]; p += 64) {
        unsigned long long key = ~0ull;
        if (p > dr) {
            const i32 c = cols[p];
            key = (static_cast<unsigned long long>(c) << 32) | static_cast<unsigned long long>(r);
            atomicAdd(reinterpret_cast<unsigned long long*>(&lower_cnt[c]), 1ull);
        }
        keys[p] = key;
    }
}

__global__ void k_sym_rows(i64 rows, const i64* __restrict__ outer, const i64* __restrict__ diag,
                           const i64* __restrict__ lower_cnt, i64* __restrict__ len) {
    const i64 r = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < rows) len[r] = lower_cnt[r] + (outer[r + 1] - diag[r]);
}

__global__ void k_sym_fill(i64 rows, const i64* __restrict__ outer, const i32* __restrict__ cols,
                           const i64* __restrict__ diag, const i64* __restrict__ lower_off,
                           const unsigned long long* __restrict__ sorted, const i64* __restrict__ outer2,
                           i32* __restrict__ cols2, i64* __restrict__ diag2) {
    const i64 r = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (r >= rows) return;
    const int lane = threadIdx.x & 63;
    const i64 lb = lower_off[r], nl = lower_off[r + 1] - lb, o2 = outer2[r];
    for (i64 k = lane; k < nl; k += 64) cols2[o2 + k] = i32(sorted[lb + k] & 0xffffffffull);
    const i64 ub = diag[r], nu = outer[r + 1] - ub;
    for (i64 k = lane; k < nu; k += 64) cols2[o2 + nl + k] = cols[ub + k];
    if (lane == 0) diag2[r] = o2 + nl;
}

void symmetrize_pattern(nlf_matrix_s* mat, hipStream_t s) {
    const i64 rows = mat->rows, nnz = mat->nnz_full;
    DevBuf<unsigned long long> keys{size_t(std::max<i64>(nnz, 1))}, sorted{size_t(std::max<i64>(nnz, 1))};
    DevBuf<i64> lower_off{size_t(rows + 1)}, len{size_t(rows + 1)};
    NLF_HIP(hipMemsetAsync(lower_off.p, 0, lower_off.bytes(), s));
    hipLaunchKernelGGL(k_lower_keys, dim3(blocks_for(rows, 4)), dim3(256), 0, s, rows, mat->outer.p, mat->cols.p,
                       mat->diag.p, keys.p, lower_off.p);
    NLF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sym_rows, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows, mat->outer.p, mat->diag.p,
                       lower_off.p, len.p);
    NLF_HIP(hipGetLastError());
    exclusive_scan_i64(lower_off.p, rows, s);
    exclusive_scan_i64(len.p, rows, s);
    sort_keys_u64(keys.p, sorted.p, nnz, s);
    i64 nnz2 = 0;
    download(&nnz2, len.p + rows, 1, s);
    DevBuf<i32> cols2{size_t(std::max<i64>(nnz2, 1))};
    DevBuf<i64> diag2{size_t(rows)};
    hipLaunchKernelGGL(k_sym_fill, dim3(blocks_for(rows, 4)), dim3(256), 0, s, rows, mat->outer.p, mat->cols.p,
                       mat->diag.p, lower_off.p, sorted.p, len.p, cols2.p, diag2.p);
    NLF_HIP(hipGetLastError());
    NLF_HIP(hipStreamSynchronize(s));
    mat->outer = std::move(len);
    mat->cols = std::move(cols2);
    mat->diag = std::move(diag2);
    mat->nnz_full = nnz2;
}

// Largest |col - row| of a sorted-row pattern (rows are ascending: first and last entries decide).
__global__ void k_bandwidth(i64 rows, const i64* __restrict__ outer, const i32* __restrict__ cols, unsigned* __restrict__ mx) {
    const i64 r = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows || outer[r] == outer[r + 1]) return;
    const i64 lo = r - i64(cols[outer[r]]), hi = i64(cols[outer[r + 1] - 1]) - r;
    atomicMax(mx, unsigned(std::min<i64>(std::max(lo, hi), i64(0x7fffffff))));
}

// 16-bit column offsets col - row for the SpMV (only when the bandwidth fits).
__global__ void k_to_off16(i64 rows, const i64* __restrict__ outer, const i32* __restrict__ cols, int16_t* __restrict__ off) {
    const i64 r = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (r >= rows) return;
    for (i64 p = outer[r] + (threadIdx.x & 63); p < outer[r + 1]; p += 64) off[p] = int16_t(i64(cols[p]) - r);
}

__global__ void k_elem_minnode(i64 E, int nn, const i32* __restrict__ el, i32* __restrict__ mn) {
    const i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    i32 m = el[e * nn];
    for (int i = 1; i < nn; ++i) m = min(m, el[e * nn + i]);
    mn[e] = m;
}

// The activity test's data per neighbour-list entry (AssemblyArgs::nb_pack).
// Four entries per thread and step (16-byte loads and stores; the buffers start 256-byte aligned),
// the last total % 4 one by one.
__global__ void k_nb_pack(i64 total, const i32* __restrict__ nb_ids, const i32* __restrict__ el_maxnode,
                          const uint8_t* __restrict__ el_dir, i32* __restrict__ out) {
    auto pack = [&](i32 e) { return el_maxnode[e] | (el_dir[e] ? i32(0x80000000u) : 0); };
    const i64 quads = total / 4, stride = i64(gridDim.x) * blockDim.x;
    const int4* in4 = reinterpret_cast<const int4*>(nb_ids);
    int4* out4 = reinterpret_cast<int4*>(out);
    for (i64 q = i64(blockIdx.x) * blockDim.x + threadIdx.x; q < quads; q += stride) {
        const int4 e = in4[q];
        out4[q] = make_int4(pack(e.x), pack(e.y), pack(e.z), pack(e.w));
    }
    for (i64 q = 4 * quads + i64(blockIdx.x) * blockDim.x + threadIdx.x; q < total; q += stride) out[q] = pack(nb_ids[q]);
}

__global__ void k_el_dirichlet(i64 E, int nn, int dof, const i32* __restrict__ el, const uint8_t* __restrict__ inner,
                               uint8_t* __restrict__ out) {
    const i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    uint8_t has = 0;
    for (int i = 0; i < nn; ++i)
        for (int d = 0; d < dof; ++d) has |= uint8_t(!inner[i64(dof) * el[e * nn + i] + d]);
    out[e] = has;
}

// Per node: entries of its rows (upper + bound) -> mx[0]; bitmap words of its upper rows' column
// spans -> mx[1] (the value pass's LDS).
__global__ void k_row_max(i64 rows, const i64* __restrict__ outer, const i64* __restrict__ diag,
                          const i64* __restrict__ b_outer, const i32* __restrict__ cols, int dof,
                          unsigned long long* __restrict__ mx) {
    const i64 n = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n * dof >= rows) return;
    i64 s = 0, w = 0;
    for (int d = 0; d < dof; ++d) {
        const i64 r = n * dof + d;
        s += (outer[r + 1] - diag[r]) + (b_outer[r + 1] - b_outer[r]);
        w += (cols[outer[r + 1] - 1] - r) / 32 + 1;
    }
    atomicMax(mx, static_cast<unsigned long long>(s));
    atomicMax(mx + 1, static_cast<unsigned long long>(w));
}

// slots[e] = neighbour count of e rounded up to a multiple of four (the pair-block slots of e).
__global__ void k_padded_slots(i64 E, const i64* __restrict__ nb_off, i64* __restrict__ slots) {
    const i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < E) slots[e] = (nb_off[e + 1] - nb_off[e] + 3) & ~i64(3);
}

// eoff[list[i]] = ofs[i]: chunk slots of the chunk's elements.
__global__ void k_scatter_offsets(i64 n, const i32* __restrict__ list, const i64* __restrict__ ofs, i64* __restrict__ eoff) {
    const i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) eoff[list[i]] = ofs[i];
}

// Longest upper part and longest bound part over all rows.
__global__ void k_row_lengths(i64 rows, const i64* __restrict__ outer, const i64* __restrict__ diag,
                              const i64* __restrict__ b_outer, unsigned long long* __restrict__ mx) {
    const i64 r = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    atomicMax(&mx[0], static_cast<unsigned long long>(outer[r + 1] - diag[r]));
    atomicMax(&mx[1], static_cast<unsigned long long>(b_outer[r + 1] - b_outer[r]));
}

template<class T>
__global__ void k_add_pairs(i64 n, const i64* __restrict__ rc, const double* __restrict__ v,
                            const i64* __restrict__ outer, const i32* __restrict__ cols, T* __restrict__ val) {
    const i64 k = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const i64 r = rc[2 * k], c = rc[2 * k + 1];
    const i64 b = outer[r], e = outer[r + 1];
    const int q = lower_bound_i32(cols + b, int(e - b), i32(c));
    if (q < e - b && cols[b + q] == c) val[b + q] += T(v[k]);
    if (r != c) {
        const i64 b2 = outer[c], e2 = outer[c + 1];
        const int q2 = lower_bound_i32(cols + b2, int(e2 - b2), i32(r));
        if (q2 < e2 - b2 && cols[b2 + q2] == r) val[b2 + q2] += T(v[k]);
    }
}

__global__ void k_to_f32(i64 n, const double* __restrict__ a, float* __restrict__ b) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) b[i] = float(a[i]);
}

// hi = fp32(v); lo = the remainder v - hi in units of 2^-16 ulp(hi), rounded to the nearest
// integer.  |v - hi| <= ulp(hi) / 2, so lo fits 16 bits and hi + lo carries 40 significand bits:
// |hi + lo * 2^-16 ulp(hi) - v| <= 2^-17 ulp(hi) <= 2^-40 |v|.  ulp(hi) = 2^(e - 150) for the
// biased exponent e of hi (e = 1 for subnormals), so the scale is 2^(e - 166).
__device__ __forceinline__ double split_scale(float h) {
    unsigned e = (__float_as_uint(h) >> 23) & 0xffu;
    e = e ? e : 1u;
    return __longlong_as_double(static_cast<long long>(e + 857u) << 52);      // 2^(e - 166)
}

__global__ void k_to_split(i64 n, const double* __restrict__ a, float* __restrict__ hi, int16_t* __restrict__ lo) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) {
        const double v = a[i];
        const float h = float(v);
        const double t = rint((v - double(h)) / split_scale(h));                 // exact power-of-two scaling
        hi[i] = h;
        lo[i] = int16_t(fmin(fmax(t, -32768.0), 32767.0));
    }
}

// Per-block partials of max|b - a|, max|a|, sum (b - a)^2, sum a^2 (fixed order: deterministic).
__global__ __launch_bounds__(256) void k_compare32(i64 n, const double* __restrict__ a, const float* __restrict__ b,
                                                   double* __restrict__ part) {
    __shared__ double red[4][256];
    double m0 = 0.0, m1 = 0.0, s0 = 0.0, s1 = 0.0;
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) {
        const double x = a[i], d = double(b[i]) - x;
        m0 = fmax(m0, fabs(d)); m1 = fmax(m1, fabs(x));
        s0 += d * d; s1 += x * x;
    }
    red[0][threadIdx.x] = m0; red[1][threadIdx.x] = m1; red[2][threadIdx.x] = s0; red[3][threadIdx.x] = s1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + w]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + w]);
            red[2][threadIdx.x] += red[2][threadIdx.x + w];
            red[3][threadIdx.x] += red[3][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// nodes [a.node0, a.node0 + count); threads per workgroup 64, 128, 256 or 512
template<int DOF, int NN, int NQ, class T, bool PAIRS>
void launch_assemble_kernel(const AssemblyArgs& a, bool injective, size_t lds, dim3 g, hipStream_t s, int threads) {
    if (threads == 512) {
        if (injective) hipLaunchKernelGGL((k_assemble<DOF, NN, NQ, true, T, 512, PAIRS>), g, dim3(512), lds, s, a);
        else hipLaunchKernelGGL((k_assemble<DOF, NN, NQ, false, T, 512, PAIRS>), g, dim3(512), lds, s, a);
    } else if (threads == 128) {
        if (injective) hipLaunchKernelGGL((k_assemble<DOF, NN, NQ, true, T, 128, PAIRS>), g, dim3(128), lds, s, a);
        else hipLaunchKernelGGL((k_assemble<DOF, NN, NQ, false, T, 128, PAIRS>), g, dim3(128), lds, s, a);
    } else if (threads == 64) {
        if (injective) hipLaunchKernelGGL((k_assemble<DOF, NN, NQ, true, T, 64, PAIRS>), g, dim3(64), lds, s, a);
        else hipLaunchKernelGGL((k_assemble<DOF, NN, NQ, false, T, 64, PAIRS>), g, dim3(64), lds, s, a);
    } else {
        if (injective) hipLaunchKernelGGL((k_assemble<DOF, NN, NQ, true, T, ASM_THREADS, PAIRS>), g, dim3(ASM_THREADS), lds, s, a);
        else hipLaunchKernelGGL((k_assemble<DOF, NN, NQ, false, T, ASM_THREADS, PAIRS>), g, dim3(ASM_THREADS), lds, s, a);
    }
}

// The pair-block variant runs when the arguments carry pair blocks (thermal Q4).
template<int DOF, int NN, int NQ, class T>
void launch_assemble(const AssemblyArgs& a, bool injective, size_t lds, i64 count, hipStream_t s, int threads = ASM_THREADS) {
    if (count <= 0) return;
    const dim3 g(static_cast<unsigned>(count));
    if constexpr (DOF == 1 && NN == 4 && NQ == 4) {
        if (sizeof(T) == sizeof(double) ? a.pairs != nullptr : a.pairs32 != nullptr) {
            launch_assemble_kernel<DOF, NN, NQ, T, true>(a, injective, lds, g, s, threads);
            return;
        }
    }
    launch_assemble_kernel<DOF, NN, NQ, T, false>(a, injective, lds, g, s, threads);
}

}  // namespace nlf

using namespace nlf;

namespace {

constexpr int PATTERN_MAX_WORDS = 12288;   // 48 KiB bitmap: column node ids within a span of 393k

GroupParams make_group(const nlf_model& md, int physics, const void* materials, int g, bool only_local) {
    if (physics == 2) {                                  // heat capacity: local theory, c[0] = rho c
        GroupParams gp = model_group_params(md, true);
        gp.p1 = 1.0;
        gp.c[0] = static_cast<const double*>(materials)[g];
        return gp;
    }
    GroupParams gp = model_group_params(md, only_local);
    if (physics == 0) {
        const nlf_thermal_material& t = static_cast<const nlf_thermal_material*>(materials)[g];
        gp.ckind = t.kind;
        gp.c[0] = t.c[0]; gp.c[1] = t.c[1]; gp.c[2] = t.c[2];
    } else {
        elastic_group_params(static_cast<const nlf_elastic_material*>(materials)[g], gp);
    }
    return gp;
}

// 16-bit column offsets for the SpMV when every |col - row| < 2^15, and the row-pattern dictionary
// over them (NLF_OPT_ROW_PATTERNS); any earlier ones are dropped first.
static void finish_spmv_storage(nlf_matrix_s* mat, hipStream_t s) {
    const i64 rows = mat->rows;
    mat->off16.release();
    mat->pat.release();
    mat->pstart.release();
    mat->pat_size = mat->pat_groups = 0;
    DevBuf<unsigned> bw{1};
    NLF_HIP(hipMemsetAsync(bw.p, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_bandwidth, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows, mat->outer.p, mat->cols.p, bw.p);
    unsigned hb = 0;
    download(&hb, bw.p, 1, s);
    if (hb < 32768u && mat->nnz_full > 0) {
        mat->off16.alloc(size_t(mat->nnz_full));
        hipLaunchKernelGGL(k_to_off16, dim3(blocks_for(rows, 4)), dim3(256), 0, s, rows, mat->outer.p, mat->cols.p,
                           mat->off16.p);
        NLF_HIP(hipGetLastError());
        if (mat->row_patterns) build_row_patterns(mat, s);
    }
}

// The full-storage SpMV's layout is built on the first product that reads it, not with the values:
// the FP64 solvers read the symmetric slice storage instead, and a matrix used only by them never
// needs it.  A new pattern or a rebuilt matrix drops it.
static void drop_spmv_storage(nlf_matrix_s* mat) {
    mat->off16.release();
    mat->pat.release();
    mat->pstart.release();
    mat->pat_size = mat->pat_groups = 0;
    mat->spmv_storage_ready = false;
    mat->row_patterns = mat->ctx->row_patterns;      // the option in force at assembly applies
}

// ---- the Neumann problem's Lagrange row and column (conductivity_matrix_2d.hpp:39-72, 83-104) ----
// c[node] = sum over the node's elements (ascending) of integrate_basic = sum_q w_q N_i(q) |J_q|
template<int NN, int NQ>
__global__ __launch_bounds__(256) void k_basic_integral(i64 N, const i32* __restrict__ ne_off, const i32* __restrict__ ne_ids,
                                                        const i32* __restrict__ el, const double* __restrict__ jac, FeTables t,
                                                        double* __restrict__ c) {
    const i64 node = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= N) return;
    double acc = 0.0;
    for (i32 k = ne_off[node]; k < ne_off[node + 1]; ++k) {
        const i64 e = ne_ids[k];
        int i = 0;
        for (int j = 0; j < NN; ++j)
            if (el[e * NN + j] == node) i = j;
        double integral = 0.0;
        for (int q = 0; q < NQ; ++q) integral += t.w[q] * t.qN[i * NQ + q] * jac[e * NQ + q];
        acc += integral;
    }
    c[node] = acc;
}

// The bordered full storage: row r < n copies its entries and appends (r, n) = c[r]; row n is
// (n, 0..n-1) = c and (n, n) = 0.  One wave per row.
__global__ __launch_bounds__(256) void k_border_rows(i64 n, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                     const double* __restrict__ val, const double* __restrict__ c,
                                                     i64* __restrict__ outer2, i32* __restrict__ cols2,
                                                     double* __restrict__ val2, i64* __restrict__ diag2,
                                                     const i64* __restrict__ diag) {
    const i64 r = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r > n) return;
    if (r < n) {
        const i64 b = outer[r], e = outer[r + 1], b2 = b + r;
        for (i64 p = b + lane; p < e; p += 64) {
            cols2[b2 + (p - b)] = cols[p];
            val2[b2 + (p - b)] = val[p];
        }
        if (lane == 0) {
            cols2[b2 + (e - b)] = i32(n);
            val2[b2 + (e - b)] = c[r];
            outer2[r] = b2;
            diag2[r] = diag[r] + r;
        }
    } else {
        const i64 b2 = outer[n] + n;
        for (i64 k = lane; k <= n; k += 64) {
            cols2[b2 + k] = i32(k);
            val2[b2 + k] = k < n ? c[k] : 0.0;
        }
        if (lane == 0) {
            outer2[n] = b2;
            outer2[n + 1] = b2 + n + 1;
            diag2[n] = b2 + n;
        }
    }
}

void border_neumann(nlf_matrix_s* mat, hipStream_t s) {
    const nlf_mesh m = mat->mesh;
    const i64 n = mat->rows;
    DevBuf<double> c{size_t(std::max<i64>(n, 1))};
    if (m->nn == NLF_ELEMENT_Q4)
        hipLaunchKernelGGL((k_basic_integral<4, 4>), dim3(blocks_for(n, 256)), dim3(256), 0, s, n, m->ne_off.p, m->ne_ids.p,
                           m->el.p, m->jac.p, m->fe, c.p);
    else
        hipLaunchKernelGGL((k_basic_integral<8, 9>), dim3(blocks_for(n, 256)), dim3(256), 0, s, n, m->ne_off.p, m->ne_ids.p,
                           m->el.p, m->jac.p, m->fe, c.p);
    const i64 nnz2 = mat->nnz_full + 2 * n + 1;
    NLF_CHECK(nnz2 < (i64(1) << 40), NLF_ERR_ARGUMENT, "pattern too large");
    DevBuf<i64> outer2{size_t(n + 2)}, diag2{size_t(n + 1)};
    DevBuf<i32> cols2{size_t(nnz2)};
    DevBuf<double> val2{size_t(nnz2)};
    hipLaunchKernelGGL(k_border_rows, dim3(blocks_for(n + 1, 4)), dim3(256), 0, s, n, mat->outer.p, mat->cols.p, mat->val.p,
                       c.p, outer2.p, cols2.p, val2.p, diag2.p, mat->diag.p);
    NLF_HIP(hipGetLastError());
    // the Lagrange row is an inner row with no bound entries
    std::vector<uint8_t> inner(size_t(n + 1), 1);
    DevBuf<uint8_t> inner2;
    upload(inner2, inner.data(), inner.size(), s);
    DevBuf<i64> b_outer2{size_t(n + 2)};
    NLF_HIP(hipMemsetAsync(b_outer2.p, 0, b_outer2.bytes(), s));
    NLF_HIP(hipStreamSynchronize(s));
    mat->outer = std::move(outer2);
    mat->diag = std::move(diag2);
    mat->cols = std::move(cols2);
    mat->val = std::move(val2);
    mat->inner = std::move(inner2);
    mat->b_outer = std::move(b_outer2);
    mat->rows = n + 1;
    mat->nnz_full = nnz2;
    mat->neumann = true;
    drop_spmv_storage(mat);
    NLF_HIP(hipStreamSynchronize(s));
}

// Value pass on the matrix's pattern.  T = double fills val and b_val (and the 16-bit column
// offsets); T = float fills val32 with FP32 contractions and FP32 accumulation.  Thermal Q4
// nonlocal blocks go through the MFMA pair kernel when its pair buffer (16 values per neighbour
// pair) fits in half of the free device memory; option NLF_OPT_ASSEMBLY_PAIRS = 0 forces the
// node-centric contraction.
// Node chunks of the MFMA pair assembly.  An element's blocks take its neighbour count rounded up
// to a multiple of four slots (k_pairs_mfma writes runs of four).  single: every element's blocks
// (all_pairs slots) fit the budget, one pass over all elements.  Otherwise a chunk is a run of nodes [n0, n1) whose
// elements' blocks fit the budget: list[lbeg, lend) are its elements (each once, ascending by first
// touching node) and cofs their first block slot within the chunk.
struct PairChunk {
    i64 n0, n1, lbeg, lend, total;
};
struct PairPlan {
    i64 per_pair = 0, all_pairs = 0, max_total = 0;
    i64 entries = 0;                     // neighbour-list entries (all_pairs counts them padded to 4 per element)
    bool single = true;
    std::vector<PairChunk> chunks;
    std::vector<i32> list;
    std::vector<i64> cofs;
    // order in which the pair kernel visits the elements, sorted on the device: (chunk << 32 |
    // Morton key of the element) for every entry of list (single: every element, its Morton key)
    std::vector<uint64_t> order_key;
    std::vector<i32> order_val;
};

i64 pair_budget(const nlf_mesh_s* m) {
    return m->ctx->pair_budget_mib > 0 ? i64(m->ctx->pair_budget_mib) << 20 : i64(8) << 30;
}

// nb_off: host copy of the neighbour offsets (E + 1).
PairPlan plan_pairs(const nlf_mesh_s* m, const std::vector<i64>& nb_off, i64 per_pair) {
    PairPlan plan;
    const i64 E = m->E, N = m->N, budget = pair_budget(m);
    plan.per_pair = per_pair;
    auto slots = [&](i64 e) { return (nb_off[size_t(e + 1)] - nb_off[size_t(e)] + 3) & ~i64(3); };
    for (i64 e = 0; e < E; ++e) plan.all_pairs += slots(e);
    plan.entries = nb_off[size_t(E)];
    plan.single = plan.all_pairs * per_pair <= budget;
    const std::vector<uint32_t>& key = m->el_key_h;
    if (plan.single) {
        plan.order_key.resize(size_t(E));
        plan.order_val.resize(size_t(E));
        for (i64 e = 0; e < E; ++e) {
            plan.order_key[size_t(e)] = key[size_t(e)];
            plan.order_val[size_t(e)] = i32(e);
        }
        return plan;
    }
    const std::vector<i32>& noff = m->ne_off_h;
    const std::vector<i32>& nel = m->ne_ids_h;
    std::vector<i32> stamp(size_t(E), -1);
    plan.list.reserve(size_t(E) + size_t(E) / 4);
    plan.cofs.reserve(size_t(E) + size_t(E) / 4);
    for (i64 n0 = 0; n0 < N;) {
        const int chunk = int(plan.chunks.size());
        const i64 lbeg = i64(plan.list.size());
        i64 total = 0, n1 = n0;
        for (; n1 < N; ++n1) {
            i64 add = 0;
            for (i32 k = noff[size_t(n1)]; k < noff[size_t(n1 + 1)]; ++k) {
                const i32 e = nel[size_t(k)];
                if (stamp[size_t(e)] != chunk) add += slots(e);
            }
            if (n1 > n0 && (total + add) * per_pair > budget) break;
            for (i32 k = noff[size_t(n1)]; k < noff[size_t(n1 + 1)]; ++k) {
                const i32 e = nel[size_t(k)];
                if (stamp[size_t(e)] == chunk) continue;
                stamp[size_t(e)] = chunk;
                plan.list.push_back(e);
                plan.cofs.push_back(total);
                total += slots(e);
            }
        }
        plan.chunks.push_back(PairChunk{n0, n1, lbeg, i64(plan.list.size()), total});
        plan.max_total = std::max(plan.max_total, total);
        n0 = n1;
    }
    plan.order_key.resize(plan.list.size());
    for (const PairChunk& c : plan.chunks)
        for (i64 i = c.lbeg; i < c.lend; ++i)
            plan.order_key[size_t(i)] = (uint64_t(&c - plan.chunks.data()) << 32) | key[size_t(plan.list[size_t(i)])];
    plan.order_val = plan.list;
    return plan;
}

// plan: the pair chunks, when the caller planned them (while the pattern kernels ran); null or
// planned for another value size: planned here.
template<class T>
void assemble_values(nlf_matrix_s* mat, hipStream_t s, const PairPlan* plan = nullptr) {
    constexpr bool F64 = sizeof(T) == sizeof(double);
    const nlf_mesh m = mat->mesh;
    const int dof = mat->dof;
    const i64 rows = mat->rows;
    const bool only_local = mat->only_local;
    bool any_nonlocal = false;
    for (const GroupParams& g : mat->groups) any_nonlocal |= g.nonlocal != 0;
    DevBuf<GroupParams> d_groups;
    upload(d_groups, mat->groups.data(), mat->groups.size(), s);
    DevBuf<i32> el_min{size_t(m->E)};
    hipLaunchKernelGGL(k_elem_minnode, dim3(blocks_for(m->E, 256)), dim3(256), 0, s, m->E, m->nn, m->el.p, el_min.p);
    DevBuf<uint8_t> el_dir{size_t(m->E)};
    hipLaunchKernelGGL(k_el_dirichlet, dim3(blocks_for(m->E, 256)), dim3(256), 0, s, m->E, m->nn, dof, m->el.p,
                       mat->inner.p, el_dir.p);
    DevBuf<unsigned long long> d_max{2};
    NLF_HIP(hipMemsetAsync(d_max.p, 0, 2 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_row_max, dim3(blocks_for(m->N, 256)), dim3(256), 0, s, rows, mat->outer.p, mat->diag.p,
                       mat->b_outer.p, mat->cols.p, dof, d_max.p);
    unsigned long long maxes[2] = {0, 0};
    download(maxes, d_max.p, 2, s);
    const size_t lds = size_t(maxes[0]) * (sizeof(T) + sizeof(i32)) + size_t(maxes[1]) * (sizeof(unsigned int) + 2) + 16;
    NLF_CHECK(lds <= 60 * 1024, NLF_ERR_UNSUPPORTED, "matrix rows too long for the LDS accumulator");
    if constexpr (F64) {
        mat->val.alloc(size_t(std::max<i64>(mat->nnz_full, 1)));
        mat->b_val.alloc(size_t(std::max<i64>(mat->nnz_bound, 1)));
    } else {
        if (!mat->val32.p) mat->val32.alloc(size_t(std::max<i64>(mat->nnz_full, 1)));
    }
    AssemblyArgs aa{};
    aa.N = m->N; aa.el = m->el.p; aa.ne_off = m->ne_off.p; aa.ne_ids = m->ne_ids.p; aa.group_of = m->group_of.p;
    aa.groups = d_groups.p;
    aa.nb_off = m->has_neighbours ? m->nb_off.p : nullptr;
    aa.nb_ids = m->has_neighbours ? m->nb_ids.p : nullptr;
    aa.el_maxnode = m->el_maxnode.p;
    aa.el_dirichlet = el_dir.p; aa.inner = mat->inner.p; aa.xq = m->xq.p; aa.jac = m->jac.p; aa.D = m->D.p;
    for (int q = 0; q < m->fe.nq; ++q) aa.w[q] = m->fe.w[q];
    for (int k = 0; k < m->fe.nn * m->fe.nq; ++k) aa.qN[k] = m->fe.qN[k];
    aa.mass = mat->physics == 2;
    aa.outer = mat->outer.p; aa.cols = mat->cols.p; aa.diag = mat->diag.p; aa.b_outer = mat->b_outer.p;
    aa.b_cols = mat->b_cols.p; aa.val = mat->val.p; aa.b_val = mat->b_val.p; aa.val32 = mat->val32.p;
    aa.local_only = only_local;
    aa.el_rank = m->el_rank.p;
    aa.cq = mat->cq.p;
    aa.rounds = m->max_degree;
    const bool injective = m->injective;
    const bool node_only = m->ctx->assembly_pairs == 0;
    const bool use_pairs = mat->physics == 0 && m->nn == 4 && any_nonlocal && !only_local && !node_only;
    // local-only matrices whose rows fit a wave: one wave per node
    unsigned long long max_upper = 0, max_bound = 0;
    if (!any_nonlocal || only_local) {
        DevBuf<unsigned long long> d_len{2};
        NLF_HIP(hipMemsetAsync(d_len.p, 0, d_len.bytes(), s));
        hipLaunchKernelGGL(k_row_lengths, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows, mat->outer.p, mat->diag.p,
                           mat->b_outer.p, d_len.p);
        unsigned long long h_len[2];
        download(h_len, d_len.p, 2, s);
        max_upper = h_len[0];
        max_bound = h_len[1];
    }
    if ((!any_nonlocal || only_local) && max_upper <= 64 && max_bound <= 64) {
        const unsigned nb = unsigned((m->N + 3) / 4);
        if (mat->physics != 1 && m->nn == 4) hipLaunchKernelGGL((k_assemble_local<1, 4, 4, T>), dim3(nb), dim3(256), 0, s, aa);
        else if (mat->physics != 1 && m->nn == 8) hipLaunchKernelGGL((k_assemble_local<1, 8, 9, T>), dim3(nb), dim3(256), 0, s, aa);
        else if (mat->physics == 1 && m->nn == 4) hipLaunchKernelGGL((k_assemble_local<2, 4, 4, T>), dim3(nb), dim3(256), 0, s, aa);
        else hipLaunchKernelGGL((k_assemble_local<2, 8, 9, T>), dim3(nb), dim3(256), 0, s, aa);
    } else if (use_pairs) {
        // thermal Q4 nonlocal blocks through the MFMA pair kernel.  When every element's blocks (16
        // values per neighbour pair) fit the budget -- 8 GiB, or NLF_OPT_PAIR_BUDGET_MIB -- one pass
        // lists every element once with its blocks at its neighbour-list offset; otherwise node chunk
        // by node chunk (plan_pairs): the buffer is bounded whatever the mesh size, and elements
        // shared by two chunks are contracted once in each.  The chunks run back to back on the
        // stream (each chunk's offsets are written after the previous chunk's kernels, in stream order).
        const i64 E = m->E, N = m->N;
        const i64 per_pair = 16 * i64(sizeof(T));
        PairPlan own;
        if (!plan || plan->per_pair != per_pair) {
            std::vector<i64> nb_off(size_t(E + 1));
            download(nb_off.data(), m->nb_off.p, nb_off.size(), s);
            own = plan_pairs(m, nb_off, per_pair);
            plan = &own;
        }
        const i64 all_pairs = plan->all_pairs;
        DevBuf<i32> nb_pack;
        if (NLF_ASM_NB_PACK && plan->entries > 0) {
            nb_pack.alloc(size_t(plan->entries));
            hipLaunchKernelGGL(k_nb_pack, dim3(4096), dim3(256), 0, s, plan->entries, m->nb_ids.p, m->el_maxnode.p,
                               el_dir.p, nb_pack.p);
            aa.nb_pack = nb_pack.p;
        }
        // the pair kernel's element order: Morton order within each chunk
        DevBuf<i32> d_order;
        {
            DevBuf<uint64_t> k_in, k_out{plan->order_key.size()};
            DevBuf<i32> v_in;
            upload(k_in, plan->order_key.data(), plan->order_key.size(), s);
            upload(v_in, plan->order_val.data(), plan->order_val.size(), s);
            d_order.alloc(plan->order_val.size());
            sort_pairs_u64(k_in.p, k_out.p, v_in.p, d_order.p, i64(plan->order_val.size()), s);
        }
        DevBuf<T> HL{size_t(E) * 32}, HN{size_t(E) * 32}, pairs;
        Weights4 w4{};
        for (int q = 0; q < 4; ++q) w4.w[q] = m->fe.w[q];
        hipLaunchKernelGGL(k_pair_prep<T>, dim3(blocks_for(E * 16, 256)), dim3(256), 0, s, E, m->D.p, w4, m->group_of.p,
                           d_groups.p, HL.p, HN.p);
        PairArgs<T> pr{};
        pr.nb_off = m->nb_off.p; pr.nb_ids = m->nb_ids.p; pr.group_of = m->group_of.p;
        pr.groups = d_groups.p; pr.xq = m->xq.p; pr.HL = HL.p; pr.HN = HN.p; pr.el_min = el_min.p;
        pr.el_max = m->el_maxnode.p; pr.el_dir = el_dir.p;
        const int poly = influence_class(mat->groups);
        auto contract = [&](i64 count) {
            const unsigned pb = unsigned(std::min<i64>((count + 3) / 4, 16384));
            pr.count = count;
            if constexpr (F64) {
                if (poly == 1) hipLaunchKernelGGL(k_pairs_mfma<1>, dim3(pb), dim3(256), 0, s, pr);
                else if (poly == 2) hipLaunchKernelGGL(k_pairs_mfma<2>, dim3(pb), dim3(256), 0, s, pr);
                else hipLaunchKernelGGL(k_pairs_mfma<0>, dim3(pb), dim3(256), 0, s, pr);
            } else {
                hipLaunchKernelGGL(k_pairs_mfma32, dim3(pb), dim3(256), 0, s, pr);
            }
            NLF_HIP(hipGetLastError());
        };
        if (plan->single) {
            AssemblyArgs ca = aa;
            ca.node0 = 0;
            DevBuf<i64> d_eoff{size_t(E + 1)};
            if (all_pairs > 0) {
                hipLaunchKernelGGL(k_padded_slots, dim3(blocks_for(E, 256)), dim3(256), 0, s, E, m->nb_off.p, d_eoff.p);
                exclusive_scan_i64(d_eoff.p, E, s);
                pairs.alloc(size_t(all_pairs) * 16);
                pr.list = d_order.p; pr.eoff = d_eoff.p; pr.pairs = pairs.p; pr.plane = all_pairs * 4;
                contract(E);
                if constexpr (F64) ca.pairs = pairs.p; else ca.pairs32 = pairs.p;
                ca.plane = all_pairs * 4;
                ca.eoff = d_eoff.p;
            }
            launch_assemble<1, 4, 4, T>(ca, injective, lds, N, s, m->ctx->assembly_threads);
            NLF_HIP(hipGetLastError());
        } else {
            DevBuf<i64> d_eoff{size_t(E)}, d_cofs;
            DevBuf<i32> d_list;
            upload(d_list, plan->list.data(), plan->list.size(), s);
            upload(d_cofs, plan->cofs.data(), plan->cofs.size(), s);
            if (plan->max_total > 0) pairs.alloc(size_t(plan->max_total) * 16);
            pr.eoff = d_eoff.p;
            for (const PairChunk& c : plan->chunks) {
                AssemblyArgs ca = aa;
                ca.node0 = c.n0;
                const i64 count = c.lend - c.lbeg;
                if (count > 0 && c.total > 0) {
                    hipLaunchKernelGGL(k_scatter_offsets, dim3(blocks_for(count, 256)), dim3(256), 0, s, count,
                                       d_list.p + c.lbeg, d_cofs.p + c.lbeg, d_eoff.p);
                    pr.list = d_order.p + c.lbeg; pr.pairs = pairs.p; pr.plane = c.total * 4;
                    contract(count);
                    if constexpr (F64) ca.pairs = pairs.p; else ca.pairs32 = pairs.p;
                    ca.plane = c.total * 4;
                    ca.eoff = d_eoff.p;
                }
                launch_assemble<1, 4, 4, T>(ca, injective, lds, c.n1 - c.n0, s, m->ctx->assembly_threads);
                NLF_HIP(hipGetLastError());
            }
            NLF_HIP(hipStreamSynchronize(s));          // before the chunk offsets and lists are released
        }
        NLF_HIP(hipStreamSynchronize(s));              // before the pair buffer is released
    } else if (mat->physics != 1 && m->nn == 4) launch_assemble<1, 4, 4, T>(aa, injective, lds, m->N, s, m->ctx->assembly_threads);
    else if (mat->physics != 1 && m->nn == 8) launch_assemble<1, 8, 9, T>(aa, injective, lds, m->N, s, m->ctx->assembly_threads);
    else if (mat->physics == 1 && m->nn == 4) launch_assemble<2, 4, 4, T>(aa, injective, lds, m->N, s, m->ctx->assembly_threads);
    else launch_assemble<2, 8, 9, T>(aa, injective, lds, m->N, s, m->ctx->assembly_threads);
    NLF_HIP(hipGetLastError());
    T* values = nullptr;
    if constexpr (F64) values = mat->val.p; else values = mat->val32.p;
    hipLaunchKernelGGL(k_mirror<T>, dim3(blocks_for(rows, 4)), dim3(256), 0, s, rows, mat->outer.p, mat->cols.p,
                       mat->diag.p, values);
    NLF_HIP(hipGetLastError());
    if constexpr (F64) drop_spmv_storage(mat);
    NLF_HIP(hipStreamSynchronize(s));
}

// Adds symmetric (row <= col) pair contributions to the FP64 values, and to the FP32 values when
// they exist; remembers them so that a later FP32 value pass repeats them.
void add_pairs_device(nlf_matrix_s* a, const std::vector<i64>& rc, const std::vector<double>& v) {
    if (v.empty()) return;
    hipStream_t s = a->ctx->stream;
    DevBuf<i64> d_rc;
    DevBuf<double> d_v;
    upload(d_rc, rc.data(), rc.size(), s);
    upload(d_v, v.data(), v.size(), s);
    const i64 n = i64(v.size());
    hipLaunchKernelGGL(k_add_pairs<double>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, d_rc.p, d_v.p, a->outer.p,
                       a->cols.p, a->val.p);
    if (a->val32.p)
        hipLaunchKernelGGL(k_add_pairs<float>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, d_rc.p, d_v.p,
                           a->outer.p, a->cols.p, a->val32.p);
    NLF_HIP(hipGetLastError());
    NLF_HIP(hipStreamSynchronize(s));
    a->split_ready = false;                           // the split remainders no longer match
    ++a->val_version;
    a->extra_rc.insert(a->extra_rc.end(), rc.begin(), rc.end());
    a->extra_v.insert(a->extra_v.end(), v.begin(), v.end());
}

}  // namespace

namespace nlf {
void ensure_spmv_storage(const nlf_matrix_s* a) {
    if (a->spmv_storage_ready) return;
    nlf_matrix_s* m = const_cast<nlf_matrix_s*>(a);
    hipStream_t s = m->ctx->stream;
    finish_spmv_storage(m, s);
    NLF_HIP(hipStreamSynchronize(s));
    m->spmv_storage_ready = true;
}
}  // namespace nlf

extern "C" {

// conductivity_matrix_2d::compute / stiffness_matrix::compute (physics 0 / 1).  is_inner[rows] is
// the host mask of non-Dirichlet rows (solvers_utils.hpp:31-40).  part = NLF_PART_FULL or
// NLF_PART_LOCAL (every group LOCAL, used for the ILLT preconditioner).  Only symmetric problems
// are supported (the reference solves the asymmetric case with BiCGSTAB, out of scope here).
int nlf_assemble(nlf_mesh m, int physics, const nlf_model* models, const void* materials, const uint8_t* is_inner,
                 int symmetric, int part, nlf_matrix* out) {
    return nlf_assemble_ex(m, physics, models, materials, is_inner, symmetric, part, nullptr, out);
}

int nlf_assemble_ex(nlf_mesh m, int physics, const nlf_model* models, const void* materials, const uint8_t* is_inner,
                    int symmetric, int part, const nlf_assemble_options* opt, nlf_matrix* out) {
    if (!m) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        const bool neumann = opt && opt->neumann;
        NLF_CHECK(out && models && materials && is_inner, NLF_ERR_ARGUMENT, "null argument");
        NLF_CHECK(physics >= 0 && physics <= 2, NLF_ERR_ARGUMENT,
                  "physics must be 0 (thermal), 1 (elastic) or 2 (heat capacity)");
        if (physics == 2) part = NLF_PART_LOCAL;            // heat_capacity_matrix_2d uses local theories
        NLF_CHECK(symmetric, NLF_ERR_UNSUPPORTED, "only symmetric assembly is implemented");
        const auto t0 = std::chrono::steady_clock::now();
        hipStream_t s = m->ctx->stream;
        const int dof = physics == 1 ? 2 : 1;
        const i64 rows = dof * m->N;
        const bool only_local = part == NLF_PART_LOCAL;
        std::vector<GroupParams> gps(m->G);
        std::vector<int> gnl(m->G);
        bool any_nonlocal = false;
        for (int g = 0; g < m->G; ++g) {
            gps[g] = make_group(models[g], physics, materials, g, only_local);
            gnl[g] = gps[g].nonlocal;
            any_nonlocal |= gnl[g] != 0;
        }
        if (any_nonlocal) {
            NLF_CHECK(m->has_neighbours, NLF_ERR_STATE, "nonlocal assembly needs nlf_find_neighbours first");
            for (int g = 0; g < m->G; ++g)
                NLF_CHECK(!gnl[g] || m->searched[g] || m->gbeg_h[g] == m->gbeg_h[g + 1], NLF_ERR_STATE,
                          "a nonlocal group was not searched for neighbours");
        }
        auto mat = std::make_unique<nlf_matrix_s>();
        mat->ctx = m->ctx; mat->mesh = m; mat->dof = dof; mat->rows = rows;
        upload(mat->inner, is_inner, size_t(rows), s);
        if (neumann) {
            NLF_CHECK(physics == 0, NLF_ERR_ARGUMENT, "Neumann problems are thermal (conductivity_matrix_2d)");
            for (i64 r = 0; r < rows; ++r)
                NLF_CHECK(is_inner[r], NLF_ERR_ARGUMENT, "a Neumann problem has no first-kind boundary condition");
        }
        // coefficients per quadrature point: only where every group is local -- a nonlocal problem with
        // varying coefficients is asymmetric (problem_settings.hpp:32-34) and solved by BiCGSTAB there
        if (opt && opt->coefficients) {
            const int nq = m->fe.nq;
            std::vector<double> cq;
            for (int g = 0; g < m->G; ++g) {
                const double* v = opt->coefficients[g];
                if (!v) continue;
                NLF_CHECK(physics != 2, NLF_ERR_ARGUMENT, "the heat capacity takes no coefficients per point");
                NLF_CHECK(!any_nonlocal, NLF_ERR_UNSUPPORTED,
                          "coefficients per point make a nonlocal problem asymmetric (BiCGSTAB in the reference)");
                if (cq.empty()) cq.assign(size_t(m->E) * nq * 6, 0.0);
                int nc;
                if (physics == 0) nc = gps[g].ckind == NLF_CONDUCTIVITY_ISOTROPIC ? 1 : gps[g].ckind == NLF_CONDUCTIVITY_ORTHOTROPIC ? 2 : 3;
                else {
                    const int k = static_cast<const nlf_elastic_material*>(materials)[g].kind;
                    nc = k == NLF_ELASTIC_ISOTROPIC ? 3 : k == NLF_ELASTIC_ORTHOTROPIC ? 4 : 6;
                }
                for (i64 e = m->gbeg_h[g]; e < m->gbeg_h[g + 1]; ++e)
                    for (int q = 0; q < nq; ++q) {
                        const double* src = v + ((e - m->gbeg_h[g]) * nq + q) * nc;
                        double* dst = cq.data() + (size_t(e) * nq + q) * 6;
                        for (int k = 0; k < nc; ++k) dst[k] = src[k];
                        if (physics == 1 && nc == 3) dst[3] = src[0];       // isotropic h22 = h11
                    }
                gps[g].varying = 1;
            }
            if (!cq.empty()) upload(mat->cq, cq.data(), cq.size(), s);
        }
        DevBuf<int> d_gnl;
        upload(d_gnl, gnl.data(), gnl.size(), s);
        // ---- pattern ----
        PatternArgs pa{};
        pa.N = m->N; pa.dof = dof; pa.nn = m->nn; pa.el = m->el.p; pa.ne_off = m->ne_off.p; pa.ne_ids = m->ne_ids.p;
        pa.group_of = m->group_of.p; pa.group_nonlocal = d_gnl.p;
        pa.nb_off = m->has_neighbours ? m->nb_off.p : nullptr;
        pa.nb_ids = m->has_neighbours ? m->nb_ids.p : nullptr;
        pa.inner = mat->inner.p; pa.max_words = PATTERN_MAX_WORDS;
        mat->outer.alloc(size_t(rows + 1));
        mat->b_outer.alloc(size_t(rows + 1));
        mat->diag.alloc(size_t(rows));
        NLF_HIP(hipMemsetAsync(mat->outer.p, 0, mat->outer.bytes(), s));
        NLF_HIP(hipMemsetAsync(mat->b_outer.p, 0, mat->b_outer.bytes(), s));
        DevBuf<int> overflow{1};
        NLF_HIP(hipMemsetAsync(overflow.p, 0, sizeof(int), s));
        pa.cnt = mat->outer.p; pa.bcnt = mat->b_outer.p; pa.overflow = overflow.p;
        // thermal Q4 nonlocal assembly takes the MFMA pair path: its chunk plan needs the neighbour
        // offsets on the host and is computed while pass 0 of the pattern runs
        const bool pairs_path = physics == 0 && m->nn == 4 && any_nonlocal && !only_local && m->ctx->assembly_pairs != 0;
        PairPlan plan;
        // no nonlocal group: the columns of a node are the nodes of its own elements (k_pattern_local)
        bool local_done = false;
        if (!any_nonlocal) {
            const unsigned lb = blocks_for(m->N, PL_THREADS);
            hipLaunchKernelGGL(k_pattern_local<0>, dim3(lb), dim3(PL_THREADS), 0, s, pa);
            NLF_HIP(hipGetLastError());
            int lovf = 0;
            download(&lovf, overflow.p, 1, s);
            if (!lovf) {
                exclusive_scan_i64(mat->outer.p, rows, s);
                exclusive_scan_i64(mat->b_outer.p, rows, s);
                download(&mat->nnz_full, mat->outer.p + rows, 1, s);
                download(&mat->nnz_bound, mat->b_outer.p + rows, 1, s);
                mat->cols.alloc(size_t(std::max<i64>(mat->nnz_full, 1)));
                mat->b_cols.alloc(size_t(std::max<i64>(mat->nnz_bound, 1)));
                pa.outer = mat->outer.p; pa.b_outer = mat->b_outer.p; pa.cols = mat->cols.p; pa.b_cols = mat->b_cols.p;
                pa.diag = mat->diag.p;
                hipLaunchKernelGGL(k_pattern_local<1>, dim3(lb), dim3(PL_THREADS), 0, s, pa);
                NLF_HIP(hipGetLastError());
                local_done = true;
            } else {                                   // a node with too many column nodes: the bitmap kernels
                NLF_HIP(hipMemsetAsync(mat->outer.p, 0, mat->outer.bytes(), s));
                NLF_HIP(hipMemsetAsync(mat->b_outer.p, 0, mat->b_outer.bytes(), s));
                NLF_HIP(hipMemsetAsync(overflow.p, 0, sizeof(int), s));
            }
        }
        if (!local_done) {
            DevBuf<i32> el_min{size_t(m->E)}, lo_e{size_t(m->E)}, hi_e{size_t(m->E)};
            hipLaunchKernelGGL(k_elem_minnode, dim3(blocks_for(m->E, 256)), dim3(256), 0, s, m->E, m->nn, m->el.p, el_min.p);
            hipLaunchKernelGGL(k_elem_span, dim3(blocks_for(m->E, 4)), dim3(256), 0, s, pa, m->E, el_min.p,
                               m->el_maxnode.p, lo_e.p, hi_e.p);
            pa.lo_e = lo_e.p; pa.hi_e = hi_e.p;
            DevBuf<unsigned int> d_words{1};
            NLF_HIP(hipMemsetAsync(d_words.p, 0, sizeof(unsigned int), s));
            hipLaunchKernelGGL(k_node_words, dim3(blocks_for(m->N, 256)), dim3(256), 0, s, m->N, m->ne_off.p, m->ne_ids.p,
                               lo_e.p, hi_e.p, d_words.p);
            NLF_HIP(hipGetLastError());
            unsigned int words = 0;
            download(&words, d_words.p, 1, s);
            NLF_CHECK(int(words) <= PATTERN_MAX_WORDS, NLF_ERR_UNSUPPORTED,
                      "column node ids of a row span more than 393k ids (renumber the mesh)");
            const size_t lds_bitmap = std::max<size_t>(words, 1) * sizeof(unsigned int);
            const int pth = m->ctx->pattern_threads;
            // pass 0 keeps every node's bitmap (words per node at most the widest span) when that takes
            // at most 8 GB, so that pass 1 emits the columns without rebuilding it
            DevBuf<unsigned int> saved;
            if (size_t(m->N) * std::max<size_t>(words, 1) * sizeof(unsigned int) <= (size_t(8) << 30)) {
                saved.alloc(size_t(m->N) * std::max<size_t>(words, 1));
                pa.saved = saved.p;
                pa.stride = int(std::max<unsigned>(words, 1));
            }
            std::vector<i64> nb_off_h;
            if (pairs_path) {
                nb_off_h.resize(size_t(m->E + 1));
                download(nb_off_h.data(), m->nb_off.p, nb_off_h.size(), s);
            }
            hipLaunchKernelGGL(k_pattern<0>, dim3(unsigned(m->N)), dim3(pth), lds_bitmap, s, pa);
            NLF_HIP(hipGetLastError());
            if (pairs_path) plan = plan_pairs(m, nb_off_h, 16 * i64(sizeof(double)));
            int ovf = 0;
            download(&ovf, overflow.p, 1, s);
            NLF_CHECK(!ovf, NLF_ERR_UNSUPPORTED, "column node ids of a row span more than 393k ids (renumber the mesh)");
            exclusive_scan_i64(mat->outer.p, rows, s);
            exclusive_scan_i64(mat->b_outer.p, rows, s);
            download(&mat->nnz_full, mat->outer.p + rows, 1, s);
            download(&mat->nnz_bound, mat->b_outer.p + rows, 1, s);
            NLF_CHECK(mat->nnz_full < (i64(1) << 40), NLF_ERR_ARGUMENT, "pattern too large");
            mat->cols.alloc(size_t(std::max<i64>(mat->nnz_full, 1)));
            mat->b_cols.alloc(size_t(std::max<i64>(mat->nnz_bound, 1)));
            pa.outer = mat->outer.p; pa.b_outer = mat->b_outer.p; pa.cols = mat->cols.p; pa.b_cols = mat->b_cols.p;
            pa.diag = mat->diag.p;
            hipLaunchKernelGGL(k_pattern<1>, dim3(unsigned(m->N)), dim3(pth), lds_bitmap, s, pa);
            NLF_HIP(hipGetLastError());
            NLF_HIP(hipStreamSynchronize(s));          // before the spans and saved bitmaps are released
        }
        NLF_HIP(hipStreamSynchronize(s));
        if (any_nonlocal && !only_local && !m->nb_symmetric) symmetrize_pattern(mat.get(), s);
        const auto t1 = std::chrono::steady_clock::now();
        mat->seconds_pattern = std::chrono::duration<double>(t1 - t0).count();
        // ---- values ----
        mat->physics = physics;
        mat->only_local = only_local;
        mat->groups = gps;
        assemble_values<double>(mat.get(), s, pairs_path ? &plan : nullptr);
        if (neumann) border_neumann(mat.get(), s);
        mat->nnz_upper = (mat->nnz_full + mat->rows) / 2;
        ++mat->val_version;
        mat->seconds_values = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        *out = mat.release();
    });
}

int nlf_matrix_destroy(nlf_matrix a) {
    if (!a) return NLF_OK;
    CtxScope scope(live_ctx(a->ctx));
    delete a;
    return NLF_OK;
}

// rows, nnz of the reference's upper CSR (inner), nnz of the full storage, nnz of the bound part.
int nlf_matrix_info(nlf_matrix a, int64_t* rows, int64_t* nnz_upper, int64_t* nnz_full, int64_t* nnz_bound) {
    if (!a) return NLF_ERR_ARGUMENT;
    if (rows) *rows = a->rows;
    if (nnz_upper) *nnz_upper = a->nnz_upper;
    if (nnz_full) *nnz_full = a->nnz_full;
    if (nnz_bound) *nnz_bound = a->nnz_bound;
    return NLF_OK;
}

int nlf_matrix_timings(nlf_matrix a, double* pattern_seconds, double* values_seconds) {
    if (!a) return NLF_ERR_ARGUMENT;
    if (pattern_seconds) *pattern_seconds = a->seconds_pattern;
    if (values_seconds) *values_seconds = a->seconds_values;
    return NLF_OK;
}

// Host export in the reference's layout (finite_element_matrix::inner()/bound(): RowMajor CSR with
// int64 indices): part 0 = inner upper triangle, 1 = bound, 2 = inner full storage; parts 3 and 4
// are parts 0 and 2 of the FP32 values, widened.
// Upper CSR of rows [row0, row1) only: outer[row1 - row0 + 1] starts at 0; inner / values may be
// NULL (a first call with both NULL gives the sizes through outer).
int nlf_matrix_export_rows(nlf_matrix a, int64_t row0, int64_t row1, int64_t* outer, int64_t* inner, double* values) {
    if (!a || !outer) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        NLF_CHECK(row0 >= 0 && row0 <= row1 && row1 <= a->rows, NLF_ERR_ARGUMENT, "row range out of bounds");
        hipStream_t s = a->ctx->stream;
        const i64 n = row1 - row0;
        outer[0] = 0;
        if (n == 0) return;
        std::vector<i64> o(static_cast<size_t>(n + 1)), dg(static_cast<size_t>(n));
        download(o.data(), a->outer.p + row0, size_t(n + 1), s);
        download(dg.data(), a->diag.p + row0, size_t(n), s);
        const i64 b = dg[0], e = o[n];                 // full-storage span holding the rows' upper tails
        std::vector<i32> c;
        std::vector<double> v;
        if (inner) { c.resize(size_t(e - b)); download(c.data(), a->cols.p + b, c.size(), s); }
        if (values) { v.resize(size_t(e - b)); download(v.data(), a->val.p + b, v.size(), s); }
        i64 pos = 0;
        for (i64 r = 0; r < n; ++r) {
            for (i64 p = dg[r]; p < o[r + 1]; ++p, ++pos) {
                if (inner) inner[pos] = c[size_t(p - b)];
                if (values) values[pos] = v[size_t(p - b)];
            }
            outer[r + 1] = pos;
        }
    });
}

int nlf_matrix_export(nlf_matrix a, int part, int64_t* outer, int64_t* inner, double* values) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        hipStream_t s = a->ctx->stream;
        const i64 rows = a->rows;
        if (part == 1) {
            std::vector<i64> o(rows + 1);
            download(o.data(), a->b_outer.p, size_t(rows + 1), s);
            std::vector<i32> c(size_t(a->nnz_bound));
            download(c.data(), a->b_cols.p, c.size(), s);
            if (outer) std::copy(o.begin(), o.end(), outer);
            if (inner) for (size_t k = 0; k < c.size(); ++k) inner[k] = c[k];
            if (values) download(values, a->b_val.p, size_t(a->nnz_bound), s);
            return;
        }
        std::vector<i64> o(rows + 1), dg(rows);
        download(o.data(), a->outer.p, size_t(rows + 1), s);
        download(dg.data(), a->diag.p, size_t(rows), s);
        std::vector<i32> c(size_t(a->nnz_full));
        download(c.data(), a->cols.p, c.size(), s);
        NLF_CHECK(part >= 0 && part <= 4, NLF_ERR_ARGUMENT, "part must be 0..4");
        std::vector<double> v(size_t(a->nnz_full));
        if (part >= 3) {
            NLF_CHECK(a->val32.p, NLF_ERR_STATE, "the matrix has no FP32 values");
            std::vector<float> v32(v.size());
            download(v32.data(), a->val32.p, v32.size(), s);
            std::copy(v32.begin(), v32.end(), v.begin());
            part -= 3;
            if (part == 1) part = 2;
        } else {
            download(v.data(), a->val.p, v.size(), s);
        }
        if (part == 2) {
            if (outer) std::copy(o.begin(), o.end(), outer);
            if (inner) for (size_t k = 0; k < c.size(); ++k) inner[k] = c[k];
            if (values) std::copy(v.begin(), v.end(), values);
            return;
        }
        i64 pos = 0;
        if (outer) outer[0] = 0;
        for (i64 r = 0; r < rows; ++r) {
            for (i64 p = dg[r]; p < o[r + 1]; ++p, ++pos) {
                if (inner) inner[pos] = c[p];
                if (values) values[pos] = v[p];
            }
            if (outer) outer[r + 1] = pos;
        }
    });
}

int nlf_matrix_pattern_info(nlf_matrix a, int64_t* groups, int64_t* entries) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        ensure_spmv_storage(a);
        if (groups) *groups = a->pstart.p ? a->pat_groups : 0;
        if (entries) *entries = a->pstart.p ? a->pat_size : 0;
    });
}

// Device views of the full inner storage: outer i64[rows+1], cols i32, values f64, diag i64[rows].
int nlf_matrix_device(nlf_matrix a, const int64_t** outer, const int32_t** cols, double** values, const int64_t** diag) {
    if (!a) return NLF_ERR_ARGUMENT;
    if (outer) *outer = a->outer.p;
    if (cols) *cols = a->cols.p;
    if (values) {
        *values = a->val.p;
        a->values_exposed = true;                     // the caller may change them: derived copies are rebuilt per solve
    }
    if (diag) *diag = a->diag.p;
    return NLF_OK;
}

// Adds symmetric pair contributions (row <= col, unique pairs) to the inner part on the device.
int nlf_matrix_add_pairs(nlf_matrix a, int64_t n, const int64_t* rc_host, const double* v_host) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        if (n <= 0) return;
        NLF_CHECK(rc_host && v_host, NLF_ERR_ARGUMENT, "null argument");
        add_pairs_device(a, std::vector<i64>(rc_host, rc_host + 2 * n), std::vector<double>(v_host, v_host + n));
    });
}

// convection_condition_2d (convection_condition_2d.hpp:11-31): K(row, col) += h int N_i N_j over the
// boundary edges of convection groups, for col >= row with both rows inner (Dirichlet rows keep
// the identity, as after first_kind_matrix_fill_2d).  bcs[GB] per boundary group.
int nlf_matrix_add_convection(nlf_matrix a, const nlf_boundary* bcs) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        NLF_CHECK(a->dof == 1, NLF_ERR_ARGUMENT, "convection is a thermal condition");
        const nlf_mesh m = a->mesh;
        std::vector<uint8_t> inner(size_t(a->rows));
        download(inner.data(), a->inner.p, inner.size(), a->ctx->stream);
        std::map<std::pair<i64, i64>, double> acc;
        const BoundaryTables& t = m->bt;
        for (int b = 0; b < m->GB; ++b) {
            if (bcs[b].kind[0] != NLF_BC_CONVECTION) continue;
            for (i64 be = m->bbeg_h[b]; be < m->bbeg_h[b + 1]; ++be) {
                double jq[2];
                for (int q = 0; q < t.nq; ++q) {
                    double jx = 0, jy = 0;
                    for (int i = 0; i < t.nb; ++i) {
                        const i64 v = m->be_h[be * t.nb + i];
                        jx = jx + m->xy_h[2 * v] * t.qNxi[i * t.nq + q];
                        jy = jy + m->xy_h[2 * v + 1] * t.qNxi[i * t.nq + q];
                    }
                    jq[q] = std::sqrt(jx * jx + jy * jy);
                }
                for (int i = 0; i < t.nb; ++i) {
                    const i64 row = m->be_h[be * t.nb + i];
                    for (int j = 0; j < t.nb; ++j) {
                        const i64 col = m->be_h[be * t.nb + j];
                        if (col < row || !inner[row] || !inner[col]) continue;
                        double integral = 0.0;
                        for (int q = 0; q < t.nq; ++q) integral += t.w[q] * t.qN[i * t.nq + q] * t.qN[j * t.nq + q] * jq[q];
                        acc[{row, col}] += bcs[b].heat_transfer * integral;
                    }
                }
            }
        }
        std::vector<i64> rc;
        std::vector<double> v;
        for (const auto& [k, x] : acc) { rc.push_back(k.first); rc.push_back(k.second); v.push_back(x); }
        add_pairs_device(a, rc, v);
    });
}

// FP32 copy of the values (mixed-precision inner solves).
int nlf_matrix_make_fp32(nlf_matrix a) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        hipStream_t s = a->ctx->stream;
        if (!a->val32.p) a->val32.alloc(size_t(std::max<i64>(a->nnz_full, 1)));
        hipLaunchKernelGGL(k_to_f32, dim3(2048), dim3(256), 0, s, a->nnz_full, a->val.p, a->val32.p);
        NLF_HIP(hipGetLastError());
        a->val32_assembled = false;
        a->split_ready = false;
    });
}

int nlf_matrix_make_split(nlf_matrix a) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        hipStream_t s = a->ctx->stream;
        if (!a->val32.p) a->val32.alloc(size_t(std::max<i64>(a->nnz_full, 1)));
        if (!a->val_lo.p) a->val_lo.alloc(size_t(std::max<i64>(a->nnz_full, 1)));
        hipLaunchKernelGGL(k_to_split, dim3(2048), dim3(256), 0, s, a->nnz_full, a->val.p, a->val32.p, a->val_lo.p);
        NLF_HIP(hipGetLastError());
        a->val32_assembled = false;
        a->split_ready = true;
    });
}

// Distance of the FP32 values from the FP64 ones over the full storage: out[0] = max |v32 - v64|,
// out[1] = max |v64|, out[2] = ||v32 - v64||_2, out[3] = ||v64||_2.
int nlf_matrix_compare_fp32(nlf_matrix a, double* out) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        NLF_CHECK(out, NLF_ERR_ARGUMENT, "null argument");
        NLF_CHECK(a->val32.p, NLF_ERR_STATE, "the matrix has no FP32 values");
        hipStream_t s = a->ctx->stream;
        constexpr int B = 1024;
        DevBuf<double> part{size_t(4 * B)};
        hipLaunchKernelGGL(k_compare32, dim3(B), dim3(256), 0, s, a->nnz_full, a->val.p, a->val32.p, part.p);
        NLF_HIP(hipGetLastError());
        std::vector<double> h(4 * B);
        download(h.data(), part.p, h.size(), s);
        double m0 = 0.0, m1 = 0.0, s0 = 0.0, s1 = 0.0;
        for (int b = 0; b < B; ++b) {
            m0 = std::max(m0, h[4 * b]); m1 = std::max(m1, h[4 * b + 1]);
            s0 += h[4 * b + 2]; s1 += h[4 * b + 3];
        }
        out[0] = m0; out[1] = m1; out[2] = std::sqrt(s0); out[3] = std::sqrt(s1);
    });
}

// FP32 values of the same operator assembled in FP32: the element-pair contractions (MFMA for the
// thermal Q4 blocks) and the accumulation run in FP32; Jacobian determinants, coordinates and the
// cutoff decisions of phi stay FP64.  Later boundary additions are repeated.
int nlf_matrix_assemble_fp32(nlf_matrix a) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        NLF_CHECK(!a->combined, NLF_ERR_STATE, "the matrix was combined with another one (nlf_matrix_combine)");
        NLF_CHECK(!a->neumann, NLF_ERR_UNSUPPORTED, "the FP32 value pass does not repeat the Neumann border (use nlf_matrix_make_fp32)");
        NLF_CHECK(a->val.p && !a->groups.empty(), NLF_ERR_STATE, "the matrix has no FP64 assembly to repeat");
        hipStream_t s = a->ctx->stream;
        assemble_values<float>(a, s);
        if (!a->extra_v.empty()) {
            DevBuf<i64> d_rc;
            DevBuf<double> d_v;
            upload(d_rc, a->extra_rc.data(), a->extra_rc.size(), s);
            upload(d_v, a->extra_v.data(), a->extra_v.size(), s);
            const i64 n = i64(a->extra_v.size());
            hipLaunchKernelGGL(k_add_pairs<float>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, d_rc.p, d_v.p,
                               a->outer.p, a->cols.p, a->val32.p);
            NLF_HIP(hipGetLastError());
            NLF_HIP(hipStreamSynchronize(s));
        }
        a->val32_assembled = true;
        a->split_ready = false;
    });
}

// Boundary data on the host (O(boundary)): second-kind RHS f (boundary_condition_second_kind_2d.hpp:12-33,
// convection h*T_a), Dirichlet values xD (first wins) and the inner mask (solvers_utils.hpp:31-40).
int nlf_boundary_data_fn(nlf_mesh m, int dof, const nlf_boundary* bcs, nlf_boundary_value_fn value, void* user,
                         double* f, double* xD, uint8_t* is_inner) {
    if (!m || !bcs || (dof != 1 && dof != 2)) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        const i64 rows = dof * m->N;
        const BoundaryTables& t = m->bt;
        if (is_inner) std::fill(is_inner, is_inner + rows, uint8_t(1));
        if (xD) std::fill(xD, xD + rows, 0.0);
        if (f) std::fill(f, f + rows, 0.0);
        // first kind: the value at the node; the first group giving a node a nonzero value keeps it
        // (boundary_condition_first_kind_2d.hpp:17-26)
        for (int b = 0; b < m->GB; ++b)
            for (int d = 0; d < dof; ++d) {
                if (bcs[b].kind[d] != NLF_BC_FIRST_KIND) continue;
                for (i64 be = m->bbeg_h[b]; be < m->bbeg_h[b + 1]; ++be)
                    for (int i = 0; i < t.nb; ++i) {
                        const i64 v = m->be_h[be * t.nb + i];
                        if (xD && xD[dof * v + d] == 0.0)
                            xD[dof * v + d] = value ? value(m->xy_h[2 * v], m->xy_h[2 * v + 1], b, d, user) : bcs[b].value[d];
                        if (is_inner) is_inner[dof * v + d] = 0;
                    }
            }
        if (!f) return;
        // second kind: int_edge N_i g |J| over the edge's Gauss points, g at the point
        // (boundary_condition_second_kind_2d.hpp:16-23); convection adds h T_a the same way
        for (int b = 0; b < m->GB; ++b)
            for (int d = 0; d < dof; ++d) {
                const int kind = bcs[b].kind[d];
                const bool flux = kind == NLF_BC_FLUX;
                if (!flux && !(dof == 1 && kind == NLF_BC_CONVECTION)) continue;
                const double g0 = flux ? bcs[b].value[d] : bcs[b].heat_transfer * bcs[b].ambient;
                for (i64 be = m->bbeg_h[b]; be < m->bbeg_h[b + 1]; ++be) {
                    double jq[2], gq[2];
                    for (int q = 0; q < t.nq; ++q) {
                        double jx = 0, jy = 0, px = 0, py = 0;
                        for (int i = 0; i < t.nb; ++i) {
                            const i64 v = m->be_h[be * t.nb + i];
                            jx = jx + m->xy_h[2 * v] * t.qNxi[i * t.nq + q];
                            jy = jy + m->xy_h[2 * v + 1] * t.qNxi[i * t.nq + q];
                            px = px + m->xy_h[2 * v] * t.qN[i * t.nq + q];
                            py = py + m->xy_h[2 * v + 1] * t.qN[i * t.nq + q];
                        }
                        jq[q] = std::sqrt(jx * jx + jy * jy);
                        gq[q] = flux && value ? value(px, py, b, d, user) : g0;
                    }
                    for (int i = 0; i < t.nb; ++i) {
                        const i64 v = m->be_h[be * t.nb + i];
                        double integral = 0.0;
                        for (int q = 0; q < t.nq; ++q) integral += t.w[q] * t.qN[i * t.nq + q] * gq[q] * jq[q];
                        f[dof * v + d] += integral;
                    }
                }
            }
    });
}

int nlf_boundary_data(nlf_mesh m, int dof, const nlf_boundary* bcs, double* f, double* xD, uint8_t* is_inner) {
    return nlf_boundary_data_fn(m, dof, bcs, nullptr, nullptr, f, xD, is_inner);
}

}  // extern "C"
