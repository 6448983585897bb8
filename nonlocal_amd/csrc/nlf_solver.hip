// Krylov layer: CSR SpMV, CG (conjugate_gradient.hpp:48-79), Jacobi and ILLT preconditioners
// (eigen_preconditioner.hpp:50-57), FP32-inner / FP64-refinement mixed mode.
//
// SpMV runs on the full (both-triangle) storage so every output row is a gather (no scatter, no
// atomics, deterministic); LPR lanes cooperate on one row.  CG keeps its scalars on the device in a
// state block; each iteration is 5 launches (SpMV+dot, alpha, update+precond+dot, beta, p update)
// with no host synchronisation; the host reads the state every few iterations.
//
// ILLT: IC(0) of the local matrix, diagonal scaling s = 1/sqrt|a_ii| and shift retries (our choice:
// Eigen::IncompleteCholesky is not vendored, so its factor is not reproduced).  Factorisation and
// both triangular solves are level-scheduled: rows are grouped by dependency level (computed once on
// the host from the pattern) and one persistent 1024-thread workgroup sweeps the levels with a
// barrier between them.
#include "nlf_internal.hpp"

#include <mutex>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <cmath>
#include <memory>
#include <numeric>
#include <type_traits>

namespace nlf {

constexpr int SPMV_THREADS = 256;
constexpr int SPMV_PAIRS_IN_FLIGHT = 4;  // nonzero pairs each lane of the pair kernel loads per step (8: no faster)
constexpr int LEVEL_THREADS = 1024;

struct CGState {
    double rz, pq, alpha, beta, bnorm, res, tol;
    int done, indefinite;                // indefinite: p'Ap < 0 is not a breakdown (bordered Neumann systems)
    i64 it, max_it;
    int breakdown, pad2;
};

template<class T> __device__ __forceinline__ T shfl_down_t(T v, int o) { return __shfl_down(v, o); }

template<class Acc>
__device__ __forceinline__ Acc block_sum(Acc v, Acc* sh) {
    // fixed-order tree reduction (deterministic)
    for (int o = 32; o > 0; o >>= 1) v += shfl_down_t(v, o);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    Acc s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < int(blockDim.x >> 6); ++i) s += sh[i];
    __syncthreads();
    return s;
}

// y = A x (rows in [0, rows)), optionally partial[b] = sum_r x_r y_r over the block's rows.  LPR
// lanes per row, grid-stride over row groups (the grid is capped, so DOT partials stay few).
// Columns are absolute (cols) or, when off16 is given, 16-bit offsets from the row (col = r + off):
// 10 instead of 12 bytes per FP64 nonzero.  Each lane keeps four (value, column) pairs and then
// four x gathers in flight.
template<class T>
__device__ __forceinline__ T stream_load(const T* p, bool nt) { return nt ? __builtin_nontemporal_load(p) : *p; }

// Split values: hi (FP32) + lo * 2^-16 ulp(hi), lo a 16-bit integer (40 significand bits).
__device__ __forceinline__ double split_value(float h, int16_t l) {
    unsigned e = (__float_as_uint(h) >> 23) & 0xffu;
    e = e ? e : 1u;
    return double(h) + double(l) * __longlong_as_double(static_cast<long long>(e + 857u) << 52);
}

// SPLIT (V = float): each value is split_value(val[p], lo[p]), summed in Acc.
template<int LPR, class V, class X, class Acc, bool DOT, bool XCD = false, int UNROLL = 4, bool NT = false,
         bool SPLIT = false>
__global__ __launch_bounds__(SPMV_THREADS) void k_spmv(i64 rows, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                       const int16_t* __restrict__ off16, const i64* __restrict__ pstart,
                                                       const V* __restrict__ val, const int16_t* __restrict__ lo,
                                                       const X* __restrict__ x, X* __restrict__ y,
                                                       double* __restrict__ partial, const CGState* __restrict__ st) {
    __shared__ double sh[SPMV_THREADS / 64];
    if (st && st->done) return;
    constexpr int RPB = SPMV_THREADS / LPR;
    const int sub = threadIdx.x % LPR;
    double d = 0.0;
    // XCD: workgroups are dealt round-robin to the 8 XCDs; renumbering them so that each XCD sweeps
    // a contiguous range of row groups keeps the x window of neighbouring rows in one L2
    i64 lb = blockIdx.x;
    if (XCD && gridDim.x % 8 == 0) lb = i64(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
    const i64 rows_r = (rows + RPB - 1) / RPB * RPB;
    for (i64 r0 = lb * RPB; r0 < rows_r; r0 += i64(gridDim.x) * RPB) {
        const i64 r = r0 + threadIdx.x / LPR;
        Acc acc = 0;
        if (r < rows) {
            const i64 b = outer[r], e = outer[r + 1];
            // offsets of row r: off16[ob + p] (ob = 0 for per-nonzero offsets; with the row-pattern
            // dictionary ob moves p into the row's shared pattern, which stays in L2)
            const i64 ob = pstart ? pstart[r] : 0;
            const bool nt_off = NT && !pstart;
            i64 p = b + sub;
            for (; p + (UNROLL - 1) * LPR < e; p += UNROLL * LPR) {
                i64 c[UNROLL];
                V v[UNROLL];
                int16_t vl[UNROLL];
#pragma unroll
                for (int t = 0; t < UNROLL; ++t) {
                    c[t] = off16 ? r + i64(stream_load(off16 + ob + p + t * LPR, nt_off))
                                 : i64(stream_load(cols + p + t * LPR, NT));
                    v[t] = stream_load(val + p + t * LPR, NT);
                    if constexpr (SPLIT) vl[t] = lo[p + t * LPR];
                }
                X xv[UNROLL];
#pragma unroll
                for (int t = 0; t < UNROLL; ++t) xv[t] = x[c[t]];
#pragma unroll
                for (int t = 0; t < UNROLL; ++t) {
                    if constexpr (SPLIT) acc += Acc(split_value(v[t], vl[t])) * Acc(xv[t]);
                    else acc += Acc(v[t]) * Acc(xv[t]);
                }
            }
            for (; p < e; p += LPR) {
                const i64 c = off16 ? r + i64(stream_load(off16 + ob + p, nt_off)) : i64(stream_load(cols + p, NT));
                if constexpr (SPLIT) acc += Acc(split_value(stream_load(val + p, NT), lo[p])) * Acc(x[c]);
                else acc += Acc(stream_load(val + p, NT)) * Acc(x[c]);
            }
        }
        for (int o = LPR / 2; o > 0; o >>= 1) acc += shfl_down_t(acc, o);   // lane sub==0 of each group holds the sum
        if (r < rows && sub == 0) {
            y[r] = X(acc);
            if (DOT) d += double(x[r]) * double(X(acc));
        }
    }
    if (DOT) {
        d = block_sum<double>(d, sh);
        if (threadIdx.x == 0) partial[blockIdx.x] = d;
    }
}

// Pair SpMV over 16-bit offsets (per nonzero, or the row-pattern dictionary whose row starts are
// even, see nlf_rowpat.hip): each lane takes two consecutive nonzeros per load, so one 2-value load
// and one 4-byte offset load replace two of each; two pairs stay in flight per lane.  A row that
// starts at an odd index handles its first nonzero apart, and an odd last one at the end, so
// every pair load is aligned.  The products are summed in the same order by every launch of a
// given LPR (results are reproducible, not bitwise equal to k_spmv's).
template<class V> struct Pair;
template<> struct Pair<double> { typedef double T __attribute__((ext_vector_type(2))); };
template<> struct Pair<float> { typedef float T __attribute__((ext_vector_type(2))); };

// XCD: each block takes a contiguous chunk of row groups, and the blocks that the hardware deals
// round-robin to one XCD take adjacent chunks, so each XCD's L2 holds the x window of its own
// eighth of the rows (without it, concurrently active rows span the whole matrix on every XCD).
template<int LPR, class V, class X, class Acc, bool DOT, bool SPLIT = false, bool XCD = false, int PF = SPMV_PAIRS_IN_FLIGHT>
__global__ __launch_bounds__(SPMV_THREADS) void k_spmv2(i64 rows, const i64* __restrict__ outer,
                                                        const int16_t* __restrict__ off, const i64* __restrict__ pstart,
                                                        const V* __restrict__ val, const int16_t* __restrict__ lo,
                                                        const X* __restrict__ x, X* __restrict__ y,
                                                        double* __restrict__ partial, const CGState* __restrict__ st) {
    typedef typename Pair<V>::T V2;
    __shared__ double sh[SPMV_THREADS / 64];
    if (st && st->done) return;
    constexpr int RPB = SPMV_THREADS / LPR;
    const int sub = threadIdx.x % LPR;
    double d = 0.0;
    auto value = [&](V v, int16_t l) -> Acc {
        if constexpr (SPLIT) return Acc(split_value(v, l));
        else return Acc(v);
    };
    auto lo_pair = [&](i64 p) -> int {
        if constexpr (SPLIT) return *reinterpret_cast<const int*>(lo + p);
        else return 0;
    };
    const i64 groups = (rows + RPB - 1) / RPB;
    i64 g_begin = blockIdx.x, g_end = groups, g_step = gridDim.x;
    if (XCD) {
        const i64 lb = i64(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
        const i64 per = (groups + gridDim.x - 1) / gridDim.x;
        g_begin = lb * per;
        g_end = std::min<i64>(groups, g_begin + per);
        g_step = 1;
    }
    for (i64 grp = g_begin; grp < g_end; grp += g_step) {
        const i64 r = grp * RPB + threadIdx.x / LPR;
        Acc acc = 0;
        if (r < rows) {
            const i64 b = outer[r], e = outer[r + 1];
            const int16_t* o = off + (pstart ? pstart[r] : 0);
            if ((b & 1) && sub == 0 && b < e)
                acc += value(val[b], SPLIT ? lo[b] : int16_t(0)) * Acc(x[r + o[b]]);
            i64 p = ((b + 1) & ~i64(1)) + 2 * sub;
            // PF pairs in flight per lane; a lane's pairs are summed in order whatever PF is
            for (; p + 2 * LPR * (PF - 1) + 1 < e; p += 2 * LPR * PF) {
                V2 v[PF];
                int cw[PF], lw[PF];
#pragma unroll
                for (int u = 0; u < PF; ++u) {
                    v[u] = *reinterpret_cast<const V2*>(val + p + 2 * LPR * u);
                    cw[u] = *reinterpret_cast<const int*>(o + p + 2 * LPR * u);
                    lw[u] = lo_pair(p + 2 * LPR * u);
                }
                X xs[2 * PF];
#pragma unroll
                for (int u = 0; u < PF; ++u) {
                    xs[2 * u] = x[r + int16_t(cw[u] & 0xffff)];
                    xs[2 * u + 1] = x[r + int16_t(cw[u] >> 16)];
                }
#pragma unroll
                for (int u = 0; u < PF; ++u) {
                    acc += value(v[u].x, int16_t(lw[u] & 0xffff)) * Acc(xs[2 * u]);
                    acc += value(v[u].y, int16_t(lw[u] >> 16)) * Acc(xs[2 * u + 1]);
                }
            }
            for (; p + 1 < e; p += 2 * LPR) {
                const V2 v0 = *reinterpret_cast<const V2*>(val + p);
                const int c0 = *reinterpret_cast<const int*>(o + p);
                const int l0 = lo_pair(p);
                acc += value(v0.x, int16_t(l0 & 0xffff)) * Acc(x[r + int16_t(c0 & 0xffff)]);
                acc += value(v0.y, int16_t(l0 >> 16)) * Acc(x[r + int16_t(c0 >> 16)]);
            }
            if (p < e) acc += value(val[p], SPLIT ? lo[p] : int16_t(0)) * Acc(x[r + o[p]]);
        }
        for (int k = LPR / 2; k > 0; k >>= 1) acc += shfl_down_t(acc, k);
        if (r < rows && sub == 0) {
            y[r] = X(acc);
            if (DOT) d += double(x[r]) * double(X(acc));
        }
    }
    if (DOT) {
        d = block_sum<double>(d, sh);
        if (threadIdx.x == 0) partial[blockIdx.x] = d;
    }
}

__global__ void k_sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
    __shared__ double sh[1024 / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
    s = block_sum<double>(s, sh);
    if (threadIdx.x == 0) *out = s;
}

// alpha = rz / (p . A p)
__global__ void k_alpha(const double* __restrict__ partial, int n, CGState* __restrict__ st) {
    __shared__ double sh[1024 / 64];
    if (st->done) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
    s = block_sum<double>(s, sh);
    if (threadIdx.x == 0) {
        st->pq = s;
        st->alpha = st->rz / s;
        const bool bad_curvature = st->indefinite ? s == 0.0 : !(s > 0.0);
        if (bad_curvature || !isfinite(st->alpha)) { st->breakdown = 1; st->done = 1; }
    }
}

// x += alpha p; r -= alpha q; z = M r (identity / Jacobi, or left for a separate apply); partial rz.
template<class X, bool JACOBI, bool DOT>
__global__ __launch_bounds__(SPMV_THREADS) void k_update_xr(i64 n, X* __restrict__ x, X* __restrict__ r, const X* __restrict__ p,
                                                            const X* __restrict__ q, X* __restrict__ z,
                                                            const X* __restrict__ invd, double* __restrict__ partial,
                                                            const CGState* __restrict__ st) {
    __shared__ double sh[SPMV_THREADS / 64];
    if (st->done) return;
    const X a = X(st->alpha);
    double d = 0.0;
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) {
        x[i] += a * p[i];
        const X ri = r[i] - a * q[i];
        r[i] = ri;
        if (z) {
            const X zi = JACOBI ? invd[i] * ri : ri;
            z[i] = zi;
            if (DOT) d += double(ri) * double(zi);
        }
    }
    if (DOT) {
        d = block_sum<double>(d, sh);
        if (threadIdx.x == 0) partial[blockIdx.x] = d;
    }
}

template<class X>
__global__ __launch_bounds__(SPMV_THREADS) void k_dot(i64 n, const X* __restrict__ a, const X* __restrict__ b,
                                                      double* __restrict__ partial, const CGState* __restrict__ st) {
    __shared__ double sh[SPMV_THREADS / 64];
    if (st && st->done) return;
    double d = 0.0;
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x)
        d += double(a[i]) * double(b[i]);
    d = block_sum<double>(d, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = d;
}

// mu = rz_new / rz; residual = sqrt(rz)/||b||; stop test of the reference.
__global__ void k_beta(const double* __restrict__ partial, int n, CGState* __restrict__ st) {
    __shared__ double sh[1024 / 64];
    if (st->done) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
    s = block_sum<double>(s, sh);
    if (threadIdx.x == 0) {
        st->beta = s / st->rz;
        st->rz = s;
        st->it += 1;
        st->res = sqrt(fabs(s)) / st->bnorm;
        if (!isfinite(st->res)) { st->breakdown = 1; st->done = 1; }
        if (st->res <= st->tol || st->it >= st->max_it) st->done = 1;
    }
}

template<class X>
__global__ void k_update_p(i64 n, X* __restrict__ p, const X* __restrict__ z, const CGState* __restrict__ st) {
    if (st->done) return;
    const X b = X(st->beta);
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) p[i] = z[i] + b * p[i];
}

// CG start: rz = r.z, bnorm from partials, residual, done.
__global__ void k_init_state(const double* __restrict__ prz, const double* __restrict__ pbb, int n, double tol, i64 max_it,
                             int indefinite, CGState* __restrict__ st) {
    __shared__ double sh[1024 / 64];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) { a += prz[i]; b += pbb[i]; }
    a = block_sum<double>(a, sh);
    b = block_sum<double>(b, sh);
    if (threadIdx.x == 0) {
        st->rz = a;
        st->bnorm = sqrt(b);
        st->tol = tol;
        st->it = 0;
        st->max_it = max_it;
        st->breakdown = 0;
        st->indefinite = indefinite;
        st->res = st->bnorm > 0.0 ? sqrt(fabs(a)) / st->bnorm : 0.0;
        st->done = (st->res <= tol || max_it <= 0) ? 1 : 0;
    }
}

template<class A, class B>
__global__ void k_axpby(i64 n, A alpha, const B* __restrict__ x, A beta, B* __restrict__ y) {   // y = alpha x + beta y
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x)
        y[i] = B(alpha * A(x[i]) + beta * A(y[i]));
}

// y += x (FP32 correction accumulated into the FP64 solution)
template<class S, class Y>
__global__ void k_add_to(i64 n, const S* __restrict__ x, Y* __restrict__ y) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) y[i] += Y(x[i]);
}

// rz from partial sums; `first` also initialises the state of a reliable-update solve (no
// convergence test inside: the host decides on the FP64 true residual).
__global__ void k_set_rz(const double* __restrict__ partial, int n, int first, i64 max_it, int indefinite,
                         CGState* __restrict__ st) {
    __shared__ double sh[1024 / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
    s = block_sum<double>(s, sh);
    if (threadIdx.x == 0) {
        st->rz = s;
        st->done = 0;
        if (first) {
            st->it = 0;
            st->max_it = max_it;
            st->tol = -1.0;
            st->bnorm = 1.0;
            st->breakdown = 0;
            st->indefinite = indefinite;
        }
        st->res = sqrt(fabs(s));
    }
}

template<class A, class B>
__global__ void k_convert(i64 n, const A* __restrict__ x, B* __restrict__ y) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) y[i] = B(x[i]);
}

template<class X>
__global__ void k_scale(i64 n, const X* __restrict__ s, const X* __restrict__ x, X* __restrict__ y) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) y[i] = s[i] * x[i];
}

// r = b - A x (full storage, one wave per row) and partial sums of r.r and b.b.
template<class V, class X>
__global__ __launch_bounds__(SPMV_THREADS) void k_residual(i64 rows, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                           const V* __restrict__ val, const X* __restrict__ x,
                                                           const X* __restrict__ b, X* __restrict__ r,
                                                           double* __restrict__ prr, double* __restrict__ pbb) {
    __shared__ double sh[SPMV_THREADS / 64];
    const int lane = threadIdx.x & 63;
    constexpr int RPB = SPMV_THREADS / 64;
    double rr = 0.0, bb = 0.0;
    const i64 rows_r = (rows + RPB - 1) / RPB * RPB;
    for (i64 r0 = i64(blockIdx.x) * RPB; r0 < rows_r; r0 += i64(gridDim.x) * RPB) {     // one wave per row
        const i64 row = r0 + (threadIdx.x >> 6);
        double acc = 0.0;
        if (row < rows)
            for (i64 p = outer[row] + lane; p < outer[row + 1]; p += 64) acc += double(val[p]) * double(x[cols[p]]);
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
        if (row < rows && lane == 0) {
            const double ri = double(b[row]) - acc;
            r[row] = X(ri);
            rr += ri * ri;
            bb += double(b[row]) * double(b[row]);
        }
    }
    rr = block_sum<double>(rr, sh);
    bb = block_sum<double>(bb, sh);
    if (threadIdx.x == 0) { prr[blockIdx.x] = rr; pbb[blockIdx.x] = bb; }
}

// f -= K_bound x_D ; f[D] = x_D   (boundary_condition_first_kind_2d.hpp:50-59)
__global__ void k_first_kind(i64 rows, const i64* __restrict__ bo, const i32* __restrict__ bc, const double* __restrict__ bv,
                             const uint8_t* __restrict__ inner, const double* __restrict__ xD, double* __restrict__ f) {
    const i64 r = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double s = 0.0;
    for (i64 p = bo[r]; p < bo[r + 1]; ++p) s += bv[p] * xD[bc[p]];
    f[r] = inner[r] ? f[r] - s : xD[r];
}

__global__ void k_inv_diag(i64 rows, const i64* __restrict__ diag, const double* __restrict__ val, double* __restrict__ inv,
                           float* __restrict__ inv32) {
    const i64 r = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const double d = val[diag[r]];
    const double v = d != 0.0 ? 1.0 / d : 1.0;      // Eigen::DiagonalPreconditioner: 1 for zero diagonals
    inv[r] = v;
    inv32[r] = float(v);
}

// ---- IC(0) ------------------------------------------------------------------------------------
// One independent diagonal block of a block-diagonal factor, factored and swept by a workgroup of
// its own: its level count, sweep window depth and first level offset (into the shared offsets).
struct SweepBlock {
    i32 levels, K, lptr_off, pad;
};

struct LevelArgs {
    i64 rows;
    int levels;
    const i64* outer;
    const i32* cols;
    const i64* diag;
    const i64* tpos;       // position of the transposed entry
    const i32* lrows;      // rows sorted by level
    const i32* lptr;       // level offsets
};

__global__ void k_ic_init(i64 rows, const i64* __restrict__ outer, const i32* __restrict__ cols, const i64* __restrict__ diag,
                          const double* __restrict__ a, const double* __restrict__ s, double shift, double* __restrict__ F) {
    const i64 r = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    for (i64 p = outer[r]; p < outer[r + 1]; ++p) F[p] = s[r] * a[p] * s[cols[p]];
    F[diag[r]] += shift;
}

__global__ void k_scale_from_diag(i64 rows, const i64* __restrict__ diag, const double* __restrict__ a, double* __restrict__ s,
                                  float* __restrict__ s32) {
    const i64 r = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const double d = fabs(a[diag[r]]);
    s[r] = d > 0.0 ? 1.0 / sqrt(d) : 1.0;
    s32[r] = float(s[r]);
}

// Up-looking IC(0): for row k (all rows i < k it depends on are done and mirrored into row k's
// lower positions), U_kk = sqrt(a_kk - sum_i U_ik^2), U_kj = (a_kj - sum_{i<k} U_ik U_ij) / U_kk.
// blk: one workgroup per independent diagonal block, over the block's own levels (a.lrows block
// after block, a.lptr each block's level offsets from blk[b].lptr_off on).
__global__ __launch_bounds__(LEVEL_THREADS) void k_ic_factor(LevelArgs a, const SweepBlock* __restrict__ blk,
                                                             double* __restrict__ F, int* __restrict__ fail) {
    if (blk) {
        const SweepBlock b = blk[blockIdx.x];
        a.levels = b.levels;
        a.lptr += b.lptr_off;
    }
    for (int L = 0; L < a.levels; ++L) {
        for (i32 t = a.lptr[L] + threadIdx.x; t < a.lptr[L + 1]; t += blockDim.x) {
            const i64 k = a.lrows[t];
            const i64 b = a.outer[k], dk = a.diag[k], e = a.outer[k + 1];
            double d = F[dk];
            for (i64 p = b; p < dk; ++p) d -= F[p] * F[p];
            if (!(d > 0.0) || !isfinite(d)) { *fail = 1; d = 1.0; }
            const double ukk = sqrt(d);
            F[dk] = ukk;
            for (i64 p = dk + 1; p < e; ++p) {
                const i32 j = a.cols[p];
                double v = F[p];
                // intersection of lower(k) and lower(j) restricted to i < k
                i64 pj = a.outer[j];
                const i64 ej = a.diag[j];
                for (i64 pk = b; pk < dk; ++pk) {
                    const i32 i = a.cols[pk];
                    while (pj < ej && a.cols[pj] < i) ++pj;
                    if (pj < ej && a.cols[pj] == i) v -= F[pk] * F[pj];
                }
                v /= ukk;
                F[p] = v;
                F[a.tpos[p]] = v;
            }
        }
        __syncthreads();
    }
}

// One triangular sweep in level order.  Records (one per row) are stored contiguously in level
// order so a level's loads are coalesced; a dependency on a row solved at most K levels earlier is
// read from an LDS window of K x W slots (slot = (level mod K) * W + position in level), otherwise
// (dep < 0) from global memory.
template<class X>
struct PackedSweep {
    int levels, K, W;
    const i32* lptr;      // [levels+1] record offsets
    const i32* row;       // row of each record
    const i32* off;       // [records+1] entry offsets
    const i32* dep;       // window slot, or -(row+1)
    const X* val;         // factor entries
    const X* dinv;        // 1 / U_kk
    const X* scale;       // s_k
};

// z = S U^-1 U^-T S r : forward U^T y = S r, backward U z' = y, z = S z'.  One workgroup.
template<class X>
__global__ __launch_bounds__(LEVEL_THREADS) void k_ic_solve(PackedSweep<X> fw, PackedSweep<X> bw, const X* __restrict__ r,
                                                            X* __restrict__ y, X* __restrict__ z,
                                                            const CGState* __restrict__ st) {
    extern __shared__ unsigned char smem_raw[];
    X* win = reinterpret_cast<X*>(smem_raw);
    if (st && st->done) return;
    for (int L = 0; L < fw.levels; ++L) {
        const i32 b = fw.lptr[L], cnt = fw.lptr[L + 1] - b;
        for (i32 t = threadIdx.x; t < cnt; t += blockDim.x) {
            const i32 rec = b + t;
            const i32 k = fw.row[rec];
            X v = fw.scale[rec] * r[k];
            for (i32 e = fw.off[rec]; e < fw.off[rec + 1]; ++e) {
                const i32 d = fw.dep[e];
                v -= fw.val[e] * (d >= 0 ? win[d] : y[-d - 1]);
            }
            v *= fw.dinv[rec];
            if (fw.K) win[(L % fw.K) * fw.W + t] = v;
            y[k] = v;
        }
        __syncthreads();
    }
    for (int L = 0; L < bw.levels; ++L) {
        const i32 b = bw.lptr[L], cnt = bw.lptr[L + 1] - b;
        for (i32 t = threadIdx.x; t < cnt; t += blockDim.x) {
            const i32 rec = b + t;
            const i32 k = bw.row[rec];
            X v = y[k];
            for (i32 e = bw.off[rec]; e < bw.off[rec + 1]; ++e) {
                const i32 d = bw.dep[e];
                v -= bw.val[e] * (d >= 0 ? win[d] : y[-d - 1]);
            }
            v *= bw.dinv[rec];
            if (bw.K) win[(L % bw.K) * bw.W + t] = v;
            y[k] = v;                  // z' (row k of the backward sweep reads only its own y[k])
            z[k] = bw.scale[rec] * v;
        }
        __syncthreads();
    }
}

// ELL form of one sweep for the fast path: every level is a block of `cnt` records with exactly WE
// entry slots each, stored record-major (entry j of record rec at WE * rec + j), so that a lane
// loads a record's values and dependency slots with 16-byte loads and a wave's loads of one level
// cover one contiguous span; padding slots carry value 0 and a dependency on slot 0 of the
// (zero-initialised) window.
template<class X>
struct EllSweep {
    int levels, K, W;
    const i32* lptr;
    const i32* row;
    const i32* dep;          // window slot (K > 0) or -(row + 1) of a dependency read from the output
    const X* val;
    const X* dinv;
};


// rhs_lvl[rec] = scale[rec] * r[row[rec]] (forward) or y[row[rec]] (backward, scale == nullptr);
// also clears `clear` (the sweep's output, read through global dependencies).
template<class X>
__global__ void k_sweep_gather(i64 n, const i32* __restrict__ row, const X* __restrict__ scale, const X* __restrict__ src,
                               X* __restrict__ rhs_lvl, X* __restrict__ clear, const CGState* __restrict__ st) {
    if (st && st->done) return;
    for (i64 t = i64(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += i64(gridDim.x) * blockDim.x) {
        const i32 k = row[t];
        rhs_lvl[t] = scale ? scale[t] * src[k] : src[k];
        if (clear) clear[t] = X(0);
    }
}

// One level-scheduled triangular sweep by one workgroup.  The first `ct` threads (ct = blockDim - 64
// SW, at least the widest level) compute: thread t owns record t of every level, and the record data
// of level L + D (right-hand side, 1/U_kk, the WE entries and their dependency slots) are loaded D
// levels ahead into a ring of D register slots, while the level offsets and every dependency live in
// LDS; a level's critical path is the LDS reads of its dependencies, a few FMAs, the LDS write and
// the barrier.  The last SW waves store: after the barrier of level L they copy the level's values
// from the LDS window to out_lvl (level order), while level L + 1 is computed.
//
// Why the split: a wave's vector-memory counter covers loads and stores alike, and loads and stores
// may complete out of order relative to each other, so while a wave has stores in flight any wait
// for one of its loads must wait for everything (vmcnt 0), which drains the prefetch ring.  Compute
// waves that never store wait for exactly the level they need.  For the same reason no load of the
// level loop goes through a generic (flat) address.  The arithmetic of a record is unchanged; the
// row-order results come from k_sweep_scatter.
constexpr size_t SWEEP_LDS_BYTES = 64 * 1024;
// Compute widths of the two prefetching sweeps and their store waves (workgroups of CT + 64 SW).
#ifndef NLF_SWEEP_SW_SMALL
#define NLF_SWEEP_SW_SMALL 2
#endif
#ifndef NLF_SWEEP_SW_LARGE
#define NLF_SWEEP_SW_LARGE 2
#endif
constexpr int SWEEP_SW_SMALL = NLF_SWEEP_SW_SMALL, SWEEP_SW_LARGE = NLF_SWEEP_SW_LARGE;
constexpr int SWEEP_CT_SMALL = 512, SWEEP_CT_LARGE = 1024 - 64 * SWEEP_SW_LARGE;

// Levels the ring holds, sized so that the ring and the rest of a compute lane fit the registers a
// lane has with CT compute lanes plus the store waves per workgroup (168 VGPRs up to 12 waves, 128
// at 16) without spilling.
template<int WE, int CT>
constexpr int sweep_depth() {
    return CT <= SWEEP_CT_SMALL ? (WE <= 4 ? 8 : WE <= 8 ? 4 : WE <= 16 ? 2 : 1) : (WE <= 4 ? 4 : WE <= 8 ? 2 : 1);
}

// N consecutive elements of type T by 16-byte loads (8-byte when N * sizeof(T) is 8); p is aligned
// to the load size
template<class T, int N>
__device__ __forceinline__ void load16(const T* __restrict__ p, T (&v)[N]) {
    constexpr int B = int(N * sizeof(T));
    static_assert(B == 8 || B % 16 == 0, "record slots are whole 8- or 16-byte words");
    if constexpr (B == 8) {
        const uint2 w = *reinterpret_cast<const uint2*>(p);
        __builtin_memcpy(v, &w, 8);
    } else {
        const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
        for (int k = 0; k < B / 16; ++k) {
            const uint4 w = q[k];
            __builtin_memcpy(reinterpret_cast<unsigned char*>(v) + 16 * k, &w, 16);
        }
    }
}

template<class X, int WE, int CT, int SW>
__global__ __launch_bounds__(CT + 64 * SW) void k_ic_sweep(EllSweep<X> s, const SweepBlock* __restrict__ blk,
                                                           const X* __restrict__ rhs, X* __restrict__ out_lvl,
                                                           const CGState* __restrict__ st) {
    // blk: one sweep per workgroup over the records of independent diagonal blocks (a batch's
    // instances).  Only the counts and the level offset come from memory: the record pointers stay
    // kernel arguments, known global, so that their loads are not flat ones.
    if (blk) {
        const SweepBlock b = blk[blockIdx.x];
        s.levels = b.levels;
        s.K = b.K;
        s.lptr += b.lptr_off;
    }
    constexpr int D = sweep_depth<WE, CT>();
    extern __shared__ unsigned char smem_raw[];
    X* win = reinterpret_cast<X*>(smem_raw);
    i32* lp = reinterpret_cast<i32*>(smem_raw + ((size_t(s.K) * size_t(s.W) * sizeof(X) + 15) & ~size_t(15)));
    if (st && st->done) return;
    const int t = threadIdx.x;
    const int ct = int(blockDim.x) - 64 * SW;
    for (int i = t; i < s.K * s.W; i += blockDim.x) win[i] = X(0);
    for (int i = t; i <= s.levels; i += blockDim.x) lp[i] = s.lptr[i];
    __syncthreads();
    // the compute waves run whole trips of D levels (the phantom levels past the last one recompute
    // the last level's records into window slots no record reads): an exit inside the unrolled trip
    // would make the compiler wait for every prefetched load at the top of each trip
    const int padded = (s.levels + D - 1) / D * D;
    if (t >= ct) {
        // store waves: one barrier per level, like the compute waves, then that level's values
        constexpr int PER = (CT + 64 * SW - 1) / (64 * SW);
        const int lane = t - ct;
        for (int L = 0; L < padded; ++L) {
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
            if (L >= s.levels) continue;
            const i32 b = lp[L], cnt = lp[L + 1] - b;
            const X* src = win + (L % s.K) * s.W;
            X v[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = lane + k * 64 * SW;
                v[k] = i < cnt ? src[i] : X(0);
            }
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = lane + k * 64 * SW;
                if (i < cnt) out_lvl[b + i] = v[k];
            }
        }
        return;
    }
    // The ring.  The loads of a level are issued by every compute thread (past the level's width a
    // thread reads the level's last record again), so that each thread issues the same loads in the
    // same order.  (Masking the idle lanes' loads and 16-bit dependency slots were both measured
    // slower: the loads cost per instruction, and unpacking the slots lengthens the level.)
    X r_rhs[D], r_dinv[D], r_val[D][WE];
    i32 r_dep[D][WE];
    auto load_at = [&](i32 b, i32 cnt, X& rv, X& dv, X (&vv)[WE], i32 (&dd)[WE]) {
        const i32 rec = b + (t < cnt ? t : cnt - 1);
        rv = rhs[rec];
        dv = s.dinv[rec];
        load16(s.val + i64(WE) * rec, vv);
        load16(s.dep + i64(WE) * rec, dd);
    };
    auto load = [&](int L, X& rv, X& dv, X (&vv)[WE], i32 (&dd)[WE]) {
        load_at(lp[L], lp[L + 1] - lp[L], rv, dv, vv, dd);
    };
    // the prologue issues the ring in slot order (the scheduling barrier keeps the compiler from
    // grouping the loads by kind): the wait for slot u in the loop then counts only later slots
#pragma unroll
    for (int u = 0; u < D; ++u) {
        load(min(u, s.levels - 1), r_rhs[u], r_dinv[u], r_val[u], r_dep[u]);
        __builtin_amdgcn_sched_barrier(0);
    }
    // at least one trip (levels >= 1): a do-while has no zero-trip test, so no block for the
    // optimiser to sink the prologue's loads into, out of slot order
    int L0 = 0;
    do {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const int L = L0 + u;
            // offsets of the level this slot is refilled with, read beside the dependencies
            const int Ln = min(L + D, s.levels - 1);
            const i32 nb = lp[Ln], ncnt = lp[Ln + 1] - nb;
            // branch-free: a lane past the level's width holds the level's last record and computes
            // that record's value again into its own window slot t >= width, which no record reads
            // (the window stride is the compute width); the store waves copy only the first cnt
            X v = r_rhs[u];
#pragma unroll
            for (int j = 0; j < WE; ++j) v -= r_val[u][j] * win[r_dep[u][j]];
            v *= r_dinv[u];
            win[(L % s.K) * s.W + t] = v;
            // the slot is free once its level is computed: refill it before the barrier, so the
            // address arithmetic and the load issue overlap the wait for the other waves
            load_at(nb, ncnt, r_rhs[u], r_dinv[u], r_val[u], r_dep[u]);
            // every dependency lives in the LDS window: wait for LDS traffic only, so the global
            // loads prefetched for the next levels stay in flight across the barrier.  The
            // scheduling barriers keep the compiler from hoisting a later level's address
            // arithmetic above it, which would wait for that level's loads here.
            __builtin_amdgcn_sched_barrier(0);
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
        }
        L0 += D;
    } while (L0 < padded);
}

// Level order to row order: out[row[rec]] = v_lvl[rec]; out2[row[rec]] = scale2[rec] * v_lvl[rec].
template<class X>
__global__ void k_sweep_scatter(i64 n, const i32* __restrict__ row, const X* __restrict__ lvl,
                                const X* __restrict__ scale2, X* __restrict__ out, X* __restrict__ out2,
                                const CGState* __restrict__ st) {
    if (st && st->done) return;
    for (i64 t = i64(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += i64(gridDim.x) * blockDim.x) {
        const i32 k = row[t];
        const X v = lvl[t];
        out[k] = v;
        if (out2) out2[k] = scale2[t] * v;
    }
}

// The same sweep when the dependency window or the level offsets do not fit LDS: dependencies
// outside the window are read back from `out` (written by earlier levels of this workgroup), one
// level at a time with no prefetch.
template<class X, int WE>
__global__ __launch_bounds__(LEVEL_THREADS) void k_ic_sweep_generic(EllSweep<X> s, const SweepBlock* __restrict__ blk,
                                                                    const X* __restrict__ rhs, X* __restrict__ out,
                                                                    X* __restrict__ out2, const X* __restrict__ scale2,
                                                                    const CGState* __restrict__ st) {
    if (blk) {
        const SweepBlock b = blk[blockIdx.x];
        s.levels = b.levels;
        s.K = b.K;
        s.lptr += b.lptr_off;
    }
    extern __shared__ unsigned char smem_raw[];
    X* win = reinterpret_cast<X*>(smem_raw);
    if (st && st->done) return;
    const int t = threadIdx.x;
    for (int i = t; i < s.K * s.W; i += blockDim.x) win[i] = X(0);
    __syncthreads();
    for (int L = 0; L < s.levels; ++L) {
        const i32 b = s.lptr[L], cnt = s.lptr[L + 1] - b;
        if (t < cnt) {
            const i32 rec = b + t;
            X v = rhs[rec];
#pragma unroll
            for (int j = 0; j < WE; ++j) {
                const i32 d = s.dep[i64(WE) * rec + j];
                v -= s.val[i64(WE) * rec + j] * (s.K ? win[d] : out[-d - 1]);
            }
            v *= s.dinv[rec];
            if (s.K) win[(L % s.K) * s.W + t] = v;
            const i32 k = s.row[rec];
            out[k] = v;
            if (out2) out2[k] = scale2[rec] * v;
        }
        __syncthreads();
    }
}

__global__ void k_pack_ell(i64 n, const i64* __restrict__ src, const double* __restrict__ F, double* __restrict__ v64,
                           float* __restrict__ v32) {
    for (i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += i64(gridDim.x) * blockDim.x) {
        const double v = src[e] >= 0 ? F[src[e]] : 0.0;
        v64[e] = v;
        v32[e] = float(v);
    }
}

// Gathers the factor into the packed sweep layout (values by source position, per-record 1/U_kk, s_k).
__global__ void k_pack_entries(i64 n, const i64* __restrict__ src, const double* __restrict__ F, double* __restrict__ v64,
                               float* __restrict__ v32) {
    for (i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += i64(gridDim.x) * blockDim.x) {
        const double v = F[src[e]];
        v64[e] = v;
        v32[e] = float(v);
    }
}

__global__ void k_pack_rows(i64 n, const i32* __restrict__ row, const i64* __restrict__ diag, const double* __restrict__ F,
                            const double* __restrict__ s, double* __restrict__ dinv64, float* __restrict__ dinv32,
                            double* __restrict__ s64, float* __restrict__ s32) {
    for (i64 t = i64(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += i64(gridDim.x) * blockDim.x) {
        const i32 k = row[t];
        const double d = 1.0 / F[diag[k]];
        dinv64[t] = d;
        dinv32[t] = float(d);
        s64[t] = s[k];
        s32[t] = float(s[k]);
    }
}

__global__ void k_transpose_pos(i64 rows, const i64* __restrict__ outer, const i32* __restrict__ cols, i64* __restrict__ tpos) {
    const i64 r = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (r >= rows) return;
    const int lane = threadIdx.x & 63;
    for (i64 p = outer[r] + lane; p < outer[r + 1]; p += 64) {
        const i32 c = cols[p];
        i64 lo = outer[c], hi = outer[c + 1];
        while (lo < hi) {
            const i64 mid = (lo + hi) >> 1;
            if (cols[mid] < r) lo = mid + 1; else hi = mid;
        }
        tpos[p] = (lo < outer[c + 1] && cols[lo] == r) ? lo : p;   // symmetric pattern: always found
    }
}

template<class X>
__global__ void k_zero(i64 n, X* __restrict__ x) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) x[i] = X(0);
}

}  // namespace nlf

// ---- preconditioner object ---------------------------------------------------------------------
struct nlf_precond_s {
    nlf_ctx ctx = nullptr;
    int kind = NLF_PRECOND_IDENTITY;
    nlf::i64 rows = 0;
    bool ok = true;
    double shift = 0.0;
    double setup_seconds = 0.0;
    // Jacobi
    nlf::DevBuf<double> invd;
    nlf::DevBuf<float> invd32;
    // ILLT
    nlf::DevBuf<nlf::i64> outer, diag, tpos;
    nlf::DevBuf<nlf::i32> cols, fw_rows, fw_ptr, bw_rows, bw_ptr;
    int fw_levels = 0, bw_levels = 0;
    // independent diagonal blocks (more than one: fw_rows / fw_ptr hold each block's levels in turn,
    // and the factorisation runs a workgroup per block)
    int blocks = 1;
    nlf::DevBuf<nlf::SweepBlock> fw_blocks;
    nlf::DevBuf<double> F, scale, work;
    nlf::DevBuf<float> work32;
    struct Sweep {
        int levels = 0, K = 0, W = 0;
        nlf::DevBuf<nlf::i32> lptr, row, off, dep;
        nlf::DevBuf<double> val, dinv, scale;
        nlf::DevBuf<float> val32, dinv32, scale32;
        nlf::PackedSweep<double> f64() const {
            return {levels, K, W, lptr.p, row.p, off.p, dep.p, val.p, dinv.p, scale.p};
        }
        nlf::PackedSweep<float> f32() const {
            return {levels, K, W, lptr.p, row.p, off.p, dep.p, val32.p, dinv32.p, scale32.p};
        }
    } fw, bw;
    int solve_threads = 64;
    struct Ell {
        int levels = 0, K = 0, W = 0, WE = 0, threads = 64;
        nlf::i64 rows = 0;
        nlf::DevBuf<nlf::i32> lptr, row, dep;
        nlf::DevBuf<double> val, dinv, scale;
        nlf::DevBuf<float> val32, dinv32, scale32;
        // blocks > 1: independent diagonal blocks, one sweep each (one workgroup per block): bdesc
        // holds each block's levels, window depth and first level offset; levels and K above are
        // the largest
        int blocks = 1;
        nlf::DevBuf<nlf::SweepBlock> bdesc;
        template<class X> nlf::EllSweep<X> get() const;
    } efw, ebw;
    bool ell = false;
    nlf::DevBuf<double> rhs_lvl, y_lvl, y1, y2;
    nlf::DevBuf<float> rhs_lvl32, y_lvl32, y1_32, y2_32;
    // learned (set by nlf_learned.hip)
    void* learned = nullptr;
    void (*learned_destroy)(void*) = nullptr;
    ~nlf_precond_s() {
        if (learned && learned_destroy) learned_destroy(learned);
    }
};

template<> nlf::EllSweep<double> nlf_precond_s::Ell::get<double>() const {
    return {levels, K, W, lptr.p, row.p, dep.p, val.p, dinv.p};
}
template<> nlf::EllSweep<float> nlf_precond_s::Ell::get<float>() const {
    return {levels, K, W, lptr.p, row.p, dep.p, val32.p, dinv32.p};
}

namespace nlf {

static LevelArgs fw_args(const nlf_precond_s* P) {
    return LevelArgs{P->rows, P->fw_levels, P->outer.p, P->cols.p, P->diag.p, P->tpos.p, P->fw_rows.p, P->fw_ptr.p};
}
static LevelArgs bw_args(const nlf_precond_s* P) {
    return LevelArgs{P->rows, P->bw_levels, P->outer.p, P->cols.p, P->diag.p, P->tpos.p, P->bw_rows.p, P->bw_ptr.p};
}

// Level sets of the lower (forward) and upper (backward) dependency DAGs of a symmetric pattern.
static void level_sets(const std::vector<i64>& outer, const std::vector<i32>& cols, const std::vector<i64>& diag,
                       std::vector<i32>& fr, std::vector<i32>& fp, std::vector<i32>& br, std::vector<i32>& bp) {
    const i64 n = i64(diag.size());
    std::vector<i32> lev(n, 0);
    int nl = 0;
    for (i64 k = 0; k < n; ++k) {
        int l = 0;
        for (i64 p = outer[k]; p < diag[k]; ++p) l = std::max(l, lev[cols[p]] + 1);
        lev[k] = l;
        nl = std::max(nl, l + 1);
    }
    auto bucket = [&](int levels, std::vector<i32>& rowsv, std::vector<i32>& ptr, bool ascending) {
        ptr.assign(levels + 1, 0);
        for (i64 k = 0; k < n; ++k) ptr[lev[k] + 1]++;
        for (int l = 0; l < levels; ++l) ptr[l + 1] += ptr[l];
        rowsv.assign(n, 0);
        std::vector<i32> fill(ptr.begin(), ptr.end() - 1);
        if (ascending) for (i64 k = 0; k < n; ++k) rowsv[fill[lev[k]]++] = i32(k);
        else for (i64 k = n - 1; k >= 0; --k) rowsv[fill[lev[k]]++] = i32(k);
    };
    bucket(nl, fr, fp, true);
    std::fill(lev.begin(), lev.end(), 0);
    nl = 0;
    for (i64 k = n - 1; k >= 0; --k) {
        int l = 0;
        for (i64 p = diag[k] + 1; p < outer[k + 1]; ++p) l = std::max(l, lev[cols[p]] + 1);
        lev[k] = l;
        nl = std::max(nl, l + 1);
    }
    bucket(nl, br, bp, false);
}

unsigned grid_stride_blocks(i64 n) { return unsigned(std::min<i64>(std::max<i64>((n + 255) / 256, 1), 4096)); }

constexpr size_t IC_WINDOW_BYTES = 64 * 1024;

// Builds one packed sweep: records in level order, entries (lower part for the forward sweep,
// strict upper part for the backward one) with their dependency slots; values are gathered on the
// device from the factor by source position.
static void build_sweep(const std::vector<i64>& outer, const std::vector<i32>& cols, const std::vector<i64>& diag,
                        const std::vector<i32>& lrows, const std::vector<i32>& lptr, bool forward, const double* F,
                        const double* scale, const i64* d_diag, nlf_precond_s::Sweep& sw, hipStream_t s) {
    const i64 n = i64(diag.size());
    const int levels = int(lptr.size()) - 1;
    std::vector<i32> level_of(n), pos_of(n);
    int W = 1;
    for (int L = 0; L < levels; ++L) {
        W = std::max(W, lptr[L + 1] - lptr[L]);
        for (i32 t = lptr[L]; t < lptr[L + 1]; ++t) { level_of[lrows[t]] = L; pos_of[lrows[t]] = t - lptr[L]; }
    }
    int K = 1;
    std::vector<i32> off(n + 1, 0);
    for (i64 t = 0; t < n; ++t) {
        const i32 k = lrows[t];
        const i64 b = forward ? outer[k] : diag[k] + 1, e = forward ? diag[k] : outer[k + 1];
        off[t + 1] = off[t] + i32(e - b);
        for (i64 p = b; p < e; ++p) K = std::max(K, level_of[k] - level_of[cols[p]]);
    }
    // K + 1 slot blocks: the level being written must not share its block with a dependency K
    // levels back that other threads of the same level are still reading
    K += 1;
    const bool window = size_t(K) * size_t(W) * sizeof(double) <= IC_WINDOW_BYTES;
    std::vector<i32> dep(off[n]);
    std::vector<i64> src(off[n]);
    for (i64 t = 0; t < n; ++t) {
        const i32 k = lrows[t];
        const i64 b = forward ? outer[k] : diag[k] + 1, e = forward ? diag[k] : outer[k + 1];
        i32 q = off[t];
        for (i64 p = b; p < e; ++p, ++q) {
            const i32 i = cols[p];
            dep[q] = window ? (level_of[i] % K) * W + pos_of[i] : -(i + 1);
            src[q] = p;
        }
    }
    sw.levels = levels;
    sw.K = window ? K : 0;
    sw.W = W;
    upload(sw.lptr, lptr.data(), lptr.size(), s);
    upload(sw.row, lrows.data(), lrows.size(), s);
    upload(sw.off, off.data(), off.size(), s);
    upload(sw.dep, dep.data(), std::max<size_t>(dep.size(), 1), s);
    DevBuf<i64> d_src;
    upload(d_src, src.data(), std::max<size_t>(src.size(), 1), s);
    const size_t ne = std::max<size_t>(src.size(), 1);
    sw.val.alloc(ne); sw.val32.alloc(ne);
    sw.dinv.alloc(size_t(n)); sw.dinv32.alloc(size_t(n)); sw.scale.alloc(size_t(n)); sw.scale32.alloc(size_t(n));
    if (!src.empty())
        hipLaunchKernelGGL(k_pack_entries, dim3(grid_stride_blocks(i64(src.size()))), dim3(256), 0, s, i64(src.size()),
                           d_src.p, F, sw.val.p, sw.val32.p);
    hipLaunchKernelGGL(k_pack_rows, dim3(grid_stride_blocks(n)), dim3(256), 0, s, n, sw.row.p, d_diag, F, scale,
                       sw.dinv.p, sw.dinv32.p, sw.scale.p, sw.scale32.p);
    NLF_HIP(hipGetLastError());
    NLF_HIP(hipStreamSynchronize(s));
}

// ELL sweep for the fast kernel; false when a level is wider than one workgroup or rows have too
// many entries (the generic packed sweep is used then).
static bool build_ell(const std::vector<i64>& outer, const std::vector<i32>& cols, const std::vector<i64>& diag,
                      const std::vector<i32>& lrows, const std::vector<i32>& level_ptr, bool forward, const double* F,
                      const double* scale, const i64* d_diag, nlf_precond_s::Ell& el, hipStream_t s) {
    const i64 n = i64(diag.size());
    // Levels wider than the sweep's compute width are split into consecutive parts: the records of a
    // level are independent, so any split keeps the dependency order.  The widest level is often
    // level 0, the rows with no dependency at all (first-kind rows among them), several times wider
    // than the rest.  The width is the one of the two prefetching sweeps that needs fewer levels,
    // preferring the narrower (deeper prefetch) unless it needs over 25% more.
    std::vector<i32> lptr(1, 0);
    {
        i64 c_small = 0, c_large = 0;
        for (size_t L = 0; L + 1 < level_ptr.size(); ++L) {
            const i64 w = level_ptr[L + 1] - level_ptr[L];
            c_small += (w + SWEEP_CT_SMALL - 1) / SWEEP_CT_SMALL;
            c_large += (w + SWEEP_CT_LARGE - 1) / SWEEP_CT_LARGE;
        }
        const i32 cap = 4 * c_small <= 5 * c_large ? SWEEP_CT_SMALL : SWEEP_CT_LARGE;
        for (size_t L = 0; L + 1 < level_ptr.size(); ++L) {
            for (i32 x = level_ptr[L] + cap; x < level_ptr[L + 1]; x += cap) lptr.push_back(x);
            lptr.push_back(level_ptr[L + 1]);
        }
    }
    const int levels = int(lptr.size()) - 1;
    std::vector<i32> level_of(n), pos_of(n);
    int W = 1, WEmax = 0;
    for (int L = 0; L < levels; ++L) {
        W = std::max(W, lptr[L + 1] - lptr[L]);
        for (i32 t = lptr[L]; t < lptr[L + 1]; ++t) { level_of[lrows[t]] = L; pos_of[lrows[t]] = t - lptr[L]; }
    }
    int K = 1;
    for (i64 t = 0; t < n; ++t) {
        const i32 k = lrows[t];
        const i64 b = forward ? outer[k] : diag[k] + 1, e = forward ? diag[k] : outer[k + 1];
        WEmax = std::max(WEmax, int(e - b));
        for (i64 p = b; p < e; ++p) K = std::max(K, level_of[k] - level_of[cols[p]]);
    }
    K += 1;
    int WE = 4;
    while (WE < WEmax) WE *= 2;
    if (W > LEVEL_THREADS || WE > 32) return false;
    // one thread per record of the widest level; the window's slot blocks are one workgroup wide,
    // so that the sweep's lanes past a level's width have slots of their own to write
    const int threads = std::min(LEVEL_THREADS, std::max(64, (W + 63) / 64 * 64));
    W = threads;
    const bool window = size_t(K) * size_t(W) * sizeof(double) <= IC_WINDOW_BYTES;
    std::vector<i32> dep(size_t(WE) * n);
    std::vector<i64> src(size_t(WE) * n, -1);
    for (int L = 0; L < levels; ++L) {
        const i32 b = lptr[L], cnt = lptr[L + 1] - b;
        for (i32 t = 0; t < cnt; ++t) {
            const i32 k = lrows[b + t];
            const i64 pb = forward ? outer[k] : diag[k] + 1, pe = forward ? diag[k] : outer[k + 1];
            for (int j = 0; j < WE; ++j) {
                const size_t e = size_t(WE) * size_t(b + t) + size_t(j);
                if (pb + j < pe) {
                    const i32 i = cols[pb + j];
                    dep[e] = window ? (level_of[i] % K) * W + pos_of[i] : -(i + 1);
                    src[e] = pb + j;
                } else {
                    dep[e] = window ? 0 : -1;       // value 0 times a finite slot
                }
            }
        }
    }
    el.levels = levels;
    el.K = window ? K : 0;
    el.W = W;
    el.WE = WE;
    el.threads = threads;
    el.rows = n;
    el.blocks = 1;                                   // one sweep (a blocked layout tried before is replaced)
    el.bdesc.release();
    upload(el.lptr, lptr.data(), lptr.size(), s);
    upload(el.row, lrows.data(), lrows.size(), s);
    upload(el.dep, dep.data(), dep.size(), s);
    DevBuf<i64> d_src;
    upload(d_src, src.data(), src.size(), s);
    el.val.alloc(src.size()); el.val32.alloc(src.size());
    hipLaunchKernelGGL(k_pack_ell, dim3(grid_stride_blocks(i64(src.size()))), dim3(256), 0, s, i64(src.size()), d_src.p,
                       F, el.val.p, el.val32.p);
    el.dinv.alloc(size_t(n)); el.dinv32.alloc(size_t(n)); el.scale.alloc(size_t(n)); el.scale32.alloc(size_t(n));
    hipLaunchKernelGGL(k_pack_rows, dim3(grid_stride_blocks(n)), dim3(256), 0, s, n, el.row.p, d_diag, F, scale, el.dinv.p,
                       el.dinv32.p, el.scale.p, el.scale32.p);
    NLF_HIP(hipGetLastError());
    NLF_HIP(hipStreamSynchronize(s));
    return true;
}

// Contiguous row ranges with no entry between them: the diagonal blocks of a block-diagonal pattern
// (a batch of independent instances), merged so that each holds at least min_rows rows; {0, n}
// for one block.  The pattern is symmetric, so no row after a boundary reaches back across it
// once no row before it reaches past it.
static std::vector<i64> independent_blocks(const std::vector<i64>& outer, const std::vector<i32>& cols, i64 min_rows) {
    const i64 n = i64(outer.size()) - 1;
    std::vector<i64> b(1, 0);
    i64 reach = -1;                 // the largest column of the rows so far
    for (i64 r = 0; r + 1 < n; ++r) {
        if (outer[r + 1] > outer[r]) reach = std::max<i64>(reach, cols[outer[r + 1] - 1]);
        if (reach <= r && r + 1 - b.back() >= min_rows) b.push_back(r + 1);
    }
    if (n - b.back() < min_rows && b.size() > 1) b.pop_back();    // a short tail joins the last block
    b.push_back(n);
    return b;
}

// The global levels restricted to each block, in order: rows block after block (level order within
// a block), the blocks' level offsets one after another (global record offsets; empty levels
// dropped, a row's dependencies stay at lower levels), seg[b] the index of block b's first offset.
struct BlockLevels {
    std::vector<i32> rows, ptr, blk;
    std::vector<i64> seg;
    std::vector<i32> levels;
};
static BlockLevels block_levels(const std::vector<i32>& lrows, const std::vector<i32>& level_ptr,
                                const std::vector<i64>& bounds) {
    const i64 n = bounds.back();
    const int B = int(bounds.size()) - 1;
    BlockLevels bl;
    bl.blk.resize(size_t(n));
    for (int b = 0; b < B; ++b)
        for (i64 r = bounds[b]; r < bounds[b + 1]; ++r) bl.blk[size_t(r)] = b;
    bl.rows.resize(size_t(n));
    std::vector<i64> pos(bounds.begin(), bounds.end() - 1);
    std::vector<std::vector<i32>> bptr(static_cast<size_t>(B));
    for (int b = 0; b < B; ++b) bptr[size_t(b)].push_back(i32(bounds[b]));
    std::vector<int> touched;
    std::vector<char> is_touched(static_cast<size_t>(B), 0);
    for (size_t L = 0; L + 1 < level_ptr.size(); ++L) {
        for (i32 t = level_ptr[L]; t < level_ptr[L + 1]; ++t) {
            const int b = bl.blk[size_t(lrows[t])];
            bl.rows[size_t(pos[size_t(b)]++)] = lrows[t];
            if (!is_touched[size_t(b)]) { is_touched[size_t(b)] = 1; touched.push_back(b); }
        }
        for (int b : touched) { bptr[size_t(b)].push_back(i32(pos[size_t(b)])); is_touched[size_t(b)] = 0; }
        touched.clear();
    }
    bl.seg.resize(size_t(B) + 1);
    bl.levels.resize(size_t(B));
    for (int b = 0; b < B; ++b) {
        bl.seg[size_t(b)] = i64(bl.ptr.size());
        bl.levels[size_t(b)] = i32(bptr[size_t(b)].size()) - 1;
        bl.ptr.insert(bl.ptr.end(), bptr[size_t(b)].begin(), bptr[size_t(b)].end());
    }
    bl.seg[size_t(B)] = i64(bl.ptr.size());
    return bl;
}

// The ELL layout for a factor made of independent diagonal blocks: every block gets its own levels
// and its own sweep, one workgroup each, so that the blocks' levels run side by side instead of one
// after another.  Records are block after block; all blocks share the compute width W (their widest
// split level) and the record arrays, each has its level offsets and window depth.  Each record's
// arithmetic is the single-sweep one, so the results are the same bits.
static bool build_ell_blocks(const std::vector<i64>& outer, const std::vector<i32>& cols, const std::vector<i64>& diag,
                             const BlockLevels& bl, bool forward, const double* F, const double* scale, const i64* d_diag,
                             nlf_precond_s::Ell& el, hipStream_t s) {
    const i64 n = i64(diag.size());
    const int B = int(bl.levels.size());
    for (int b = 0; b < B; ++b)
        if (bl.levels[size_t(b)] == 0) return false;                // an empty block
    // the split width, as in build_ell, over every block's levels
    i32 cap;
    {
        i64 c_small = 0, c_large = 0;
        for (int b = 0; b < B; ++b)
            for (i64 L = bl.seg[size_t(b)]; L + 1 < bl.seg[size_t(b) + 1]; ++L) {
                const i64 w = bl.ptr[size_t(L) + 1] - bl.ptr[size_t(L)];
                c_small += (w + SWEEP_CT_SMALL - 1) / SWEEP_CT_SMALL;
                c_large += (w + SWEEP_CT_LARGE - 1) / SWEEP_CT_LARGE;
            }
        cap = 4 * c_small <= 5 * c_large ? SWEEP_CT_SMALL : SWEEP_CT_LARGE;
    }
    // split levels; records keep the block-major order of bl.rows
    std::vector<i32> lptr;
    std::vector<i64> seg(static_cast<size_t>(B) + 1, 0);
    std::vector<int> blevels(static_cast<size_t>(B), 0);
    std::vector<i32> level_of(static_cast<size_t>(n)), pos_of(static_cast<size_t>(n));
    int W = 1;
    for (int b = 0; b < B; ++b) {
        seg[size_t(b)] = i64(lptr.size());
        lptr.push_back(bl.ptr[size_t(bl.seg[size_t(b)])]);
        int lv = 0;
        for (i64 L = bl.seg[size_t(b)]; L + 1 < bl.seg[size_t(b) + 1]; ++L) {
            const i32 end = bl.ptr[size_t(L) + 1];
            for (i32 x = bl.ptr[size_t(L)]; x < end; x += cap) {
                const i32 y = std::min(x + cap, end);
                for (i32 t = x; t < y; ++t) {
                    level_of[size_t(bl.rows[size_t(t)])] = lv;
                    pos_of[size_t(bl.rows[size_t(t)])] = t - x;
                }
                W = std::max(W, y - x);
                lptr.push_back(y);
                ++lv;
            }
        }
        blevels[size_t(b)] = lv;
    }
    seg[size_t(B)] = i64(lptr.size());
    int K = 1, WEmax = 0;
    std::vector<int> bK(static_cast<size_t>(B), 1);
    for (i64 t = 0; t < n; ++t) {
        const i32 k = bl.rows[size_t(t)];
        const int b = bl.blk[size_t(k)];
        const i64 pb = forward ? outer[k] : diag[k] + 1, pe = forward ? diag[k] : outer[k + 1];
        WEmax = std::max(WEmax, int(pe - pb));
        for (i64 p = pb; p < pe; ++p) {
            if (bl.blk[size_t(cols[p])] != b) return false;         // not block diagonal
            bK[size_t(b)] = std::max(bK[size_t(b)], level_of[size_t(k)] - level_of[size_t(cols[p])] + 1);
        }
    }
    for (int b = 0; b < B; ++b) K = std::max(K, bK[size_t(b)]);
    int WE = 4;
    while (WE < WEmax) WE *= 2;
    if (W > LEVEL_THREADS || WE > 32) return false;
    const int threads = std::min(LEVEL_THREADS, std::max(64, (W + 63) / 64 * 64));
    W = threads;
    const bool window = size_t(K) * size_t(W) * sizeof(double) <= IC_WINDOW_BYTES;
    std::vector<i32> dep(size_t(WE) * size_t(n));
    std::vector<i64> src(size_t(WE) * size_t(n), -1);
    for (i64 t = 0; t < n; ++t) {
        const i32 k = bl.rows[size_t(t)];
        const int Kb = bK[size_t(bl.blk[size_t(k)])];
        const i64 pb = forward ? outer[k] : diag[k] + 1, pe = forward ? diag[k] : outer[k + 1];
        for (int j = 0; j < WE; ++j) {
            const size_t e = size_t(WE) * size_t(t) + size_t(j);
            if (pb + j < pe) {
                const i32 i = cols[pb + j];
                dep[e] = window ? (level_of[size_t(i)] % Kb) * W + pos_of[size_t(i)] : -(i + 1);
                src[e] = pb + j;
            } else {
                dep[e] = window ? 0 : -1;           // value 0 times a finite slot
            }
        }
    }
    el.levels = *std::max_element(blevels.begin(), blevels.end());
    el.K = window ? K : 0;
    el.W = W;
    el.WE = WE;
    el.threads = threads;
    el.rows = n;
    el.blocks = B;
    upload(el.lptr, lptr.data(), lptr.size(), s);
    upload(el.row, bl.rows.data(), bl.rows.size(), s);
    upload(el.dep, dep.data(), dep.size(), s);
    DevBuf<i64> d_src;
    upload(d_src, src.data(), src.size(), s);
    el.val.alloc(src.size()); el.val32.alloc(src.size());
    hipLaunchKernelGGL(k_pack_ell, dim3(grid_stride_blocks(i64(src.size()))), dim3(256), 0, s, i64(src.size()), d_src.p,
                       F, el.val.p, el.val32.p);
    el.dinv.alloc(size_t(n)); el.dinv32.alloc(size_t(n)); el.scale.alloc(size_t(n)); el.scale32.alloc(size_t(n));
    hipLaunchKernelGGL(k_pack_rows, dim3(grid_stride_blocks(n)), dim3(256), 0, s, n, el.row.p, d_diag, F, scale, el.dinv.p,
                       el.dinv32.p, el.scale.p, el.scale32.p);
    NLF_HIP(hipGetLastError());
    std::vector<SweepBlock> bd(static_cast<size_t>(B));
    for (int b = 0; b < B; ++b)
        bd[size_t(b)] = SweepBlock{blevels[size_t(b)], window ? bK[size_t(b)] : 0, i32(seg[size_t(b)]), 0};
    upload(el.bdesc, bd.data(), bd.size(), s);
    NLF_HIP(hipStreamSynchronize(s));
    return true;
}

template<class X>
const X* ell_scale(const nlf_precond_s::Ell& e);
template<> const double* ell_scale<double>(const nlf_precond_s::Ell& e) { return e.scale.p; }
template<> const float* ell_scale<float>(const nlf_precond_s::Ell& e) { return e.scale32.p; }

// LDS bytes of the dependency window, and whether the prefetching kernel takes the sweep (it needs
// the window and the level offsets in LDS and at most SWEEP_CT_LARGE compute lanes).
static size_t ell_window_bytes(const nlf_precond_s::Ell& e, size_t xsize) {
    return (size_t(std::max(e.K * e.W, 1)) * xsize + 15) & ~size_t(15);
}
static bool ell_prefetching(const nlf_precond_s::Ell& e, size_t xsize) {
    return e.K != 0 && ell_window_bytes(e, xsize) + size_t(e.levels + 1) * sizeof(i32) <= SWEEP_LDS_BYTES &&
           e.threads <= SWEEP_CT_LARGE;
}

template<class X>
void ell_sweep(const nlf_precond_s::Ell& e, const X* rhs, X* lvl, X* out, X* out2, const X* scale2, const CGState* st,
               hipStream_t s) {
    // LDS: the dependency window, then the level offsets; the prefetching kernel needs both there
    const size_t win = ell_window_bytes(e, sizeof(X));
    const size_t offs = size_t(e.levels + 1) * sizeof(i32);
    const EllSweep<X> a = e.template get<X>();
    const SweepBlock* blk = e.blocks > 1 ? e.bdesc.p : nullptr;
    const dim3 g(unsigned(e.blocks));
    if (e.levels == 0) return;
    if (!ell_prefetching(e, sizeof(X))) {
        const size_t lds = e.K ? win : 16;
        const dim3 b(unsigned(e.threads));
        switch (e.WE) {
            case 4: hipLaunchKernelGGL((k_ic_sweep_generic<X, 4>), g, b, lds, s, a, blk, rhs, out, out2, scale2, st); break;
            case 8: hipLaunchKernelGGL((k_ic_sweep_generic<X, 8>), g, b, lds, s, a, blk, rhs, out, out2, scale2, st); break;
            case 16: hipLaunchKernelGGL((k_ic_sweep_generic<X, 16>), g, b, lds, s, a, blk, rhs, out, out2, scale2, st); break;
            default: hipLaunchKernelGGL((k_ic_sweep_generic<X, 32>), g, b, lds, s, a, blk, rhs, out, out2, scale2, st); break;
        }
        return;
    }
    const size_t lds = win + offs;
    if (e.threads <= SWEEP_CT_SMALL) {
        constexpr int CT = SWEEP_CT_SMALL, SW = SWEEP_SW_SMALL;
        const dim3 b(unsigned(e.threads + 64 * SW));
        switch (e.WE) {
            case 4: hipLaunchKernelGGL((k_ic_sweep<X, 4, CT, SW>), g, b, lds, s, a, blk, rhs, lvl, st); break;
            case 8: hipLaunchKernelGGL((k_ic_sweep<X, 8, CT, SW>), g, b, lds, s, a, blk, rhs, lvl, st); break;
            case 16: hipLaunchKernelGGL((k_ic_sweep<X, 16, CT, SW>), g, b, lds, s, a, blk, rhs, lvl, st); break;
            default: hipLaunchKernelGGL((k_ic_sweep<X, 32, CT, SW>), g, b, lds, s, a, blk, rhs, lvl, st); break;
        }
    } else {
        constexpr int CT = SWEEP_CT_LARGE, SW = SWEEP_SW_LARGE;
        const dim3 b(unsigned(e.threads + 64 * SW));
        switch (e.WE) {
            case 4: hipLaunchKernelGGL((k_ic_sweep<X, 4, CT, SW>), g, b, lds, s, a, blk, rhs, lvl, st); break;
            case 8: hipLaunchKernelGGL((k_ic_sweep<X, 8, CT, SW>), g, b, lds, s, a, blk, rhs, lvl, st); break;
            case 16: hipLaunchKernelGGL((k_ic_sweep<X, 16, CT, SW>), g, b, lds, s, a, blk, rhs, lvl, st); break;
            default: hipLaunchKernelGGL((k_ic_sweep<X, 32, CT, SW>), g, b, lds, s, a, blk, rhs, lvl, st); break;
        }
    }
    const i64 n = i64(e.rows);
    hipLaunchKernelGGL(k_sweep_scatter<X>, dim3(grid_stride_blocks(n)), dim3(256), 0, s, n, e.row.p, (const X*)lvl,
                       scale2, out, out2, st);
}

// z = S U^-1 U^-T S r through the ELL sweeps: gather, forward sweep, gather, backward sweep.
template<class X>
void ell_apply(const nlf_precond_s* P, i64 n, const X* r, X* z, X* rhs, X* lvl, X* y1, X* y2, const CGState* st,
               hipStream_t s) {
    const unsigned g = grid_stride_blocks(n);
    hipLaunchKernelGGL(k_sweep_gather<X>, dim3(g), dim3(256), 0, s, n, P->efw.row.p, ell_scale<X>(P->efw), r, rhs, y1, st);
    ell_sweep<X>(P->efw, rhs, lvl, y1, (X*)nullptr, (const X*)nullptr, st, s);
    hipLaunchKernelGGL(k_sweep_gather<X>, dim3(g), dim3(256), 0, s, n, P->ebw.row.p, (const X*)nullptr, y1, rhs, y2, st);
    ell_sweep<X>(P->ebw, rhs, lvl, y2, z, ell_scale<X>(P->ebw), st, s);
}

// Applies the preconditioner z = M^-1 r on the stream (FP64 or FP32 vectors).
template<class X>
void precond_apply(const nlf_precond_s* P, i64 n, const X* r, X* z, const CGState* st, hipStream_t s);

template<>
void precond_apply<double>(const nlf_precond_s* P, i64 n, const double* r, double* z, const CGState* st, hipStream_t s) {
    if (P && P->kind == NLF_PRECOND_JACOBI)
        hipLaunchKernelGGL(k_scale<double>, dim3(grid_stride_blocks(n)), dim3(256), 0, s, n, P->invd.p, r, z);
    else if (P && P->kind == NLF_PRECOND_ILLT && P->ell)
        ell_apply<double>(P, n, r, z, P->rhs_lvl.p, P->y_lvl.p, P->y1.p, P->y2.p, st, s);
    else if (P && P->kind == NLF_PRECOND_ILLT) {
        const size_t lds = size_t(std::max(P->fw.K * P->fw.W, P->bw.K * P->bw.W)) * sizeof(double);
        hipLaunchKernelGGL((k_ic_solve<double>), dim3(1), dim3(P->solve_threads), lds, s, P->fw.f64(), P->bw.f64(), r,
                           P->work.p, z, st);
    } else
        NLF_HIP(hipMemcpyAsync(z, r, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
}

template<>
void precond_apply<float>(const nlf_precond_s* P, i64 n, const float* r, float* z, const CGState* st, hipStream_t s) {
    if (P && P->kind == NLF_PRECOND_JACOBI)
        hipLaunchKernelGGL(k_scale<float>, dim3(grid_stride_blocks(n)), dim3(256), 0, s, n, P->invd32.p, r, z);
    else if (P && P->kind == NLF_PRECOND_ILLT && P->ell)
        ell_apply<float>(P, n, r, z, P->rhs_lvl32.p, P->y_lvl32.p, P->y1_32.p, P->y2_32.p, st, s);
    else if (P && P->kind == NLF_PRECOND_ILLT) {
        const size_t lds = size_t(std::max(P->fw.K * P->fw.W, P->bw.K * P->bw.W)) * sizeof(float);
        hipLaunchKernelGGL((k_ic_solve<float>), dim3(1), dim3(P->solve_threads), lds, s, P->fw.f32(), P->bw.f32(), r,
                           P->work32.p, z, st);
    } else
        NLF_HIP(hipMemcpyAsync(z, r, size_t(n) * sizeof(float), hipMemcpyDeviceToDevice, s));
}

// Lanes per row: the smallest power of two (4..64) giving each lane at most 16 nonzeros for the
// one-nonzero-per-load kernel, 64 for the pair kernel (16-bit offsets).  Measured on the 1000 x 1000
// composite (500 nonzeros per row, FP64): one-per-load 32 lanes 1.19 ms, 64 lanes 1.29 ms, 16 lanes
// 1.31 ms; pair kernel with the row-pattern dictionary 8 lanes 0.905 ms, 16 lanes 0.93 ms, 32 lanes
// 1.00 ms, 64 lanes 1.12 ms.
int pick_lpr(double avg, bool pairs = false) {
    const double per_lane = pairs ? 64.0 : 16.0;
    int l = 4;
    while (l < 64 && double(l) * per_lane < avg) l *= 2;
    return l;
}

int pick_lpr(const nlf_matrix_s* a) {
    ensure_spmv_storage(a);
    return pick_lpr(double(a->nnz_full) / double(std::max<i64>(a->rows, 1)), a->off16.p != nullptr);
}

// SpMV grid: one row group per block iteration, capped so that DOT partials stay <= SPMV_MAX_BLOCKS
constexpr i64 SPMV_MAX_BLOCKS = 4096;
unsigned spmv_blocks(int lpr, i64 rows) {
    return unsigned(std::min<i64>(std::max<i64>((rows + (SPMV_THREADS / lpr) - 1) / (SPMV_THREADS / lpr), 1), SPMV_MAX_BLOCKS));
}

// cs: column source (spmv_cols of the matrix, or plain offsets / columns with cs.pstart = null).
// lo: 16-bit remainders of split values (V = float, NLF_MODE_SPLIT), else null.  With 16-bit
// offsets (per nonzero or the dictionary) the pair kernel k_spmv2 runs unless pairs = false; xcd
// selects its XCD-contiguous row chunks (the one-nonzero-per-load kernel's XCD numbering when
// pairs = false).
template<class V, class X, class Acc, bool DOT>
void launch_spmv(int lpr, i64 rows, const i64* outer, SpmvCols cs, const V* val, const X* x, X* y,
                 double* partial, const CGState* st, hipStream_t s, bool xcd = false, const int16_t* lo = nullptr,
                 bool pairs = true, int pf = SPMV_PAIRS_IN_FLIGHT) {
    const i32* cols = cs.cols;
    const int16_t* off16 = cs.off;
    const i64* pstart = cs.pstart;
    const unsigned blocks = spmv_blocks(lpr, rows);
    if (off16 && pairs && xcd && blocks % 8 == 0 && !lo) {
#define NLF_SPMV2_XCASE(L)                                                                                         \
    case L:                                                                                                        \
        hipLaunchKernelGGL((k_spmv2<L, V, X, Acc, DOT, false, true>), dim3(blocks), dim3(SPMV_THREADS), 0, s, rows, outer, \
                           off16, pstart, val, static_cast<const int16_t*>(nullptr), x, y, partial, st);          \
        return;
        switch (lpr) {
            NLF_SPMV2_XCASE(4)
            NLF_SPMV2_XCASE(8)
            NLF_SPMV2_XCASE(16)
            NLF_SPMV2_XCASE(32)
            default: NLF_SPMV2_XCASE(64)
        }
#undef NLF_SPMV2_XCASE
    }
    if (off16 && pairs && !xcd) {
        if (lo) {
            if constexpr (sizeof(V) == sizeof(float)) {
#define NLF_SPMV2_SCASE(L)                                                                                         \
    case L:                                                                                                        \
        hipLaunchKernelGGL((k_spmv2<L, V, X, Acc, DOT, true>), dim3(blocks), dim3(SPMV_THREADS), 0, s, rows, outer, off16, \
                           pstart, val, lo, x, y, partial, st);                                                    \
        return;
                switch (lpr) {
                    NLF_SPMV2_SCASE(4)
                    NLF_SPMV2_SCASE(8)
                    NLF_SPMV2_SCASE(16)
                    NLF_SPMV2_SCASE(32)
                    default: NLF_SPMV2_SCASE(64)
                }
#undef NLF_SPMV2_SCASE
            }
        } else {
#define NLF_SPMV2_CASE(L)                                                                                          \
    case L:                                                                                                        \
        if (pf == 2)                                                                                               \
            hipLaunchKernelGGL((k_spmv2<L, V, X, Acc, DOT, false, false, 2>), dim3(blocks), dim3(SPMV_THREADS), 0, s, rows, \
                               outer, off16, pstart, val, static_cast<const int16_t*>(nullptr), x, y, partial, st);   \
        else                                                                                                       \
            hipLaunchKernelGGL((k_spmv2<L, V, X, Acc, DOT, false, false, SPMV_PAIRS_IN_FLIGHT>), dim3(blocks),        \
                               dim3(SPMV_THREADS), 0, s, rows, outer, off16, pstart, val,                           \
                               static_cast<const int16_t*>(nullptr), x, y, partial, st);                           \
        return;
            switch (lpr) {
                NLF_SPMV2_CASE(4)
                NLF_SPMV2_CASE(8)
                NLF_SPMV2_CASE(16)
                NLF_SPMV2_CASE(32)
                default: NLF_SPMV2_CASE(64)
            }
#undef NLF_SPMV2_CASE
        }
    }
    if constexpr (sizeof(V) == sizeof(float)) {
        if (lo) {
#define NLF_SPMV_SCASE(L)                                                                                          \
    case L:                                                                                                        \
        hipLaunchKernelGGL((k_spmv<L, V, X, Acc, DOT, false, 4, false, true>), dim3(blocks), dim3(SPMV_THREADS), 0, s,  \
                           rows, outer, cols, off16, pstart, val, lo, x, y, partial, st);                          \
        return;
            switch (lpr) {
                NLF_SPMV_SCASE(4)
                NLF_SPMV_SCASE(8)
                NLF_SPMV_SCASE(16)
                NLF_SPMV_SCASE(32)
                default: NLF_SPMV_SCASE(64)
            }
#undef NLF_SPMV_SCASE
        }
    }
    if (xcd) {
#define NLF_SPMV_XCASE(L)                                                                                          \
    case L:                                                                                                        \
        hipLaunchKernelGGL((k_spmv<L, V, X, Acc, DOT, true>), dim3(blocks), dim3(SPMV_THREADS), 0, s, rows, outer, cols, \
                           off16, pstart, val, static_cast<const int16_t*>(nullptr), x, y, partial, st);          \
        return;
        switch (lpr) {
            NLF_SPMV_XCASE(4)
            NLF_SPMV_XCASE(8)
            NLF_SPMV_XCASE(16)
            NLF_SPMV_XCASE(32)
            default: NLF_SPMV_XCASE(64)
        }
#undef NLF_SPMV_XCASE
    }
#define NLF_SPMV_CASE(L)                                                                                           \
    case L:                                                                                                        \
        hipLaunchKernelGGL((k_spmv<L, V, X, Acc, DOT>), dim3(blocks), dim3(SPMV_THREADS), 0, s, rows, outer, cols, off16, \
                           pstart, val, static_cast<const int16_t*>(nullptr), x, y, partial, st);                 \
        break;
    switch (lpr) {
        NLF_SPMV_CASE(4)
        NLF_SPMV_CASE(8)
        NLF_SPMV_CASE(16)
        NLF_SPMV_CASE(32)
        default: NLF_SPMV_CASE(64)
    }
#undef NLF_SPMV_CASE
}


// Pairs of HIP events around sampled kernel launches on the solver's stream (slot 0: SpMV, 1: preconditioner).
struct KernelTimer {
    std::vector<hipEvent_t> ev[2][2];
    double total_ms[2] = {0.0, 0.0};
    i64 count[2] = {0, 0};
    ~KernelTimer() {
        for (auto& a : ev)
            for (auto& b : a)
                for (hipEvent_t e : b) (void)hipEventDestroy(e);
    }
    void start(int slot, hipStream_t s) { record(slot, 0, s); }
    void stop(int slot, hipStream_t s) { record(slot, 1, s); }
    void record(int slot, int which, hipStream_t s) {
        hipEvent_t e;
        NLF_HIP(hipEventCreate(&e));
        NLF_HIP(hipEventRecord(e, s));
        ev[slot][which].push_back(e);
    }
    void collect() {
        for (int k = 0; k < 2; ++k)
            for (size_t i = 0; i < ev[k][1].size(); ++i) {
                float ms = 0.f;
                NLF_HIP(hipEventSynchronize(ev[k][1][i]));
                NLF_HIP(hipEventElapsedTime(&ms, ev[k][0][i], ev[k][1][i]));
                total_ms[k] += ms;
                count[k] += 1;
            }
        for (auto& a : ev)
            for (auto& b : a) {
                for (hipEvent_t e : b) (void)hipEventDestroy(e);
                b.clear();
            }
    }
    double avg(int k) const { return count[k] ? total_ms[k] / double(count[k]) : 0.0; }
};

// r = b - r (r holds A x on entry); per-block partials of r.r and b.b.
__global__ __launch_bounds__(SPMV_THREADS) void k_resid_finish(i64 n, const double* __restrict__ b, double* __restrict__ r,
                                                               double* __restrict__ prr, double* __restrict__ pbb) {
    __shared__ double sh[SPMV_THREADS / 64];
    double rr = 0.0, bb = 0.0;
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) {
        const double bi = b[i], ri = bi - r[i];
        r[i] = ri;
        rr += ri * ri;
        bb += bi * bi;
    }
    rr = block_sum<double>(rr, sh);
    __syncthreads();
    bb = block_sum<double>(bb, sh);
    if (threadIdx.x == 0) { prr[blockIdx.x] = rr; pbb[blockIdx.x] = bb; }
}

// FP64 true residual r = b - A x through the SpMV kernels, with nb partials of r.r and b.b
// (nb <= SPMV_MAX_BLOCKS; k_sum_partials reduces them).
void residual64(i64 n, const i64* outer, SpmvCols cs, const double* val, int lpr, const double* x, const double* b,
                double* r, double* prr, double* pbb, unsigned nb, hipStream_t s) {
    launch_spmv<double, double, double, false>(lpr, n, outer, cs, val, x, r, nullptr, nullptr, s);
    hipLaunchKernelGGL(k_resid_finish, dim3(nb), dim3(SPMV_THREADS), 0, s, n, b, r, prr, pbb);
}


// Work memory of one CG solve: the vectors r, z, p (n entries of X each, 256-byte aligned), five
// arrays of per-block partial sums (SPMV_MAX_BLOCKS doubles each: every launch of the solve writes at
// most that many), two reduced sums and the solver state.
struct CGWork {
    static size_t align(size_t b) { return (b + 255) & ~size_t(255); }
    template<class X>
    static size_t bytes(i64 n) {
        return 3 * align(size_t(n) * sizeof(X)) + 5 * align(SPMV_MAX_BLOCKS * sizeof(double)) + align(2 * sizeof(double)) +
               align(sizeof(CGState));
    }
};

}  // namespace nlf

#include "nlf_symv.inl"

nlf_matrix_s::~nlf_matrix_s() { delete symv; }

namespace nlf {
// residual64 from the matrix: the product from the symmetric storage when it is built (symv_prepare),
// else from the full storage.
void residual64(const nlf_matrix_s* a, const double* x, const double* b, double* r, double* prr, double* pbb, unsigned nb,
                hipStream_t s) {
    if (a->symv && a->symv->version == a->val_version && a->ctx->spmv_symmetric) {
        launch_symv<false>(a, x, r, nullptr, nullptr, s, SPMV_MAX_BLOCKS);
        hipLaunchKernelGGL(k_resid_finish, dim3(nb), dim3(SPMV_THREADS), 0, s, a->rows, b, r, prr, pbb);
    } else {
        residual64(a->rows, a->outer.p, spmv_cols(a), a->val.p, pick_lpr(a), x, b, r, prr, pbb, nb, s);
    }
}


// Preconditioned CG on device vectors (conjugate_gradient::solve).  Vectors of type X, matrix values V.
// The work memory (CGWork) is allocated by the object, or given by the caller (no allocation then;
// `timed` = false also records no HIP events).
template<class V, class X>
struct CG {
    hipStream_t s;
    i64 n;
    int lpr;
    const i64* outer;
    const i32* cols;
    SpmvCols cs{};                      // column source of the SpMV (spmv_cols of the matrix)
    const V* val;
    const int16_t* lo = nullptr;        // split values: 16-bit remainders (NLF_MODE_SPLIT)
    const nlf_precond_s* P;
    DevBuf<unsigned char> own;
    X *r = nullptr, *z = nullptr, *p = nullptr;
    double *part1 = nullptr, *part2 = nullptr, *pbb = nullptr, *prz = nullptr, *pb2 = nullptr, *sums = nullptr;
    CGState* st = nullptr;
    unsigned sblocks, vblocks;
    bool timed = true;
    bool indefinite = false;            // a bordered (Neumann) system: negative curvature allowed
    const nlf_matrix_s* sym = nullptr;  // FP64 products from the symmetric storage (symv_prepare done)
    KernelTimer timer;

    CG(hipStream_t s_, i64 n_, i64 nnz, const i64* o, const i32* c, const V* v, const nlf_precond_s* P_,
       void* work = nullptr)
        : s(s_), n(n_), outer(o), cols(c), val(v), P(P_) {
        cs.cols = c;
        lpr = pick_lpr(double(nnz) / double(std::max<i64>(n, 1)));
        if (!work) {
            own.alloc(CGWork::bytes<X>(n));
            work = own.p;
        } else {
            timed = false;
        }
        unsigned char* w = static_cast<unsigned char*>(work);
        auto take = [&](size_t b) { unsigned char* q = w; w += CGWork::align(b); return q; };
        r = reinterpret_cast<X*>(take(size_t(n) * sizeof(X)));
        z = reinterpret_cast<X*>(take(size_t(n) * sizeof(X)));
        p = reinterpret_cast<X*>(take(size_t(n) * sizeof(X)));
        double** parts[5] = {&part1, &part2, &pbb, &prz, &pb2};
        for (double** q : parts) *q = reinterpret_cast<double*>(take(SPMV_MAX_BLOCKS * sizeof(double)));
        sums = reinterpret_cast<double*>(take(2 * sizeof(double)));
        st = reinterpret_cast<CGState*>(take(sizeof(CGState)));
        sblocks = spmv_blocks(lpr, n);
        vblocks = grid_stride_blocks(n);
    }

    bool jacobi() const { return P && P->kind == NLF_PRECOND_JACOBI; }

    // Column source of the matrix; the lane count follows the kernel it selects.
    void set_cols(const nlf_matrix_s* a) {
        cs = spmv_cols(a);
        lpr = pick_lpr(a);
        sblocks = spmv_blocks(lpr, n);
    }

    const X* invd() const;

    // x: initial guess in, solution out; b: right-hand side.  Returns iterations, sets residual.
    i64 solve(const X* b, X* x, double tol, i64 max_it, double* residual, int* breakdown, int check_every = 8) {
        // r = b - A x ; z = M r ; rz = r.z ; bnorm
        hipLaunchKernelGGL((k_residual<V, X>), dim3(spmv_blocks(64, n)), dim3(SPMV_THREADS), 0, s, n, outer, cols, val, x, b,
                           r, part1, part2);
        const int nb_res = int(spmv_blocks(64, n));
        NLF_HIP(hipMemcpyAsync(pbb, part2, size_t(nb_res) * sizeof(double), hipMemcpyDeviceToDevice, s));
        precond_apply<X>(P, n, r, p, nullptr, s);
        hipLaunchKernelGGL(k_dot<X>, dim3(vblocks), dim3(SPMV_THREADS), 0, s, n, r, p, part1, (const CGState*)nullptr);
        // pad partial arrays: k_init_state sums vblocks of r.z and nb_res of b.b (cleared by a kernel,
        // not a memset, so that the sequence may be captured)
        const int npad = int(std::max<unsigned>(vblocks, unsigned(nb_res)));
        hipLaunchKernelGGL(k_zero<double>, dim3(grid_stride_blocks(npad)), dim3(256), 0, s, i64(npad), prz);
        hipLaunchKernelGGL(k_zero<double>, dim3(grid_stride_blocks(npad)), dim3(256), 0, s, i64(npad), pb2);
        NLF_HIP(hipMemcpyAsync(prz, part1, size_t(vblocks) * sizeof(double), hipMemcpyDeviceToDevice, s));
        NLF_HIP(hipMemcpyAsync(pb2, pbb, size_t(nb_res) * sizeof(double), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1024), 0, s, prz, pb2, npad, tol, max_it, indefinite ? 1 : 0, st);
        CGState hs{};
        NLF_HIP(hipMemcpyAsync(&hs, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
        NLF_HIP(hipStreamSynchronize(s));
        while (!hs.done) {
            // the first iteration of each chunk starts with done == 0, so its kernels run in full:
            // HIP events around its SpMV and preconditioner give the live kernel timings
            for (int k = 0; k < check_every; ++k) iteration(x, timed && k == 0);
            NLF_HIP(hipMemcpyAsync(&hs, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
            NLF_HIP(hipStreamSynchronize(s));
        }
        timer.collect();
        if (residual) *residual = hs.res;
        if (breakdown) *breakdown = hs.breakdown;
        return hs.it;
    }

    // Mixed precision with reliable updates: the CG recurrences run in X (FP32) on the X matrix;
    // whenever the recursive residual has dropped by `delta` since the last update, the FP32
    // solution increment is folded into the FP64 solution, the residual is recomputed in FP64 with
    // the FP64 matrix and replaces the recursive one; the search direction is kept, so the Krylov
    // process is not restarted.  Stops on the FP64 true residual ||b - A x|| / ||b|| <= tol.
    i64 solve_reliable(const i64* o64, const i32* c64, const double* v64, const double* b, double* x, double tol,
                       i64 max_it, double delta, double* true_res, i64* updates, int* breakdown) {
        const unsigned rb = spmv_blocks(64, n);
        DevBuf<double> r64{size_t(n)};
        DevBuf<X> xh{size_t(n)};
        auto true_residual = [&]() {
            residual64(n, o64, cs, v64, lpr, x, b, r64.p, part1, part2, rb, s);
            hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, part1, int(rb), sums);
            hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, part2, int(rb), sums + 1);
            double h[2];
            download(h, sums, 2, s);
            return h[1] > 0.0 ? std::sqrt(h[0] / h[1]) : 0.0;
        };
        // r = M^-1-preconditioned restart data from the FP64 residual: r32 = r64, z = M r32, rz = r32.z
        auto load_residual = [&](bool first) {
            hipLaunchKernelGGL((k_convert<double, X>), dim3(vblocks), dim3(256), 0, s, n, r64.p, r);
            precond_apply<X>(P, n, r, z, nullptr, s);
            if (first) NLF_HIP(hipMemcpyAsync(p, z, size_t(n) * sizeof(X), hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(k_dot<X>, dim3(vblocks), dim3(SPMV_THREADS), 0, s, n, r, z, part1,
                               (const CGState*)nullptr);
            hipLaunchKernelGGL(k_set_rz, dim3(1), dim3(1024), 0, s, part1, int(vblocks), first ? 1 : 0, max_it,
                               indefinite ? 1 : 0, st);
        };
        i64 nupd = 0;
        double rel = true_residual();
        hipLaunchKernelGGL(k_zero<X>, dim3(vblocks), dim3(256), 0, s, n, xh.p);
        CGState hs{};
        if (rel > tol && max_it > 0) {
            load_residual(true);
            NLF_HIP(hipMemcpyAsync(&hs, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
            NLF_HIP(hipStreamSynchronize(s));
            double last = std::sqrt(std::fabs(hs.rz));
            // the recursive estimate sqrt(r'z) is preconditioned; q maps it to the relative true
            // residual as measured at the last update, so that the true residual is also checked
            // once it is predicted to meet tol (not only at the next tenfold drop)
            double q = last > 0.0 ? rel / last : 0.0;
            for (;;) {
                for (int k = 0; k < 8; ++k) iteration(xh.p, k == 0);
                NLF_HIP(hipMemcpyAsync(&hs, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
                NLF_HIP(hipStreamSynchronize(s));
                const double est = std::sqrt(std::fabs(hs.rz));
                const bool stop = hs.it >= max_it || hs.breakdown;
                if (!(est < delta * last) && !(q * est <= tol) && !stop) continue;
                hipLaunchKernelGGL((k_add_to<X, double>), dim3(vblocks), dim3(256), 0, s, n, xh.p, x);
                hipLaunchKernelGGL(k_zero<X>, dim3(vblocks), dim3(256), 0, s, n, xh.p);
                ++nupd;
                rel = true_residual();
                if (rel <= tol || stop) break;
                load_residual(false);
                NLF_HIP(hipMemcpyAsync(&hs, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
                NLF_HIP(hipStreamSynchronize(s));
                last = std::sqrt(std::fabs(hs.rz));
                q = last > 0.0 ? rel / last : 0.0;
            }
        }
        timer.collect();
        if (true_res) *true_res = rel;
        if (updates) *updates = nupd;
        if (breakdown) *breakdown = hs.breakdown;
        return hs.it;
    }

    void iteration(X* x, bool sample = false) {
        // z = A p ; alpha = rz / p.z
        if (sample) timer.start(0, s);
        bool done_sym = false;
        if constexpr (std::is_same_v<V, double> && std::is_same_v<X, double>) {
            if (sym) {
                launch_symv<true>(sym, p, z, part1, st, s, sblocks);
                done_sym = true;
            }
        }
        if (!done_sym) launch_spmv<V, X, double, true>(lpr, n, outer, cs, val, p, z, part1, st, s, false, lo);
        if (sample) timer.stop(0, s);
        hipLaunchKernelGGL(k_alpha, dim3(1), dim3(1024), 0, s, part1, int(sblocks), st);
        if (!P || P->kind == NLF_PRECOND_IDENTITY || jacobi()) {
            // x += a p ; r -= a z ; z = M r ; partial r.z  (fused)
            if (jacobi())
                hipLaunchKernelGGL((k_update_xr<X, true, true>), dim3(vblocks), dim3(SPMV_THREADS), 0, s, n, x, r, p, z,
                                   z, invd(), part2, st);
            else
                hipLaunchKernelGGL((k_update_xr<X, false, true>), dim3(vblocks), dim3(SPMV_THREADS), 0, s, n, x, r, p,
                                   z, z, (const X*)nullptr, part2, st);
        } else {
            hipLaunchKernelGGL((k_update_xr<X, false, false>), dim3(vblocks), dim3(SPMV_THREADS), 0, s, n, x, r, p, z,
                               (X*)nullptr, (const X*)nullptr, part2, st);
            if (sample) timer.start(1, s);
            precond_apply<X>(P, n, r, z, st, s);
            if (sample) timer.stop(1, s);
            hipLaunchKernelGGL(k_dot<X>, dim3(vblocks), dim3(SPMV_THREADS), 0, s, n, r, z, part2, st);
        }
        hipLaunchKernelGGL(k_beta, dim3(1), dim3(1024), 0, s, part2, int(vblocks), st);
        hipLaunchKernelGGL(k_update_p<X>, dim3(vblocks), dim3(256), 0, s, n, p, z, st);
    }
};

template<> const double* CG<double, double>::invd() const { return P->invd.p; }
template<> const float* CG<float, float>::invd() const { return P->invd32.p; }
template<> const double* CG<float, double>::invd() const { return P->invd.p; }

}  // namespace nlf

#include "nlf_learned.inl"
#include "nlf_batch.inl"

using namespace nlf;

namespace nlf {

// HBM bytes one SpMV streams: values, column source, row pointers, x once and y.
i64 spmv_stream_bytes(const nlf_matrix_s* a, size_t sv, size_t sx) {
    const i64 n = a->rows;
    if (a->symv && sv == sizeof(double) && a->symv->version == a->val_version) {
        // slice values, the strips written and read, the direct sums written and read, x once and y
        const nlf_symv_s& v = *a->symv;
        return v.slots * 64 * i64(sizeof(double)) + 2 * v.part_size * i64(sizeof(double)) + n * i64(4 * sizeof(double));
    }
    ensure_spmv_storage(a);
    i64 cols = a->nnz_full * i64(a->off16.p ? sizeof(int16_t) : sizeof(i32));
    if (a->pstart.p) cols = a->pat_size * i64(sizeof(int16_t)) + n * i64(sizeof(i64));
    return a->nnz_full * i64(sv) + cols + n * i64(sizeof(i64) + 2 * sx);
}

template<int LPR, int UNROLL, bool NT>
void launch_spmv_variant(nlf_matrix a, const double* x, double* y) {
    ensure_spmv_storage(a);
    hipLaunchKernelGGL((k_spmv<LPR, double, double, double, false, false, UNROLL, NT>), dim3(spmv_blocks(LPR, a->rows)),
                       dim3(SPMV_THREADS), 0, a->ctx->stream, a->rows, a->outer.p, a->cols.p, a->off16.p,
                       static_cast<const i64*>(nullptr), a->val.p, static_cast<const int16_t*>(nullptr), x, y, nullptr,
                       nullptr);
}

}  // namespace nlf

extern "C" {

// y = A x on the inner part (the reference's selfadjointView<Upper>() product), device vectors.
int nlf_spmv(nlf_matrix a, const double* x, double* y) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        if (symv_prepare(a, a->ctx->stream)) {
            launch_symv<false>(a, x, y, nullptr, nullptr, a->ctx->stream, SPMV_MAX_BLOCKS);
            NLF_HIP(hipGetLastError());
            return;
        }
        const int lpr = pick_lpr(a);
        launch_spmv<double, double, double, false>(lpr, a->rows, a->outer.p, spmv_cols(a), a->val.p, x, y, nullptr,
                                                   nullptr,
                                                   a->ctx->stream);
        NLF_HIP(hipGetLastError());
    });
}

int nlf_matrix_symv_info(nlf_matrix a, int64_t* slots, int64_t* strips) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        NLF_CHECK(symv_prepare(a, a->ctx->stream), NLF_ERR_STATE, "the symmetric SpMV storage does not apply");
        if (slots) *slots = a->symv->slots;
        if (strips) *strips = a->symv->part_size;
    });
}

// FP32 SpMV on the FP32 copy (nlf_matrix_make_fp32 first).
int nlf_spmv_fp32(nlf_matrix a, const float* x, float* y) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        NLF_CHECK(a->val32.p, NLF_ERR_STATE, "no FP32 copy");
        const int lpr = pick_lpr(a);
        launch_spmv<float, float, float, false>(lpr, a->rows, a->outer.p, spmv_cols(a), a->val32.p, x, y, nullptr,
                                                nullptr,
                                                a->ctx->stream);
        NLF_HIP(hipGetLastError());
    });
}


// Same SpMV with an explicit lanes-per-row choice (autotuning / benchmarking).  lpr bit 8 selects
// the XCD-contiguous workgroup numbering, bit 9 the 32-bit column indices instead of the offsets,
// bit 12 the per-nonzero 16-bit offsets instead of the row-pattern dictionary, bit 13 the split
// values (FP32 + 16-bit remainder, FP64 vectors; fp32 = 0), bit 14 the one-nonzero-per-load kernel
// instead of the pair kernel, bit 15 two nonzero pairs in flight per lane in the pair kernel instead of
// SPMV_PAIRS_IN_FLIGHT; bits 10 / 11 (FP64 only, LPR 32 or 64, per-nonzero offsets) select 8
// nonzeros in flight per lane / non-temporal loads of the matrix stream.
int nlf_spmv_lpr(nlf_matrix a, int lpr, int fp32, const void* x, void* y) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        const int variant = (lpr >> 10) & 3;
        if (variant && !fp32 && ((lpr & 0xff) == 32 || (lpr & 0xff) == 64)) {
            const double* xd = static_cast<const double*>(x);
            double* yd = static_cast<double*>(y);
            const bool l32 = (lpr & 0xff) == 32;
            if (variant == 1) l32 ? launch_spmv_variant<32, 8, false>(a, xd, yd) : launch_spmv_variant<64, 8, false>(a, xd, yd);
            else if (variant == 2) l32 ? launch_spmv_variant<32, 4, true>(a, xd, yd) : launch_spmv_variant<64, 4, true>(a, xd, yd);
            else l32 ? launch_spmv_variant<32, 8, true>(a, xd, yd) : launch_spmv_variant<64, 8, true>(a, xd, yd);
            NLF_HIP(hipGetLastError());
            return;
        }
        const bool xcd = (lpr & 0x100) != 0;
        const bool split = (lpr & 0x2000) != 0;
        const bool pairs = (lpr & 0x4000) == 0;
        const int pf = (lpr & 0x8000) ? 2 : SPMV_PAIRS_IN_FLIGHT;
        SpmvCols cs = spmv_cols(a);
        if (lpr & 0x1000) cs = SpmvCols{a->cols.p, a->off16.p, nullptr};
        if (lpr & 0x200) cs = SpmvCols{a->cols.p, nullptr, nullptr};
        lpr &= 0xff;
        if (split) {
            NLF_CHECK(a->split_ready, NLF_ERR_STATE, "no split values (nlf_matrix_make_split)");
            launch_spmv<float, double, double, false>(lpr, a->rows, a->outer.p, cs, a->val32.p,
                                                      static_cast<const double*>(x), static_cast<double*>(y), nullptr,
                                                      nullptr, a->ctx->stream, false, a->val_lo.p, pairs);
        } else if (fp32) {
            NLF_CHECK(a->val32.p, NLF_ERR_STATE, "no FP32 values");
            launch_spmv<float, float, float, false>(lpr, a->rows, a->outer.p, cs, a->val32.p, static_cast<const float*>(x),
                                                    static_cast<float*>(y), nullptr, nullptr, a->ctx->stream, xcd, nullptr,
                                                    pairs, pf);
        } else {
            launch_spmv<double, double, double, false>(lpr, a->rows, a->outer.p, cs, a->val.p, static_cast<const double*>(x),
                                                       static_cast<double*>(y), nullptr, nullptr, a->ctx->stream, xcd, nullptr,
                                                       pairs, pf);
        }
        NLF_HIP(hipGetLastError());
    });
}

// f -= K_bound x_D; f[D] = x_D  (boundary_condition_first_kind_2d), device vectors.
int nlf_apply_first_kind(nlf_matrix a, const double* xD, double* f) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        hipLaunchKernelGGL(k_first_kind, dim3(blocks_for(a->rows, 256)), dim3(256), 0, a->ctx->stream, a->rows,
                           a->b_outer.p, a->b_cols.p, a->b_val.p, a->inner.p, xD, f);
        NLF_HIP(hipGetLastError());
    });
}

// init_preconditioner<...> (iterative_solver_base.hpp:203-207): JACOBI from A's diagonal, ILLT from
// the matrix passed (the reference factors the LOCAL matrix, stationary_heat_equation_solver_2d.hpp:91-96).
int nlf_precond_create(nlf_matrix a, int kind, nlf_precond* out) {
    if (!out) return NLF_ERR_ARGUMENT;
    nlf_ctx ctx = a ? a->ctx : nullptr;
    return guarded(ctx, [&] {
        const auto t0 = std::chrono::steady_clock::now();
        auto P = std::make_unique<nlf_precond_s>();
        P->ctx = ctx;
        P->kind = kind;
        if (kind == NLF_PRECOND_IDENTITY) { P->rows = a ? a->rows : 0; *out = P.release(); return; }
        NLF_CHECK(a, NLF_ERR_ARGUMENT, "matrix required");
        hipStream_t s = ctx->stream;
        P->rows = a->rows;
        const i64 n = a->rows;
        if (kind == NLF_PRECOND_JACOBI) {
            P->invd.alloc(size_t(n));
            P->invd32.alloc(size_t(n));
            hipLaunchKernelGGL(k_inv_diag, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a->diag.p, a->val.p, P->invd.p,
                               P->invd32.p);
            NLF_HIP(hipGetLastError());
        } else if (kind == NLF_PRECOND_ILLT) {
            // pattern copy + level sets (host analysis of the pattern only)
            std::vector<i64> outer(n + 1), diag(n);
            std::vector<i32> cols(size_t(a->nnz_full));
            download(outer.data(), a->outer.p, outer.size(), s);
            download(diag.data(), a->diag.p, diag.size(), s);
            download(cols.data(), a->cols.p, cols.size(), s);
            std::vector<i32> fr, fp, br, bp;
            level_sets(outer, cols, diag, fr, fp, br, bp);
            P->fw_levels = int(fp.size()) - 1;
            P->bw_levels = int(bp.size()) - 1;
            // a block-diagonal matrix (a batch of independent instances) factors and sweeps block by
            // block, the blocks side by side: one workgroup each instead of one for all
            const std::vector<i64> bounds = independent_blocks(outer, cols, std::max<i64>(1024, n / 2048));
            P->blocks = int(bounds.size()) - 1;
            BlockLevels bfw, bbw;
            if (P->blocks > 1) {
                bfw = block_levels(fr, fp, bounds);
                bbw = block_levels(br, bp, bounds);
                std::vector<SweepBlock> bd(static_cast<size_t>(P->blocks));
                for (int b = 0; b < P->blocks; ++b)
                    bd[size_t(b)] = SweepBlock{bfw.levels[size_t(b)], 0, i32(bfw.seg[size_t(b)]), 0};
                upload(P->fw_blocks, bd.data(), bd.size(), s);
                upload(P->fw_rows, bfw.rows.data(), bfw.rows.size(), s);
                upload(P->fw_ptr, bfw.ptr.data(), bfw.ptr.size(), s);
            } else {
                upload(P->fw_rows, fr.data(), fr.size(), s);
                upload(P->fw_ptr, fp.data(), fp.size(), s);
            }
            upload(P->outer, outer.data(), outer.size(), s);
            upload(P->diag, diag.data(), diag.size(), s);
            upload(P->cols, cols.data(), cols.size(), s);
            upload(P->bw_rows, br.data(), br.size(), s);
            upload(P->bw_ptr, bp.data(), bp.size(), s);
            P->tpos.alloc(cols.size());
            hipLaunchKernelGGL(k_transpose_pos, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, P->outer.p, P->cols.p, P->tpos.p);
            P->scale.alloc(size_t(n));
            P->work.alloc(size_t(n));
            P->work32.alloc(size_t(n));
            P->F.alloc(cols.size());
            DevBuf<float> scale32{size_t(n)};
            hipLaunchKernelGGL(k_scale_from_diag, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, P->diag.p, a->val.p,
                               P->scale.p, scale32.p);
            DevBuf<int> fail{1};
            double shift = 0.0;
            P->ok = false;
            for (int attempt = 0; attempt < 40; ++attempt) {
                NLF_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), s));
                hipLaunchKernelGGL(k_ic_init, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, P->outer.p, P->cols.p, P->diag.p,
                                   a->val.p, P->scale.p, shift, P->F.p);
                hipLaunchKernelGGL(k_ic_factor, dim3(unsigned(P->blocks)), dim3(LEVEL_THREADS), 0, s, fw_args(P.get()),
                                   P->blocks > 1 ? P->fw_blocks.p : nullptr, P->F.p, fail.p);
                NLF_HIP(hipGetLastError());
                int f = 0;
                download(&f, fail.p, 1, s);
                if (!f) { P->ok = true; P->shift = shift; break; }
                shift = shift == 0.0 ? 1e-3 : 2 * shift;
            }
            P->ell = P->blocks > 1 &&
                     build_ell_blocks(outer, cols, diag, bfw, true, P->F.p, P->scale.p, P->diag.p, P->efw, s) &&
                     build_ell_blocks(outer, cols, diag, bbw, false, P->F.p, P->scale.p, P->diag.p, P->ebw, s);
            if (!P->ell)
                P->ell = build_ell(outer, cols, diag, fr, fp, true, P->F.p, P->scale.p, P->diag.p, P->efw, s) &&
                         build_ell(outer, cols, diag, br, bp, false, P->F.p, P->scale.p, P->diag.p, P->ebw, s);
            // the packed sweeps (one workgroup, any level width) where the ELL layout does not apply
            if (!P->ell) {
                build_sweep(outer, cols, diag, fr, fp, true, P->F.p, P->scale.p, P->diag.p, P->fw, s);
                build_sweep(outer, cols, diag, br, bp, false, P->F.p, P->scale.p, P->diag.p, P->bw, s);
                const int widest = std::max(P->fw.W, P->bw.W);
                P->solve_threads = std::min(LEVEL_THREADS, std::max(64, (widest + 63) / 64 * 64));
            }
            if (P->ell) {
                P->rhs_lvl.alloc(size_t(n)); P->y_lvl.alloc(size_t(n)); P->y1.alloc(size_t(n)); P->y2.alloc(size_t(n));
                P->rhs_lvl32.alloc(size_t(n)); P->y_lvl32.alloc(size_t(n)); P->y1_32.alloc(size_t(n));
                P->y2_32.alloc(size_t(n));
            }
        } else {
            throw Error(NLF_ERR_UNSUPPORTED, "use nlf_learned_precond_create for the learned preconditioner");
        }
        NLF_HIP(hipStreamSynchronize(s));
        P->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = P.release();
    });
}

int nlf_precond_destroy(nlf_precond p) {
    if (!p) return NLF_OK;
    CtxScope scope(live_ctx(p->ctx));
    delete p;
    return NLF_OK;
}

// computation_info(): ok = 1 on success; shift used; number of forward / backward levels.
int nlf_precond_info(nlf_precond p, int* ok, double* shift, int* fw_levels, int* bw_levels, double* setup_seconds) {
    if (!p) return NLF_ERR_ARGUMENT;
    if (ok) *ok = p->ok;
    if (shift) *shift = p->shift;
    if (fw_levels) *fw_levels = p->fw_levels;
    if (bw_levels) *bw_levels = p->bw_levels;
    if (setup_seconds) *setup_seconds = p->setup_seconds;
    return NLF_OK;
}

int nlf_precond_sweep_info(nlf_precond p, int32_t out[8]) {
    if (!p || !out) return NLF_ERR_ARGUMENT;
    for (int i = 0; i < 8; ++i) out[i] = 0;
    if (p->kind != NLF_PRECOND_ILLT) return NLF_ERR_ARGUMENT;
    const nlf_precond_s::Ell* e[2] = {&p->efw, &p->ebw};
    for (int k = 0; k < 2; ++k) {
        if (p->ell) {
            out[4 * k] = e[k]->levels;
            out[4 * k + 1] = e[k]->threads;
            out[4 * k + 2] = e[k]->K;
            out[4 * k + 3] = nlf::ell_prefetching(*e[k], sizeof(double)) ? 2 : 1;
        } else {
            const auto& w = k ? p->bw : p->fw;
            out[4 * k] = w.levels;
            out[4 * k + 1] = p->solve_threads;
            out[4 * k + 2] = w.K;
            out[4 * k + 3] = 0;
        }
    }
    return NLF_OK;
}

int nlf_precond_blocks(nlf_precond p, int32_t* blocks) {
    if (!p || !blocks || p->kind != NLF_PRECOND_ILLT) return NLF_ERR_ARGUMENT;
    *blocks = p->blocks;
    return NLF_OK;
}

// preconditioner_base::solve: z = M^-1 r on device vectors.
int nlf_precond_apply(nlf_precond p, const double* r, double* z) {
    if (!p) return NLF_ERR_ARGUMENT;
    return guarded(p->ctx, [&] {
        NLF_CHECK(p->rows > 0, NLF_ERR_STATE, "preconditioner has no size");
        if (p->kind == NLF_PRECOND_LEARNED) {
            LearnedGraph& g = *learned_of(p);
            learned_apply<double>(g, g.w.p, r, z, p->ctx->stream);
        } else {
            precond_apply<double>(p, p->rows, r, z, nullptr, p->ctx->stream);
        }
        NLF_HIP(hipGetLastError());
    });
}

// Host export of the ILLT factor in full storage (upper: U, lower: U^T) and the scaling.
int nlf_precond_export_factor(nlf_precond p, double* values, double* scale) {
    if (!p || p->kind != NLF_PRECOND_ILLT) return NLF_ERR_ARGUMENT;
    return guarded(p->ctx, [&] {
        if (values) download(values, p->F.p, p->F.n, p->ctx->stream);
        if (scale) download(scale, p->scale.p, p->scale.n, p->ctx->stream);
    });
}

// conjugate_gradient::solve on device vectors.  mode NLF_MODE_FP64: the reference's PCG with its
// stopping rule sqrt(r'z)/||b|| <= tol.  NLF_MODE_MIXED: FP64 iterative refinement around FP32 PCG
// (x += A32^-1 (b - A x)) until the FP64 true residual ||b - Ax||/||b|| <= tol.  NLF_MODE_SPLIT: the
// same with the values read as FP32 + 16-bit remainder.  x holds the initial guess on entry.
// ---- CG with no allocation -------------------------------------------------------------------
// The work memory of nlf_cg_solve_noalloc: one file-scope device array (every GPU the library runs
// on holds its own copy until the process ends).  Host code never passes the array's name to a
// kernel: its device address, per device, comes from hipGetSymbolAddress at context creation
// (noalloc_scratch_init) and is kept in g_scratch_ptr.
constexpr size_t NOALLOC_SCRATCH_BYTES = 3 * (size_t(NLF_NOALLOC_MAX_ROWS) * sizeof(double)) +
                                         5 * (SPMV_MAX_BLOCKS * sizeof(double)) + 3 * 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

}  // extern "C"

namespace nlf {
// Called by nlf_ctx_create with `device` current: looks up this device's copy of the scratch array once.
int noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return NLF_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return NLF_OK;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return NLF_ERR_HIP;
    g_scratch_ptr[device] = ptr;
    return NLF_OK;
}
}  // namespace nlf

extern "C" {

int nlf_cg_solve_noalloc(nlf_matrix a, nlf_precond p, const double* b, double* x, double tol, int64_t max_it,
                         nlf_solve_stats* stats) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        NLF_CHECK(b && x, NLF_ERR_ARGUMENT, "null vector");
        NLF_CHECK(!p || p->kind != NLF_PRECOND_LEARNED, NLF_ERR_ARGUMENT, "the learned preconditioner needs nlf_fcg_solve");
        NLF_CHECK(a->rows <= NLF_NOALLOC_MAX_ROWS, NLF_ERR_ARGUMENT,
                  "nlf_cg_solve_noalloc: the matrix has more rows than its fixed work memory holds (NLF_NOALLOC_MAX_ROWS = "
                  "2097152); use nlf_cg_solve");
        NLF_CHECK(CGWork::bytes<double>(a->rows) <= NOALLOC_SCRATCH_BYTES, NLF_ERR_ARGUMENT, "work memory too small");
        const int dev = a->ctx->device;
        void* work = (dev >= 0 && dev < NOALLOC_MAX_DEVICES) ? g_scratch_ptr[dev] : nullptr;
        NLF_CHECK(work, NLF_ERR_STATE, "the fixed work memory of this device was not set up (nlf_ctx_create)");
        hipStream_t s = a->ctx->stream;
        const auto t0 = std::chrono::steady_clock::now();
        const i64 n = a->rows;
        nlf_solve_stats out{};
        out.preconditioner = p ? p->kind : NLF_PRECOND_IDENTITY;
        int breakdown = 0;
        CG<double, double> cg(s, n, a->nnz_full, a->outer.p, a->cols.p, a->val.p, p, work);
        cg.set_cols(a);
        cg.indefinite = a->neumann;
        out.iterations = cg.solve(b, x, tol, max_it, &out.residual, &breakdown);
        out.spmv_bytes = spmv_stream_bytes(a, sizeof(double), sizeof(double));
        // FP64 true residual ||b - A x|| / ||b||, with z and the partial arrays of the work memory
        const unsigned rb = spmv_blocks(64, n);
        residual64(n, a->outer.p, spmv_cols(a), a->val.p, pick_lpr(a), x, b, cg.z, cg.part1, cg.part2, rb, s);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, cg.part1, int(rb), cg.sums);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, cg.part2, int(rb), cg.sums + 1);
        double h[2];
        download(h, cg.sums, 2, s);
        out.true_residual = h[1] > 0 ? std::sqrt(h[0]) / std::sqrt(h[1]) : 0.0;
        out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = out;
        NLF_CHECK(!breakdown, NLF_ERR_NUMERIC, "CG breakdown (non-positive curvature or non-finite residual)");
    });
}

int nlf_cg_solve(nlf_matrix a, nlf_precond p, const double* b, double* x, double tol, int64_t max_it, int mode,
                 nlf_solve_stats* stats) {
    if (!a) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        NLF_CHECK(b && x, NLF_ERR_ARGUMENT, "null vector");
        NLF_CHECK(!p || p->kind != NLF_PRECOND_LEARNED, NLF_ERR_ARGUMENT, "the learned preconditioner needs nlf_fcg_solve");
        hipStream_t s = a->ctx->stream;
        const auto t0 = std::chrono::steady_clock::now();
        const i64 n = a->rows;
        nlf_solve_stats out{};
        out.preconditioner = p ? p->kind : NLF_PRECOND_IDENTITY;
        int breakdown = 0;
        auto timings = [&](const KernelTimer& t, size_t sv, size_t sx) {
            out.spmv_ms = t.avg(0);
            out.precond_ms = t.avg(1);
            out.spmv_samples = t.count[0];
            out.spmv_bytes = spmv_stream_bytes(a, sv, sx);
        };
        if (mode == NLF_MODE_FP64) {
            CG<double, double> cg(s, n, a->nnz_full, a->outer.p, a->cols.p, a->val.p, p);
            cg.set_cols(a);
            cg.indefinite = a->neumann;
            if (symv_prepare(a, s)) cg.sym = a;
            out.iterations = cg.solve(b, x, tol, max_it, &out.residual, &breakdown);
            timings(cg.timer, sizeof(double), sizeof(double));
        } else {
            NLF_CHECK(mode == NLF_MODE_MIXED || mode == NLF_MODE_SPLIT, NLF_ERR_ARGUMENT, "unknown mode");
            NLF_CHECK(a->val32.p, NLF_ERR_STATE, "mixed mode needs nlf_matrix_make_fp32 first");
            NLF_CHECK(mode != NLF_MODE_SPLIT || a->split_ready, NLF_ERR_STATE, "split mode needs nlf_matrix_make_split first");
            // FP32 matrix values (the bytes that bound the SpMV) with FP64 vectors and recurrences: FP32
            // CG recurrences lose orthogonality at the conditioning of these operators (kappa ~ 1e6)
            // and need several times the iterations, while FP32 storage costs none; the FP64 true
            // residual with the FP64 matrix replaces the recursive one at every reliable update.
            CG<float, double> cg(s, n, a->nnz_full, a->outer.p, a->cols.p, a->val32.p, p);
            cg.set_cols(a);
            cg.indefinite = a->neumann;
            if (mode == NLF_MODE_SPLIT) cg.lo = a->val_lo.p;
            out.iterations = cg.solve_reliable(a->outer.p, a->cols.p, a->val.p, b, x, tol, max_it, 0.1,
                                               &out.true_residual, &out.outer_iterations, &breakdown);
            out.residual = out.true_residual;
            timings(cg.timer, mode == NLF_MODE_SPLIT ? sizeof(float) + sizeof(int16_t) : sizeof(float), sizeof(double));
        }
        // FP64 true residual ||b - A x|| / ||b||
        {
            const unsigned rb = spmv_blocks(64, n);
            DevBuf<double> r{size_t(n)}, prr{rb}, pbb{rb}, sums{2};
            residual64(a, x, b, r.p, prr.p, pbb.p, rb, s);
            hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, prr.p, int(rb), sums.p);
            hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, pbb.p, int(rb), sums.p + 1);
            double h[2];
            download(h, sums.p, 2, s);
            out.true_residual = h[1] > 0 ? std::sqrt(h[0]) / std::sqrt(h[1]) : 0.0;
        }
        out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = out;
        NLF_CHECK(!breakdown, NLF_ERR_NUMERIC, "CG breakdown (non-positive curvature or non-finite residual)");
    });
}

}  // extern "C"
