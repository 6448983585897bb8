// Influence-radius element neighbour search (find_neighbours.hpp:49-160), bit-exact.
//
// Compiled with -ffp-contract=off.  Output order is the reference's: for each element, the 3x3
// window of cells is scanned x-cell outer, y-cell inner, each cell's elements in ascending id
// (split_elements_by_subareas pushes them in id order, find_neighbours.hpp:95-117,152-155).
//
// Layout: cells of every searched group are numbered consecutively; (cell << 32 | element) keys are
// radix-sorted once, which yields every cell's element list in ascending id order.  A wave (64
// lanes) handles one element and tests 64 candidates per step; ballot + popcount give the ordered
// write position.  The Euclidean predicate is evaluated exactly as powered_distance<2>(x, y, r) <= 1
// (distance.hpp:17-23, norm.hpp:15-22): z_i = (a_i - b_i) / r_i, s = (0 + z0*z0) + z1*z1.  A
// reciprocal-multiply estimate decides candidates whose value is more than 1e-12 away from 1
// (the estimate differs from the exact value by < 2e-15 relative), the rest take the division path.
// The other distances (Lp with n != 2, the maximum norm, the rotated ellipse) go through
// general_distance, operation for operation like the reference; the 3x3 cell window is the
// reference's for every distance (each keeps |z_i| <= 1 inside its unit ball, so a cell side of
// max(r0, r1) covers it).
#include "nlf_internal.hpp"

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>

namespace nlf {

void exclusive_scan_i64(i64* data, i64 n, hipStream_t s) {
    size_t bytes = 0;
    NLF_HIP(rocprim::exclusive_scan(nullptr, bytes, data, data, i64(0), size_t(n + 1), rocprim::plus<i64>(), s));
    DevBuf<uint8_t> tmp{bytes + 16};
    NLF_HIP(rocprim::exclusive_scan(tmp.p, bytes, data, data, i64(0), size_t(n + 1), rocprim::plus<i64>(), s));
    NLF_HIP(hipStreamSynchronize(s));
}

void sort_pairs_u64(const uint64_t* keys_in, uint64_t* keys_out, const i32* vals_in, i32* vals_out, i64 n, hipStream_t s) {
    if (n <= 0) return;
    size_t bytes = 0;
    NLF_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, size_t(n), 0, 64, s));
    DevBuf<uint8_t> tmp{bytes + 16};
    NLF_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, keys_in, keys_out, vals_in, vals_out, size_t(n), 0, 64, s));
    NLF_HIP(hipStreamSynchronize(s));          // before the temporary storage is released
}

void sort_keys_u64(const unsigned long long* keys_in, unsigned long long* keys_out, i64 n, hipStream_t s) {
    if (n <= 0) return;
    size_t bytes = 0;
    NLF_HIP(rocprim::radix_sort_keys(nullptr, bytes, keys_in, keys_out, size_t(n), 0, 64, s));
    DevBuf<uint8_t> tmp{bytes + 16};
    NLF_HIP(rocprim::radix_sort_keys(tmp.p, bytes, keys_in, keys_out, size_t(n), 0, 64, s));
    NLF_HIP(hipStreamSynchronize(s));          // before the temporary storage is released
}

struct SearchGroup {
    double r[2];        // expanded radii
    double inv_r[2];
    double R;           // cell side max(r0, r1)
    double left, down;
    i64 rows, cols;     // cells along x and y
    i64 base;           // first global cell id
    i64 begin, end;     // element range
    int searched;
    DistanceParams dist;
};

__device__ __forceinline__ unsigned long long ordered_bits(double v) {
    const unsigned long long b = __double_as_longlong(v);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}

__host__ __device__ inline double from_ordered_bits(unsigned long long o) {
    const unsigned long long b = (o & 0x8000000000000000ull) ? (o & ~0x8000000000000000ull) : ~o;
    double v;
    memcpy(&v, &b, sizeof(v));
    return v;
}

// max (or min) over element-node-adjacent pairs of |c_e - c_a|^2 per group
// (diameter_between_elements_centers, find_neighbours.hpp:49-72).
// Wave-level pre-reduction for the per-group atomics below: when every active lane of a wave
// works on the same group (the usual case: groups are contiguous element ranges), the wave
// reduces first and one lane issues the atomic, instead of 64 atomics on one address.
__device__ __forceinline__ bool wave_same_group(int g, bool act, int& lead) {
    const unsigned long long m = __ballot(act);
    if (!m) { lead = -1; return true; }
    lead = __ffsll(static_cast<long long>(m)) - 1;
    const int gl = __shfl(g, lead);
    return __ballot(act && g != gl) == 0;
}

template<bool MIN>
__device__ __forceinline__ unsigned long long wave_reduce_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o);
        v = MIN ? (w < v ? w : v) : (w > v ? w : v);
    }
    return v;
}

__global__ void k_expand(i64 E, int nn, int strategy, const i32* __restrict__ el, const i32* __restrict__ group_of,
                         const int* __restrict__ searched, const i32* __restrict__ ne_off, const i32* __restrict__ ne_ids,
                         const double* __restrict__ c, unsigned long long* __restrict__ out) {
    const i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool is_min = strategy == NLF_EXPAND_MIN;
    int g = -1;
    bool act = false;
    if (e < E) {
        g = group_of[e];
        act = searched[g] != 0;
    }
    // non-negative doubles: bit order == value order; identities for the wave reduction
    unsigned long long bits = is_min ? ~0ull : 0ull;
    if (act) {
        double best = is_min ? 1.0e308 : 0.0;
        const double cx = c[2 * e], cy = c[2 * e + 1];
        for (int i = 0; i < nn; ++i) {
            const i32 v = el[e * nn + i];
            for (i32 k = ne_off[v]; k < ne_off[v + 1]; ++k) {
                const i32 a = ne_ids[k];
                const double z0 = cx - c[2 * a], z1 = cy - c[2 * a + 1];
                double s = 0.0;
                s = s + z0 * z0;
                s = s + z1 * z1;
                best = is_min ? fmin(best, s) : fmax(best, s);
            }
        }
        bits = __double_as_longlong(best);
    }
    int lead;
    if (wave_same_group(g, act, lead)) {
        if (lead < 0) return;
        bits = is_min ? wave_reduce_u64<true>(bits) : wave_reduce_u64<false>(bits);
        if (int(threadIdx.x & 63) != lead) return;
    } else if (!act) {
        return;
    }
    if (is_min) atomicMin(&out[g], bits); else atomicMax(&out[g], bits);
}

// MEAN strategy (find_neighbours.hpp:55-71): the mean of |c_e - c_a| over the distinct unordered
// pairs {e, a} met by the loop "e in the group, node of e, element a at that node" (self pairs
// included), summed in the order of that loop.  A pair is counted where the loop meets it first:
// at its lower element when both lie in the group (the outer loop reaches that one first), and at
// the first node of e that a shares.  PASS 0 counts the terms of each element, PASS 1 writes them
// (sqrt of the squared distance, evaluated as the reference does) at the element's scanned offset,
// so each group's terms are one contiguous run in loop order.
template<int PASS>
__global__ void k_mean_terms(i64 E, int nn, const i32* __restrict__ el, const i32* __restrict__ group_of,
                             const int* __restrict__ searched, const i32* __restrict__ ne_off,
                             const i32* __restrict__ ne_ids, const double* __restrict__ c, i64* __restrict__ cnt,
                             const i64* __restrict__ off, double* __restrict__ terms) {
    const i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int g = group_of[e];
    if (!searched[g]) {
        if (PASS == 0) cnt[e] = 0;
        return;
    }
    i64 n = 0;
    const i64 pos = PASS ? off[e] : 0;
    for (int i = 0; i < nn; ++i) {
        const i32 v = el[e * nn + i];
        for (i32 k = ne_off[v]; k < ne_off[v + 1]; ++k) {
            const i32 a = ne_ids[k];
            if (a != e && a < e && group_of[a] == g) continue;     // met first in a's own iteration
            bool earlier = false;                                  // a shares an earlier node of e
            for (int i2 = 0; i2 < i && !earlier; ++i2) {
                const i32 v2 = el[e * nn + i2];
                for (int j = 0; j < nn; ++j) earlier |= el[i64(a) * nn + j] == v2;
            }
            if (earlier) continue;
            if (PASS == 1) {
                const double z0 = c[2 * e] - c[2 * i64(a)], z1 = c[2 * e + 1] - c[2 * i64(a) + 1];
                double d = 0.0;
                d = d + z0 * z0;
                d = d + z1 * z1;
                terms[pos + n] = sqrt(d);
            }
            ++n;
        }
    }
    if (PASS == 0) cnt[e] = n;
}

// Serial sum of each group's run of MEAN terms in loop order (one block per group): the block
// stages tiles of the run in LDS with coalesced loads and one thread adds them one after the
// other, so the rounding is that of the reference's running sum.  out[g] = sum / count.
constexpr int MEAN_TILE = 2048;
__global__ __launch_bounds__(256) void k_mean_sum(const i64* __restrict__ gbeg, const int* __restrict__ searched,
                                                  const i64* __restrict__ off, const double* __restrict__ terms,
                                                  double* __restrict__ out) {
    __shared__ double tile[MEAN_TILE];
    const int g = blockIdx.x;
    if (!searched[g]) {
        if (threadIdx.x == 0) out[g] = 0.0;
        return;
    }
    const i64 b = off[gbeg[g]], e = off[gbeg[g + 1]];
    double sum = 0.0;
    for (i64 t0 = b; t0 < e; t0 += MEAN_TILE) {
        const int len = int(e - t0 < MEAN_TILE ? e - t0 : MEAN_TILE);
        for (int t = threadIdx.x; t < len; t += blockDim.x) tile[t] = terms[t0 + t];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int t = 0; t < len; ++t) sum = sum + tile[t];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[g] = e > b ? sum / double(e - b) : 0.0;
}

// group_corners (find_neighbours.hpp:40-47): min/max over the nodes of the group's elements.
__global__ void k_bbox(i64 E, int nn, const i32* __restrict__ el, const i32* __restrict__ group_of,
                       const int* __restrict__ searched, const double* __restrict__ xy,
                       unsigned long long* __restrict__ box) {
    const i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    int g = -1;
    bool act = false;
    if (e < E) {
        g = group_of[e];
        act = searched[g] != 0;
    }
    unsigned long long lo_x = ~0ull, lo_y = ~0ull, hi_x = 0ull, hi_y = 0ull;   // ordered bits
    if (act) {
        double l = 1e308, d = 1e308, r = -1e308, u = -1e308;
        for (int i = 0; i < nn; ++i) {
            const i32 v = el[e * nn + i];
            l = fmin(l, xy[2 * v]); d = fmin(d, xy[2 * v + 1]);
            r = fmax(r, xy[2 * v]); u = fmax(u, xy[2 * v + 1]);
        }
        lo_x = ordered_bits(l); lo_y = ordered_bits(d); hi_x = ordered_bits(r); hi_y = ordered_bits(u);
    }
    int lead;
    if (wave_same_group(g, act, lead)) {
        if (lead < 0) return;
        lo_x = wave_reduce_u64<true>(lo_x);
        lo_y = wave_reduce_u64<true>(lo_y);
        hi_x = wave_reduce_u64<false>(hi_x);
        hi_y = wave_reduce_u64<false>(hi_y);
        if (int(threadIdx.x & 63) != lead) return;
    } else if (!act) {
        return;
    }
    atomicMin(&box[4 * g + 0], lo_x);
    atomicMin(&box[4 * g + 1], lo_y);
    atomicMax(&box[4 * g + 2], hi_x);
    atomicMax(&box[4 * g + 3], hi_y);
}

__device__ __forceinline__ void cell_of(const SearchGroup& sg, double cx, double cy, i64& a, i64& b) {
    // subarea_id (find_neighbours.hpp:86-93): size_t((c - corner) / R); clamped to the grid so a
    // centre rounded onto the far corner cannot index past the last cell
    const double u = (cx - sg.left) / sg.R, v = (cy - sg.down) / sg.R;
    a = u > 0.0 ? i64((unsigned long long)u) : 0;
    b = v > 0.0 ? i64((unsigned long long)v) : 0;
    a = a < sg.rows ? a : sg.rows - 1;
    b = b < sg.cols ? b : sg.cols - 1;
}

__global__ void k_cell_keys(i64 E, const i32* __restrict__ group_of, const SearchGroup* __restrict__ groups,
                            const double* __restrict__ c, unsigned long long* __restrict__ keys,
                            i64* __restrict__ cell_count, i64* __restrict__ n_keys_unused) {
    const i64 e = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const SearchGroup sg = groups[group_of[e]];
    unsigned long long key = ~0ull;                               // not searched: sorts last
    if (sg.searched) {
        i64 a, b;
        cell_of(sg, c[2 * e], c[2 * e + 1], a, b);
        const i64 cell = sg.base + a * sg.cols + b;
        key = (static_cast<unsigned long long>(cell) << 32) | static_cast<unsigned long long>(e);
        atomicAdd(reinterpret_cast<unsigned long long*>(&cell_count[cell]), 1ull);
    }
    keys[e] = key;
}

__device__ __forceinline__ bool inside(const SearchGroup& sg, double ax, double ay, double bx, double by) {
    if (sg.dist.kind != DIST_L2) return general_distance(sg.dist, sg.r[0], sg.r[1], ax, ay, bx, by) <= 1.0;
    const double dx = ax - bx, dy = ay - by;
    const double ex = dx * sg.inv_r[0], ey = dy * sg.inv_r[1];
    double est = 0.0;
    est = est + ex * ex;
    est = est + ey * ey;
    if (est < 1.0 - 1e-12) return true;
    if (est > 1.0 + 1e-12) return false;
    const double z0 = dx / sg.r[0], z1 = dy / sg.r[1];
    double s = 0.0;
    s = s + z0 * z0;
    s = s + z1 * z1;
    return s <= 1.0;
}

// One wave per element.  PASS 0 counts, PASS 1 writes ids at offsets.
template<int PASS>
__global__ __launch_bounds__(256) void k_neighbours(i64 E, const i32* __restrict__ group_of,
                                                    const SearchGroup* __restrict__ groups, const double* __restrict__ c,
                                                    const i64* __restrict__ cell_off, const i32* __restrict__ cell_ids,
                                                    i64* __restrict__ count_or_off, i32* __restrict__ ids) {
    const int lane = threadIdx.x & 63;
    const i64 e = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (e >= E) return;
    const SearchGroup sg = groups[group_of[e]];
    if (!sg.searched) {
        if (PASS == 0 && lane == 0) count_or_off[e] = 0;
        return;
    }
    const double ax = c[2 * e], ay = c[2 * e + 1];
    i64 row, col;
    cell_of(sg, ax, ay, row, col);
    const i64 r0 = row ? row - 1 : row, r1 = row != sg.rows - 1 ? row + 2 : row + 1;   // subareas_ids :109-117
    const i64 c0 = col ? col - 1 : col, c1 = col != sg.cols - 1 ? col + 2 : col + 1;
    i64 pos = PASS == 0 ? 0 : count_or_off[e];
    for (i64 i = r0; i < r1; ++i)
        for (i64 j = c0; j < c1; ++j) {
            const i64 cell = sg.base + i * sg.cols + j;
            const i64 b = cell_off[cell], end = cell_off[cell + 1];
            for (i64 k = b; k < end; k += 64) {
                bool take = false;
                i32 cand = 0;
                if (k + lane < end) {
                    cand = cell_ids[k + lane];
                    take = inside(sg, ax, ay, c[2 * cand], c[2 * cand + 1]);
                }
                const unsigned long long mask = __ballot(take);
                if (PASS == 1 && take) ids[pos + __popcll(mask & ((1ull << lane) - 1ull))] = cand;
                pos += __popcll(mask);
            }
        }
    if (PASS == 0 && lane == 0) count_or_off[e] = pos;
}

__global__ void k_extract_ids(i64 n, const unsigned long long* __restrict__ keys, i32* __restrict__ ids) {
    const i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) ids[i] = i32(keys[i] & 0xffffffffull);
}

}  // namespace nlf

using namespace nlf;

extern "C" {

// find_neighbours (find_neighbours.hpp:119-160) + mesh_2d::neighbours (mesh_2d.hpp:205-212).
// search_radii[2G]: per-group search radius (search_radius, else nonlocal_radius; read_model.hpp:226-244);
// a group with a radius <= 0 is not searched and its elements get empty lists.  distances[G] (null:
// Euclidean) the distance function of each group's influence.  expanded_out[2G] (optional) receives
// the expanded radii.  Strategies NO, MAX, MIN and MEAN.
int nlf_find_neighbours(nlf_mesh m, const double* search_radii, int strategy, double* expanded_out) {
    return nlf_find_neighbours_ex(m, search_radii, nullptr, strategy, expanded_out);
}

int nlf_find_neighbours_ex(nlf_mesh m, const double* search_radii, const nlf_distance* distances, int strategy,
                           double* expanded_out) {
    if (!m) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        NLF_CHECK(search_radii, NLF_ERR_ARGUMENT, "null radii");
        std::vector<DistanceParams> dist(size_t(m->G));
        if (distances)
            for (int g = 0; g < m->G; ++g)
                dist[size_t(g)] = make_distance(distances[g].kind, distances[g].n, distances[g].n_real);
        NLF_CHECK(strategy == NLF_EXPAND_NO || strategy == NLF_EXPAND_MAX || strategy == NLF_EXPAND_MIN ||
                      strategy == NLF_EXPAND_MEAN, NLF_ERR_ARGUMENT, "unknown diameter expanding strategy");
        const auto t0 = std::chrono::steady_clock::now();
        hipStream_t s = m->ctx->stream;
        const int G = m->G;
        const i64 E = m->E;
        std::vector<int> searched(G);
        for (int g = 0; g < G; ++g)
            searched[g] = search_radii[2 * g] > 0.0 && search_radii[2 * g + 1] > 0.0 && m->gbeg_h[g + 1] > m->gbeg_h[g];
        DevBuf<int> d_searched;
        upload(d_searched, searched.data(), size_t(G), s);
        // expansion
        std::vector<double> add(G, 0.0);
        if (strategy == NLF_EXPAND_MEAN) {
            DevBuf<i64> d_off{size_t(E + 1)};
            hipLaunchKernelGGL(k_mean_terms<0>, dim3(blocks_for(E, 256)), dim3(256), 0, s, E, m->nn, m->el.p,
                               m->group_of.p, d_searched.p, m->ne_off.p, m->ne_ids.p, m->centres.p, d_off.p,
                               (const i64*)nullptr, (double*)nullptr);
            NLF_HIP(hipGetLastError());
            exclusive_scan_i64(d_off.p, E, s);
            i64 terms_total = 0;
            download(&terms_total, d_off.p + E, 1, s);
            DevBuf<double> d_terms{size_t(std::max<i64>(terms_total, 1))};
            hipLaunchKernelGGL(k_mean_terms<1>, dim3(blocks_for(E, 256)), dim3(256), 0, s, E, m->nn, m->el.p,
                               m->group_of.p, d_searched.p, m->ne_off.p, m->ne_ids.p, m->centres.p, (i64*)nullptr,
                               d_off.p, d_terms.p);
            NLF_HIP(hipGetLastError());
            std::vector<i64> gbeg(m->gbeg_h.begin(), m->gbeg_h.end());
            DevBuf<i64> d_gbeg;
            upload(d_gbeg, gbeg.data(), gbeg.size(), s);
            DevBuf<double> d_mean{size_t(G)};
            hipLaunchKernelGGL(k_mean_sum, dim3(G), dim3(256), 0, s, d_gbeg.p, d_searched.p, d_off.p, d_terms.p,
                               d_mean.p);
            NLF_HIP(hipGetLastError());
            download(add.data(), d_mean.p, size_t(G), s);
        } else if (strategy != NLF_EXPAND_NO) {
            std::vector<unsigned long long> init(G, strategy == NLF_EXPAND_MIN ? 0x7fefffffffffffffull : 0ull);
            DevBuf<unsigned long long> d_exp;
            upload(d_exp, init.data(), size_t(G), s);
            hipLaunchKernelGGL(k_expand, dim3(blocks_for(E, 256)), dim3(256), 0, s, E, m->nn, strategy, m->el.p,
                               m->group_of.p, d_searched.p, m->ne_off.p, m->ne_ids.p, m->centres.p, d_exp.p);
            NLF_HIP(hipGetLastError());
            std::vector<unsigned long long> got(G);
            download(got.data(), d_exp.p, size_t(G), s);
            for (int g = 0; g < G; ++g) {
                double v;
                std::memcpy(&v, &got[g], sizeof(v));
                add[g] = searched[g] ? std::sqrt(v) : 0.0;
            }
        }
        // corners
        std::vector<unsigned long long> box_init(4 * G);
        for (int g = 0; g < G; ++g) {
            box_init[4 * g + 0] = box_init[4 * g + 1] = ~0ull;
            box_init[4 * g + 2] = box_init[4 * g + 3] = 0ull;
        }
        DevBuf<unsigned long long> d_box;
        upload(d_box, box_init.data(), box_init.size(), s);
        hipLaunchKernelGGL(k_bbox, dim3(blocks_for(E, 256)), dim3(256), 0, s, E, m->nn, m->el.p, m->group_of.p,
                           d_searched.p, m->xy.p, d_box.p);
        NLF_HIP(hipGetLastError());
        std::vector<unsigned long long> box(4 * G);
        download(box.data(), d_box.p, box.size(), s);
        std::vector<SearchGroup> sg(G);
        m->expanded.assign(2 * G, 0.0);
        bool nb_symmetric = true;
        i64 cells = 0;
        for (int g = 0; g < G; ++g) {
            SearchGroup& q = sg[g];
            std::memset(&q, 0, sizeof(q));
            q.begin = m->gbeg_h[g]; q.end = m->gbeg_h[g + 1];
            q.r[0] = search_radii[2 * g] + add[g];
            q.r[1] = search_radii[2 * g + 1] + add[g];
            q.searched = searched[g] && q.r[0] > 0.0 && q.r[1] > 0.0;
            q.dist = dist[size_t(g)];
            if (!q.searched) continue;
            nb_symmetric = nb_symmetric && q.dist.symmetric();
            m->expanded[2 * g] = q.r[0]; m->expanded[2 * g + 1] = q.r[1];
            q.inv_r[0] = 1.0 / q.r[0]; q.inv_r[1] = 1.0 / q.r[1];
            q.R = std::max(q.r[0], q.r[1]);
            q.left = from_ordered_bits(box[4 * g + 0]); q.down = from_ordered_bits(box[4 * g + 1]);
            const double right = from_ordered_bits(box[4 * g + 2]), up = from_ordered_bits(box[4 * g + 3]);
            q.rows = i64(size_t((right - q.left) / q.R)) + 1;          // split_elements_by_subareas :101
            q.cols = i64(size_t((up - q.down) / q.R)) + 1;
            q.base = cells;
            cells += q.rows * q.cols;
        }
        NLF_CHECK(cells < (i64(1) << 31), NLF_ERR_UNSUPPORTED, "too many search cells");
        m->searched = searched;
        DevBuf<SearchGroup> d_sg;
        upload(d_sg, sg.data(), size_t(G), s);
        // bin elements into cells: count + radix sort of (cell, element) keys
        DevBuf<i64> cell_off{size_t(cells + 1)};
        NLF_HIP(hipMemsetAsync(cell_off.p, 0, cell_off.bytes(), s));
        DevBuf<unsigned long long> keys{size_t(E)}, sorted{size_t(E)};
        hipLaunchKernelGGL(k_cell_keys, dim3(blocks_for(E, 256)), dim3(256), 0, s, E, m->group_of.p, d_sg.p,
                           m->centres.p, keys.p, cell_off.p, (i64*)nullptr);
        NLF_HIP(hipGetLastError());
        exclusive_scan_i64(cell_off.p, cells, s);
        {
            int end_bit = 32;
            while ((i64(1) << (end_bit - 32)) <= cells && end_bit < 64) ++end_bit;
            end_bit = std::min(64, end_bit + 1);
            size_t bytes = 0;
            NLF_HIP(rocprim::radix_sort_keys(nullptr, bytes, keys.p, sorted.p, size_t(E), 0, end_bit, s));
            DevBuf<uint8_t> tmp{bytes + 16};
            NLF_HIP(rocprim::radix_sort_keys(tmp.p, bytes, keys.p, sorted.p, size_t(E), 0, end_bit, s));
        }
        DevBuf<i32> cell_ids{size_t(E)};
        hipLaunchKernelGGL(k_extract_ids, dim3(blocks_for(E, 256)), dim3(256), 0, s, E, sorted.p, cell_ids.p);
        // count, scan, fill
        m->nb_off.alloc(size_t(E + 1));
        NLF_HIP(hipMemsetAsync(m->nb_off.p, 0, m->nb_off.bytes(), s));
        const unsigned wblocks = blocks_for(E, 4);
        hipLaunchKernelGGL(k_neighbours<0>, dim3(wblocks), dim3(256), 0, s, E, m->group_of.p, d_sg.p, m->centres.p,
                           cell_off.p, cell_ids.p, m->nb_off.p, (i32*)nullptr);
        NLF_HIP(hipGetLastError());
        exclusive_scan_i64(m->nb_off.p, E, s);
        i64 total = 0;
        download(&total, m->nb_off.p + E, 1, s);
        m->nb_ids.alloc(size_t(std::max<i64>(total, 1)));
        hipLaunchKernelGGL(k_neighbours<1>, dim3(wblocks), dim3(256), 0, s, E, m->group_of.p, d_sg.p, m->centres.p,
                           cell_off.p, cell_ids.p, m->nb_off.p, m->nb_ids.p);
        NLF_HIP(hipGetLastError());
        NLF_HIP(hipStreamSynchronize(s));
        m->has_neighbours = true;
        m->nb_symmetric = nb_symmetric;
        if (expanded_out) std::copy(m->expanded.begin(), m->expanded.end(), expanded_out);
        m->last_neighbour_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
}

int nlf_neighbours_count(nlf_mesh m, int64_t* total) {
    if (!m || !total) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        NLF_CHECK(m->has_neighbours, NLF_ERR_STATE, "neighbours not computed");
        download(total, m->nb_off.p + m->E, 1, m->ctx->stream);
    });
}

// Host export of the lists as CSR: offsets[E+1], ids[total] (for bit-exact parity checks).
int nlf_neighbours_export(nlf_mesh m, int64_t* offsets, int32_t* ids) {
    if (!m) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        NLF_CHECK(m->has_neighbours, NLF_ERR_STATE, "neighbours not computed");
        hipStream_t s = m->ctx->stream;
        i64 total = 0;
        download(&total, m->nb_off.p + m->E, 1, s);
        if (offsets) download(offsets, m->nb_off.p, size_t(m->E + 1), s);
        if (ids && total) download(ids, m->nb_ids.p, size_t(total), s);
    });
}

// Device pointers of the lists (offsets i64[E+1], ids i32) for callers that keep them on the GPU.
int nlf_neighbours_device(nlf_mesh m, const int64_t** offsets, const int32_t** ids) {
    if (!m || !m->has_neighbours) return NLF_ERR_STATE;
    if (offsets) *offsets = m->nb_off.p;
    if (ids) *ids = m->nb_ids.p;
    return NLF_OK;
}

}  // extern "C"
