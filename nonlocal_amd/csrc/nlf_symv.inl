// Symmetric SpMV from the upper triangle (independent_symmetric_matrix_vector_product,
// slae/independent_symmetric_matrix_vector_product.hpp:55-79, reads the upper CSR of the reference and
// adds each off-diagonal entry twice: y_i += a_ij x_j and y_j += a_ij x_i).  Included by nlf_solver.hip.
//
// The full (both-triangle) storage reads every off-diagonal value twice from HBM; this form reads it
// once.  Layout, per slice of 64 consecutive rows (one wave):
//   * U_s, the sorted union of the upper column offsets (col - row) of the slice's rows;
//   * values column-major over U_s: sval[(kptr[s] + k) * 64 + lane] is row 64 s + lane's entry at
//     offset U_s[k], 0 where the row has none (the padding of rows with other patterns);
//   * meta[kptr[s] + k] = {U_s[k], window index of lane 0's transposed target, or -1 for the diagonal}.
// The wave accumulates a_ij x_j of its rows in registers and the transposed products a_ij x_i in a
// wave-private LDS window.  The window holds one strip per band (a run of consecutive offsets
// u_b .. u_b + w_b - 1 of U_s): the targets 64 s + u_b + [0, 64 + w_b - 1).  One wave owns the window
// and its LDS accesses execute in program order, so no atomics and no barriers are needed and the
// sums run in a fixed order.  The window is then written out (part[pbase[s] + i]).  A second kernel
// adds, per target row, the strips that cover it in (slice, band) order to the row's direct sum.
// The result is bitwise reproducible; it differs from the full-storage kernels' in summation order.
//
// Traffic per product (1000 x 1000 composite): the upper values once (2.1 GB with the padding) and
// the strips written and read (about 0.3 GB), against 4.0 GB of values in the full storage.

namespace nlf {

struct SymEntry {                        // one strip covering a target slice
    i64 base;                            // part index of lane 0 (lanes lo..hi-1 are in the strip)
    int lo, hi;
};

struct SymBand {                         // a run of consecutive offsets of a slice's union
    int u, wb, w;                        // first offset, window index of the strip, run length
    int pad_;
};

#ifndef NLF_SYMV_WAVES
#define NLF_SYMV_WAVES 4
#endif
#ifndef NLF_SYMV_UNROLL
#define NLF_SYMV_UNROLL 8
#endif
constexpr int SYMV_WAVES = NLF_SYMV_WAVES;    // slices per block of the product kernel
constexpr int SYMV_UNROLL = NLF_SYMV_UNROLL;  // entries each lane has in flight

template<bool STOP>
__global__ __launch_bounds__(64 * SYMV_WAVES) void k_symv_slices(i64 rows, i64 ns, int wmax, const i64* __restrict__ kptr,
                                                                 const int2* __restrict__ meta, const double* __restrict__ sval,
                                                                 const i64* __restrict__ pbase, const double* __restrict__ x,
                                                                 double* __restrict__ direct, double* __restrict__ part,
                                                                 const CGState* __restrict__ st) {
    extern __shared__ double win[];
    if (STOP && st->done) return;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const i64 s = i64(blockIdx.x) * SYMV_WAVES + w;
    if (s >= ns) return;
    double* W = win + size_t(w) * size_t(wmax);
    const i64 k0 = kptr[s], k1 = kptr[s + 1];
    const i64 pb = pbase[s];
    const int wl = int(pbase[s + 1] - pb);
    for (int i = lane; i < wl; i += 64) W[i] = 0.0;
    const i64 r = s * 64 + lane;
    const i64 rc = r < rows ? r : rows - 1;
    const double xr = r < rows ? x[r] : 0.0;
    double acc = 0.0;
    // chunks of SYMV_UNROLL entries, the next chunk's values loaded while the current one is summed: the
    // current chunk's x gathers are issued first, so waiting for them leaves the next values in flight
    int2 mc[SYMV_UNROLL], mn[SYMV_UNROLL];
    double ac[SYMV_UNROLL], an[SYMV_UNROLL];
    auto load = [&](i64 kk, int2 (&m)[SYMV_UNROLL], double (&a)[SYMV_UNROLL]) {
#pragma unroll
        for (int t = 0; t < SYMV_UNROLL; ++t) {
            const bool in = kk + t < k1;
            const i64 q = in ? kk + t : k1 - 1;
            m[t] = meta[q];
            a[t] = __builtin_nontemporal_load(sval + q * 64 + lane);
            if (!in) { m[t].y = -1; a[t] = 0.0; }
        }
    };
    load(k0, mc, ac);
    for (i64 k = k0; k < k1; k += SYMV_UNROLL) {
        double xc[SYMV_UNROLL];
#pragma unroll
        for (int t = 0; t < SYMV_UNROLL; ++t) {
            const i64 c = rc + mc[t].x;                        // padding (a = 0) may point past the end
            xc[t] = x[c < rows ? c : rows - 1];
        }
        const bool more = k + SYMV_UNROLL < k1;
        if (more) load(k + SYMV_UNROLL, mn, an);
#pragma unroll
        for (int t = 0; t < SYMV_UNROLL; ++t) {
            acc += ac[t] * xc[t];
            if (mc[t].y >= 0) W[mc[t].y + lane] += ac[t] * xr;
        }
        if (more) {
#pragma unroll
            for (int t = 0; t < SYMV_UNROLL; ++t) { mc[t] = mn[t]; ac[t] = an[t]; }
        }
    }
    if (r < rows) direct[r] = acc;
    for (int i = lane; i < wl; i += 64) part[pb + i] = W[i];
}

// y_t = direct_t + the strips covering t in (slice, band) order; DOT: partial[block] = sum x_t y_t.
template<bool DOT, bool STOP>
__global__ __launch_bounds__(256) void k_symv_reduce(i64 rows, i64 ns, const i64* __restrict__ rptr,
                                                     const SymEntry* __restrict__ ent, const double* __restrict__ direct,
                                                     const double* __restrict__ part, const double* __restrict__ x,
                                                     double* __restrict__ y, double* __restrict__ partial,
                                                     const CGState* __restrict__ st) {
    __shared__ double sh[256 / 64];
    if (STOP && st->done) return;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    double d = 0.0;
    for (i64 tau = i64(blockIdx.x) * 4 + w; tau < ns; tau += i64(gridDim.x) * 4) {
        const i64 t = tau * 64 + lane;
        double acc = t < rows ? direct[t] : 0.0;
        const i64 e1 = rptr[tau + 1];
        for (i64 e = rptr[tau]; e < e1; ++e) {
            const SymEntry q = ent[e];
            if (lane >= q.lo && lane < q.hi) acc += part[q.base + lane];
        }
        if (t < rows) {
            y[t] = acc;
            if (DOT) d += x[t] * acc;
        }
    }
    if (DOT) {
        d = block_sum<double>(d, sh);
        if (threadIdx.x == 0) partial[blockIdx.x] = d;
    }
}

// ---- build ----------------------------------------------------------------------------------------
// widest upper offset (col - row) of any row
__global__ void k_upper_width(i64 rows, const i64* __restrict__ outer, const i32* __restrict__ cols,
                              unsigned* __restrict__ mx) {
    const i64 r = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows || outer[r] == outer[r + 1]) return;
    atomicMax(mx, unsigned(std::min<i64>(i64(cols[outer[r + 1] - 1]) - r, i64(0x7fffffff))));
}

constexpr int SYMB_THREADS = 256;        // a block (four waves) per slice in the build

// The union of the upper offsets of slice s's rows as a bitmap in LDS (offsets < 2^15: `words`
// 32-bit words).  Wave w takes rows w, w + 4, ...; each row's upper entries are read 256 at a time,
// coalesced, before they are marked.
__device__ __forceinline__ void sym_bitmap(i64 rows, i64 s, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                           const i64* __restrict__ diag, int words, unsigned* bm) {
    const int lane = threadIdx.x & 63, w = int(threadIdx.x >> 6);
    for (int i = threadIdx.x; i < words; i += SYMB_THREADS) bm[i] = 0u;
    __syncthreads();
    for (int j = w; j < 64; j += SYMB_THREADS / 64) {
        const i64 r = s * 64 + j;
        if (r >= rows) break;
        const i64 b = diag[r], e = outer[r + 1];
        for (i64 p0 = b; p0 < e; p0 += 256) {
            i32 c[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const i64 p = p0 + t * 64 + lane;
                c[t] = p < e ? cols[p] : -1;
            }
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (c[t] >= 0) {
                    const int u = int(c[t] - r);
                    atomicOr(&bm[u >> 5], 1u << (u & 31));
                }
        }
    }
    __syncthreads();
}

// set bits of word i whose lower neighbour (bit i*32 - 1 for bit 0) is clear: the starts of runs
__device__ __forceinline__ unsigned run_starts(const unsigned* bm, int i) {
    const unsigned v = bm[i], prev = i ? bm[i - 1] : 0u;
    return v & ~((v << 1) | (prev >> 31));
}

// Exclusive block prefix of two counters (a, b) per thread; totals to ta, tb.
__device__ __forceinline__ void block_scan2(int& a, int& b, int& ta, int& tb, int* sh /* >= 16 */) {
    const int lane = threadIdx.x & 63, w = int(threadIdx.x >> 6), W = SYMB_THREADS / 64;
    int xa = a, xb = b;
    for (int o = 1; o < 64; o <<= 1) {
        const int ya = __shfl_up(xa, o), yb = __shfl_up(xb, o);
        if (lane >= o) { xa += ya; xb += yb; }
    }
    if (lane == 63) { sh[w] = xa; sh[W + w] = xb; }
    __syncthreads();
    int oa = 0, ob = 0;
    ta = 0; tb = 0;
    for (int i = 0; i < W; ++i) {
        if (i < w) { oa += sh[i]; ob += sh[W + i]; }
        ta += sh[i]; tb += sh[W + i];
    }
    a = xa - a + oa;
    b = xb - b + ob;
    __syncthreads();
}

// L = |U_s|, the number of bands, and the window length sum_b (64 + w_b - 1) = L + 63 bands.
__global__ __launch_bounds__(SYMB_THREADS) void k_sym_count(i64 rows, const i64* __restrict__ outer,
                                                            const i32* __restrict__ cols, const i64* __restrict__ diag,
                                                            int words, i64* __restrict__ L, i64* __restrict__ nb,
                                                            i64* __restrict__ W) {
    extern __shared__ unsigned bm[];
    __shared__ int sh[16];
    const i64 s = blockIdx.x;
    sym_bitmap(rows, s, outer, cols, diag, words, bm);
    int count = 0, runs = 0;
    for (int i = threadIdx.x; i < words; i += SYMB_THREADS) {
        count += __popc(bm[i]);
        runs += __popc(run_starts(bm, i));
    }
    int tc, tr;
    block_scan2(count, runs, tc, tr, sh);
    if (threadIdx.x == 0) {
        L[s] = tc;
        nb[s] = tr;
        W[s] = tc + 63 * i64(tr);
    }
}

// meta, bands and the values of slice s (sval zeroed beforehand).  With rank(u) the position of u in
// U_s and band(u) the number of run starts <= u less one, the window index of u is 63 band + rank:
// the strips before band b take sum_{b' < b} (64 + w_b' - 1) = 63 b + rank(u_b) entries, and ranks
// run on consecutively inside a band.
// Slots per LDS tile of the value placement in k_sym_fill (64 rows x 8 B each: 32 KB), and the tile's
// offset in the dynamic LDS, in words, after the three word arrays (bitmap, rank and band prefixes),
// rounded up to 8 bytes.
constexpr int SYMF_CHUNK = 64;
__host__ __device__ constexpr size_t symf_tile_offset(int words) { return (3 * size_t(words) + 1) & ~size_t(1); }
__host__ __device__ constexpr size_t symf_lds_bytes(int words) {
    return symf_tile_offset(words) * sizeof(unsigned) + size_t(SYMF_CHUNK) * 64 * sizeof(double);
}

__global__ __launch_bounds__(SYMB_THREADS) void k_sym_fill(i64 rows, const i64* __restrict__ outer,
                                                           const i32* __restrict__ cols, const i64* __restrict__ diag,
                                                           const double* __restrict__ val, int words,
                                                           const i64* __restrict__ kptr, const i64* __restrict__ bptr,
                                                           int2* __restrict__ meta, SymBand* __restrict__ band,
                                                           double* __restrict__ sval) {
    extern __shared__ unsigned bm[];          // words bitmap, words rank prefixes, words band prefixes
    unsigned* pre = bm + words;
    unsigned* bpre = pre + words;
    __shared__ int sh[16];
    __shared__ int start_rank[256], start_u[256];
    const i64 s = blockIdx.x;
    sym_bitmap(rows, s, outer, cols, diag, words, bm);
    const i64 kb = kptr[s], bb = bptr[s];
    // prefixes over the words: thread t takes the contiguous run [t cw, t cw + cw)
    const int cw = (words + SYMB_THREADS - 1) / SYMB_THREADS;
    const int w0 = min(words, int(threadIdx.x) * cw), w1 = min(words, w0 + cw);
    int cnt = 0, st = 0;
    for (int i = w0; i < w1; ++i) { cnt += __popc(bm[i]); st += __popc(run_starts(bm, i)); }
    int total, nbands;
    block_scan2(cnt, st, total, nbands, sh);
    for (int i = w0; i < w1; ++i) {
        pre[i] = unsigned(cnt);
        bpre[i] = unsigned(st);
        cnt += __popc(bm[i]);
        st += __popc(run_starts(bm, i));
    }
    __syncthreads();
    for (int i = w0; i < w1; ++i) {
        unsigned v = bm[i];
        const unsigned starts = run_starts(bm, i);
        while (v) {
            const int bit = __ffs(v) - 1;
            v &= v - 1;
            const int u = i * 32 + bit;
            const unsigned below = (1u << bit) - 1u, upto = (2u << bit) - 1u;   // bit 31: 2u << 31 wraps to 0
            const int rank = int(pre[i] + __popc(bm[i] & below));
            const int bnd = int(bpre[i] + __popc(starts & upto)) - 1;
            meta[kb + rank] = make_int2(u, u == 0 ? -1 : 63 * bnd + rank);
            if (starts & (1u << bit)) { start_rank[bnd] = rank; start_u[bnd] = u; }
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbands; b += SYMB_THREADS) {
        const int next = b + 1 < nbands ? start_rank[b + 1] : total;
        band[bb + b] = SymBand{start_u[b], 63 * b + start_rank[b], next - start_rank[b], 0};
    }
    // values: the slice's slots SYMF_CHUNK at a time, each chunk assembled in an LDS tile (slot-major,
    // 64 rows a slot, zero where a row lacks the offset) and then stored whole, so that every store
    // is coalesced and the padding needs no clearing pass of its own.  A row's upper entries are in
    // ascending offset, so in ascending slot rank: the entries of one chunk are the next run of the
    // row, which wave w (rows w, w + 4, ...) reads 64 at a time from the row's cursor.
    double (*tile)[64] = reinterpret_cast<double (*)[64]>(bm + symf_tile_offset(words));
    __shared__ i64 cursor[64];
    const int lane = threadIdx.x & 63, wv = int(threadIdx.x >> 6);
    if (threadIdx.x < 64) {
        const i64 r = s * 64 + threadIdx.x;
        cursor[threadIdx.x] = r < rows ? diag[r] : 0;
    }
    for (int k0 = 0; k0 < total; k0 += SYMF_CHUNK) {
        for (int e = threadIdx.x; e < SYMF_CHUNK * 64; e += SYMB_THREADS) tile[e >> 6][e & 63] = 0.0;
        __syncthreads();
        for (int j = wv; j < 64; j += SYMB_THREADS / 64) {
            const i64 r = s * 64 + j;
            if (r >= rows) break;
            const i64 e = outer[r + 1];
            i64 p0 = cursor[j];
            for (;;) {
                const i64 p = p0 + lane;
                int rank = 0x7fffffff;                         // past the row's end: in no chunk
                double v = 0.0;
                if (p < e) {
                    const int u = int(cols[p] - r);
                    rank = int(pre[u >> 5] + __popc(bm[u >> 5] & ((1u << (u & 31)) - 1u)));
                    v = val[p];
                }
                const bool in = rank < k0 + SYMF_CHUNK;
                if (in) tile[rank - k0][j] = v;
                const int n_in = __popcll(__ballot(in));
                p0 += n_in;
                if (n_in < 64) break;                          // the rest of the row lies past this chunk
            }
            if (lane == 0) cursor[j] = p0;
        }
        __syncthreads();
        const int nk = min(SYMF_CHUNK, total - k0);
        for (int e = threadIdx.x; e < nk * 64; e += SYMB_THREADS) sval[(kb + k0) * 64 + e] = tile[e >> 6][e & 63];
        __syncthreads();
    }
}

// Strips per band: the target slices [T0 / 64, (T0 + len - 1) / 64] of T0 = 64 s + u_b, len = 64 + w_b - 1.
__device__ __forceinline__ void band_span(i64 s, const SymBand& b, i64 ns, i64& t0, i64& t1) {
    const i64 T0 = s * 64 + b.u, len = 64 + b.w - 1;
    t0 = T0 / 64;
    t1 = std::min<i64>((T0 + len - 1) / 64, ns - 1);
}

// PASS 0: strips per band (cnt[band index]); PASS 1: keys (tau << 36 | s << 8 | b) at off[band index].
template<int PASS>
__global__ void k_sym_keys(i64 ns, const i64* __restrict__ bptr, const SymBand* __restrict__ band,
                           i64* __restrict__ cnt_or_off, unsigned long long* __restrict__ keys) {
    const i64 s = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    for (i64 q = bptr[s]; q < bptr[s + 1]; ++q) {
        i64 t0, t1;
        band_span(s, band[q], ns, t0, t1);
        if (PASS == 0) {
            cnt_or_off[q] = t1 - t0 + 1;
        } else {
            i64 o = cnt_or_off[q];
            for (i64 tau = t0; tau <= t1; ++tau)
                keys[o++] = (static_cast<unsigned long long>(tau) << 36) | (static_cast<unsigned long long>(s) << 8) |
                            static_cast<unsigned long long>(q - bptr[s]);
        }
    }
}

__global__ void k_sym_entries(i64 n, i64 ns, const unsigned long long* __restrict__ keys, const i64* __restrict__ bptr,
                              const SymBand* __restrict__ band, const i64* __restrict__ pbase,
                              SymEntry* __restrict__ ent, i64* __restrict__ tau_count) {
    const i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long key = keys[i];
    const i64 tau = i64(key >> 36), s = i64((key >> 8) & 0xfffffffull), b = i64(key & 0xffull);
    const SymBand q = band[bptr[s] + b];
    const i64 T0 = s * 64 + q.u, len = 64 + q.w - 1;
    SymEntry e;
    e.base = pbase[s] + q.wb + tau * 64 - T0;
    e.lo = int(std::max<i64>(0, T0 - tau * 64));
    e.hi = int(std::min<i64>(64, T0 + len - tau * 64));
    ent[i] = e;
    atomicAdd(reinterpret_cast<unsigned long long*>(&tau_count[tau]), 1ull);
}

}  // namespace nlf

// The storage of the symmetric SpMV of one matrix (FP64 values).
struct nlf_symv_s {
    nlf::i64 ns = 0;                     // slices of 64 rows
    int wmax = 0;                        // longest window (doubles)
    nlf::i64 slots = 0;                  // sum of |U_s| (values = 64 slots)
    nlf::i64 part_size = 0;              // doubles of the strips
    nlf::i64 version = -1;               // the matrix value version it was built from
    nlf::DevBuf<nlf::i64> kptr, pbase, rptr;
    nlf::DevBuf<int2> meta;
    nlf::DevBuf<double> sval, part, direct;
    nlf::DevBuf<nlf::SymEntry> ent;
};

namespace nlf {

void symv_release(nlf_matrix_s* a) {
    delete a->symv;
    a->symv = nullptr;
}

// Builds (or rebuilds, when the values changed) the symmetric SpMV storage of a FP64 matrix.  Returns
// false, with nothing kept, where it does not apply: offsets of 2^15 or more, more than 255 bands in a
// slice, a window over 2048 doubles, or padding that leaves less than a quarter of the traffic saved.
bool symv_prepare(nlf_matrix_s* a, hipStream_t s) {
    if (!a->ctx->spmv_symmetric || a->rows <= 0 || !a->val.p || a->nnz_full <= 0) return false;
    if (a->symv && a->symv->version == a->val_version && !a->values_exposed) return true;
    if (a->symv_refused_version == a->val_version) return false;
    symv_release(a);
    auto refuse = [&]() { a->symv_refused_version = a->val_version; return false; };
    const i64 rows = a->rows, ns = (rows + 63) / 64;
    if (ns >= (i64(1) << 27)) return refuse();
    // the widest upper offset
    DevBuf<unsigned> bw{1};
    NLF_HIP(hipMemsetAsync(bw.p, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_upper_width, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows, a->outer.p, a->cols.p, bw.p);
    unsigned hb = 0;
    download(&hb, bw.p, 1, s);
    if (hb >= 32768u) return refuse();
    const int words = int(hb / 32 + 1);
    DevBuf<i64> L{size_t(ns + 1)}, nb{size_t(ns + 1)}, W{size_t(ns + 1)};
    hipLaunchKernelGGL(k_sym_count, dim3(unsigned(ns)), dim3(SYMB_THREADS), size_t(words) * sizeof(unsigned), s, rows,
                       a->outer.p, a->cols.p, a->diag.p, words, L.p, nb.p, W.p);
    NLF_HIP(hipGetLastError());
    std::vector<i64> hL(ns), hnb(ns), hW(ns);
    download(hL.data(), L.p, size_t(ns), s);
    download(hnb.data(), nb.p, size_t(ns), s);
    download(hW.data(), W.p, size_t(ns), s);
    i64 slots = 0, maxw = 0, maxb = 0;
    for (i64 i = 0; i < ns; ++i) {
        slots += hL[i];
        maxw = std::max(maxw, hW[i]);
        maxb = std::max(maxb, hnb[i]);
    }
    const i64 upper_offdiag = a->nnz_upper - rows;
    // bytes: 8 per slot of 64 values and the strips twice, against 8 per full-storage value
    const double sym_bytes = double(slots) * 64.0 * 8.0 + 2.0 * 8.0 * double(std::accumulate(hW.begin(), hW.end(), i64(0)));
    if (maxb > 255 || maxw > 2048 || sym_bytes > 0.75 * 8.0 * double(a->nnz_full) || upper_offdiag <= 0) return refuse();
    auto sv = std::make_unique<nlf_symv_s>();
    sv->ns = ns;
    sv->wmax = int(maxw);
    sv->slots = slots;
    sv->part_size = 0;
    std::vector<i64> kp(ns + 1, 0), pbv(ns + 1, 0), bp(ns + 1, 0);
    for (i64 i = 0; i < ns; ++i) {
        kp[i + 1] = kp[i] + hL[i];
        pbv[i + 1] = pbv[i] + hW[i];
        bp[i + 1] = bp[i] + hnb[i];
    }
    upload(sv->kptr, kp.data(), kp.size(), s);
    upload(sv->pbase, pbv.data(), pbv.size(), s);
    sv->part_size = pbv[ns];
    DevBuf<i64> bptr;
    upload(bptr, bp.data(), bp.size(), s);
    sv->meta.alloc(size_t(std::max<i64>(slots, 1)));
    sv->sval.alloc(size_t(std::max<i64>(slots, 1)) * 64);
    sv->part.alloc(size_t(std::max<i64>(pbv[ns], 1)));
    sv->direct.alloc(size_t(rows));
    DevBuf<SymBand> band{size_t(std::max<i64>(bp[ns], 1))};
    hipLaunchKernelGGL(k_sym_fill, dim3(unsigned(ns)), dim3(SYMB_THREADS), symf_lds_bytes(words), s, rows,
                       a->outer.p, a->cols.p, a->diag.p, a->val.p, words, sv->kptr.p, bptr.p, sv->meta.p, band.p,
                       sv->sval.p);
    NLF_HIP(hipGetLastError());
    // strips per target slice, in (slice, band) order
    const i64 nbands = bp[ns];
    DevBuf<i64> koff{size_t(nbands + 1)};
    hipLaunchKernelGGL(k_sym_keys<0>, dim3(blocks_for(ns, 256)), dim3(256), 0, s, ns, bptr.p, band.p, koff.p,
                       (unsigned long long*)nullptr);
    exclusive_scan_i64(koff.p, nbands, s);
    i64 nkeys = 0;
    download(&nkeys, koff.p + nbands, 1, s);
    DevBuf<unsigned long long> keys{size_t(std::max<i64>(nkeys, 1))}, sorted{size_t(std::max<i64>(nkeys, 1))};
    hipLaunchKernelGGL(k_sym_keys<1>, dim3(blocks_for(ns, 256)), dim3(256), 0, s, ns, bptr.p, band.p, koff.p, keys.p);
    NLF_HIP(hipGetLastError());
    sort_keys_u64(keys.p, sorted.p, nkeys, s);
    sv->ent.alloc(size_t(std::max<i64>(nkeys, 1)));
    sv->rptr.alloc(size_t(ns + 1));
    NLF_HIP(hipMemsetAsync(sv->rptr.p, 0, sv->rptr.bytes(), s));
    hipLaunchKernelGGL(k_sym_entries, dim3(blocks_for(nkeys, 256)), dim3(256), 0, s, nkeys, ns, sorted.p, bptr.p, band.p,
                       sv->pbase.p, sv->ent.p, sv->rptr.p);
    NLF_HIP(hipGetLastError());
    exclusive_scan_i64(sv->rptr.p, ns, s);
    NLF_HIP(hipStreamSynchronize(s));
    sv->version = a->val_version;
    a->symv = sv.release();
    return true;
}

// y = A x from the symmetric storage (symv_prepare first); DOT: partial[0 .. blocks) of x.y, the
// reduction running on `blocks` blocks (any count: the caller's partial count).
template<bool DOT>
void launch_symv(const nlf_matrix_s* a, const double* x, double* y, double* partial, const CGState* st, hipStream_t s,
                 unsigned blocks) {
    const nlf_symv_s& v = *a->symv;
    const unsigned b1 = unsigned((v.ns + SYMV_WAVES - 1) / SYMV_WAVES);
    const size_t lds = size_t(SYMV_WAVES) * size_t(v.wmax) * sizeof(double);
    if (st)
        hipLaunchKernelGGL(k_symv_slices<true>, dim3(b1), dim3(64 * SYMV_WAVES), lds, s, a->rows, v.ns, v.wmax, v.kptr.p,
                           v.meta.p, v.sval.p, v.pbase.p, x, v.direct.p, v.part.p, st);
    else
        hipLaunchKernelGGL(k_symv_slices<false>, dim3(b1), dim3(64 * SYMV_WAVES), lds, s, a->rows, v.ns, v.wmax, v.kptr.p,
                           v.meta.p, v.sval.p, v.pbase.p, x, v.direct.p, v.part.p, st);
    const unsigned b2 = std::max(1u, blocks);
    if (st)
        hipLaunchKernelGGL((k_symv_reduce<DOT, true>), dim3(b2), dim3(256), 0, s, a->rows, v.ns, v.rptr.p, v.ent.p,
                           v.direct.p, v.part.p, x, y, partial, st);
    else
        hipLaunchKernelGGL((k_symv_reduce<DOT, false>), dim3(b2), dim3(256), 0, s, a->rows, v.ns, v.rptr.p, v.ent.p,
                           v.direct.p, v.part.p, x, y, partial, st);
}

}  // namespace nlf
