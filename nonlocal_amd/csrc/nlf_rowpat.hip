// Shared row patterns for the SpMV: a dictionary of the distinct column-offset rows.
//
// On a mesh whose nodes are numbered along grid lines, most rows of the nonlocal matrix have the
// same sequence of column offsets (col - row): every interior node of a uniform grid couples with
// the same disc of nodes.  Storing each distinct offset sequence once, plus one start index per
// row, removes the 2 bytes per nonzero of the 16-bit offsets from the SpMV's HBM stream for the
// rows that share a pattern (the shared patterns stay in L2), so those rows read 8 bytes per
// nonzero (FP64 values) or 4 (FP32) from HBM.  Each pattern is stored twice, at an even and at an
// odd start, so that every row's offsets can be loaded in aligned pairs.  Rows are grouped by a 64-bit hash of (length,
// offsets); identical hashes are then compared entry by entry, and any mismatch abandons the
// dictionary (the 16-bit offsets stay in use), so a hash collision can cost speed but never change
// a result.
#include "nlf_internal.hpp"

#include <rocprim/rocprim.hpp>

namespace nlf {

namespace {

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {   // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// One wave per row: hash = mix(length) + sum_q mix(q << 16 | offset_q) (position-tagged, so the sum
// is order-sensitive), reduced across the wave.
__global__ __launch_bounds__(256) void k_row_hash(i64 rows, const i64* __restrict__ outer,
                                                  const int16_t* __restrict__ off16, unsigned long long* __restrict__ key,
                                                  i32* __restrict__ idx) {
    const i64 r = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    const i64 b = outer[r], e = outer[r + 1];
    unsigned long long h = 0;
    for (i64 p = b + lane; p < e; p += 64) {
        const unsigned long long q = static_cast<unsigned long long>(p - b);
        h += mix64((q << 16) | static_cast<unsigned short>(off16[p]));
    }
    for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
    if (lane == 0) {
        key[r] = h + mix64(0x9e3779b97f4a7c15ull ^ static_cast<unsigned long long>(e - b));
        idx[r] = i32(r);
    }
}

// Sorted position i starts a group when its hash differs from position i - 1; otherwise the two
// rows must be identical (length and every offset), else *bad is set.
__global__ __launch_bounds__(256) void k_group_heads(i64 rows, const unsigned long long* __restrict__ key,
                                                     const i32* __restrict__ idx, const i64* __restrict__ outer,
                                                     const int16_t* __restrict__ off16, i64* __restrict__ head,
                                                     int* __restrict__ bad) {
    const i64 i = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= rows) return;
    const bool h = i == 0 || key[i] != key[i - 1];
    if (lane == 0) head[i] = h ? 1 : 0;
    if (h) return;
    const i32 r = idx[i], s = idx[i - 1];
    const i64 br = outer[r], lr = outer[r + 1] - br;
    const i64 bs = outer[s], ls = outer[s + 1] - bs;
    bool same = lr == ls;
    if (same)
        for (i64 q = lane; q < lr; q += 64) same &= off16[br + q] == off16[bs + q];
    if (!same) *bad = 1;
}

// head[] holds the exclusive scan of the group-start flags: group id of sorted position i is
// head[i + 1] - 1.  Group g records its representative row and the size of its slot: two copies
// of the pattern (2 len + 2 entries, see pattern_copy).
__global__ void k_group_rows(i64 rows, const i64* __restrict__ gscan, const i32* __restrict__ idx,
                             const i64* __restrict__ outer, i32* __restrict__ grow, i64* __restrict__ gslot) {
    const i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    if (gscan[i + 1] != gscan[i]) {                 // position i starts group gscan[i]
        const i64 g = gscan[i];
        const i32 r = idx[i];
        grow[g] = r;
        gslot[g] = 2 * (outer[r + 1] - outer[r]) + 2;
    }
}

// The SpMV loads offsets in pairs at even dictionary indices; a row starting at an odd nonzero
// index needs its pattern at an odd start.  Each group's slot (even start) holds copy 0 at an
// even and copy 1 at an odd position.
__device__ __forceinline__ i64 pattern_copy(i64 slot, i64 len, int parity) {
    return parity == 0 ? slot : slot + len + 1 + (len & 1);
}

__global__ __launch_bounds__(256) void k_fill_patterns(i64 groups, const i32* __restrict__ grow,
                                                       const i64* __restrict__ gbeg, const i64* __restrict__ outer,
                                                       const int16_t* __restrict__ off16, int16_t* __restrict__ pat) {
    const i64 g = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= groups) return;
    const i32 r = grow[g];
    const i64 b = outer[r], len = outer[r + 1] - b;
    const i64 d0 = pattern_copy(gbeg[g], len, 0), d1 = pattern_copy(gbeg[g], len, 1);
    for (i64 q = lane; q < len; q += 64) {
        const int16_t o = off16[b + q];
        pat[d0 + q] = o;
        pat[d1 + q] = o;
    }
}

__global__ void k_row_starts(i64 rows, const i64* __restrict__ gscan, const i32* __restrict__ idx,
                             const i64* __restrict__ gbeg, const i64* __restrict__ outer, i64* __restrict__ pstart) {
    const i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    const i32 r = idx[i];
    // the SpMV reads pat[pstart[r] + p] for p in [outer[r], outer[r+1]): the row's copy has the
    // parity of outer[r], so pstart[r] is even and pairs at even p are 4-byte aligned
    const i64 b = outer[r];
    pstart[r] = pattern_copy(gbeg[gscan[i + 1] - 1], outer[r + 1] - b, int(b & 1)) - b;
}

}  // namespace

bool build_row_patterns(i64 rows, i64 nnz, const i64* outer, const int16_t* off16, DevBuf<int16_t>& pat,
                        DevBuf<i64>& pstart, i64& entries, i64& groups_out, hipStream_t s) {
    pat.release();
    pstart.release();
    entries = 0;
    groups_out = 0;
    if (!off16 || rows < 2 || rows >= (i64(1) << 31)) return false;
    DevBuf<unsigned long long> key{size_t(rows)}, key_s{size_t(rows)};
    DevBuf<i32> idx{size_t(rows)}, idx_s{size_t(rows)};
    hipLaunchKernelGGL(k_row_hash, dim3(blocks_for(rows, 4)), dim3(256), 0, s, rows, outer, off16, key.p, idx.p);
    NLF_HIP(hipGetLastError());
    size_t bytes = 0;
    NLF_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key.p, key_s.p, idx.p, idx_s.p, size_t(rows), 0, 64, s));
    DevBuf<uint8_t> tmp{bytes + 16};
    NLF_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, key.p, key_s.p, idx.p, idx_s.p, size_t(rows), 0, 64, s));
    DevBuf<i64> gscan{size_t(rows + 1)};
    DevBuf<int> bad{1};
    NLF_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_group_heads, dim3(blocks_for(rows, 4)), dim3(256), 0, s, rows, key_s.p, idx_s.p, outer, off16,
                       gscan.p, bad.p);
    NLF_HIP(hipGetLastError());
    int h_bad = 0;
    download(&h_bad, bad.p, 1, s);
    if (h_bad) return false;                          // hash collision: keep the per-nonzero offsets
    exclusive_scan_i64(gscan.p, rows, s);
    i64 groups = 0;
    download(&groups, gscan.p + rows, 1, s);
    DevBuf<i32> grow{size_t(groups)};
    DevBuf<i64> gbeg{size_t(groups + 1)};
    NLF_HIP(hipMemsetAsync(gbeg.p, 0, gbeg.bytes(), s));
    hipLaunchKernelGGL(k_group_rows, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows, gscan.p, idx_s.p, outer, grow.p,
                       gbeg.p);
    NLF_HIP(hipGetLastError());
    exclusive_scan_i64(gbeg.p, groups, s);
    i64 total = 0;
    download(&total, gbeg.p + groups, 1, s);
    // kept when it saves at least half of the offsets (the slots hold two copies of each pattern):
    // rows with a pattern of their own (near material interfaces, Dirichlet and outer boundaries)
    // then stream it as they streamed their offsets, the shared ones read it from L2
    if (total > nnz) return false;
    pat.alloc(size_t(std::max<i64>(total, 1)));
    pstart.alloc(size_t(rows));
    hipLaunchKernelGGL(k_fill_patterns, dim3(blocks_for(groups, 4)), dim3(256), 0, s, groups, grow.p, gbeg.p, outer, off16,
                       pat.p);
    hipLaunchKernelGGL(k_row_starts, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows, gscan.p, idx_s.p, gbeg.p, outer,
                       pstart.p);
    NLF_HIP(hipGetLastError());
    NLF_HIP(hipStreamSynchronize(s));
    entries = total;
    groups_out = groups;
    return true;
}

void build_row_patterns(nlf_matrix_s* a, hipStream_t s) {
    build_row_patterns(a->rows, a->nnz_full, a->outer.p, a->off16.p, a->pat, a->pstart, a->pat_size, a->pat_groups, s);
}

}  // namespace nlf
