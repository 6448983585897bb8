// Aggregation coarse space for the two-level learned preconditioner (no reference counterpart:
// the reference's preconditioners are Eigen's diagonal and incomplete Cholesky).
//
//   B r = M r + P A_c^-1 P^T r
//
// P maps every fine row to one aggregate: the nodes are binned into a grid of square cells over
// their bounding box (one aggregate per non-empty cell and DoF), so P^T r sums a cell's rows and
// P y copies a cell's value back.  A_c = P^T A P is assembled densely by one workgroup per coarse
// row; its entries are accumulated as 64-bit fixed-point integers in LDS, which makes the sum
// independent of the order of the atomic additions (the result is reproducible bit for bit).
// A_c^-1 is formed once by blocked Gauss-Jordan elimination on 64 x 64 tiles (A_c is symmetric
// positive definite, so no pivoting), and each application is three small kernels: restriction,
// a dense matrix-vector product with A_c^-1, and prolongation.
#include "nlf_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace nlf {

namespace {

constexpr int TB = 64;                     // tile of the blocked Gauss-Jordan elimination

// Per-block maxima over rows of sum_j |a_ij|: one wave per row, waves striding over the rows.
__global__ __launch_bounds__(256) void k_row_abs_max(i64 rows, const i64* __restrict__ outer,
                                                     const double* __restrict__ val, double* __restrict__ part) {
    __shared__ double wmax[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double best = 0.0;
    for (i64 r = i64(blockIdx.x) * 4 + wv; r < rows; r += i64(gridDim.x) * 4) {
        double s = 0.0;
        for (i64 p = outer[r] + lane; p < outer[r + 1]; p += 64) s += fabs(val[p]);
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        best = fmax(best, s);
    }
    if (lane == 0) wmax[wv] = best;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
}

// Row I of A_c = P^T A P: the nonzeros of I's fine rows, added in fixed point into an LDS row.
__global__ __launch_bounds__(256) void k_coarse_galerkin(i64 ncp, const i64* __restrict__ mem_off,
                                                         const i32* __restrict__ mem, const i64* __restrict__ outer,
                                                         const i32* __restrict__ cols, const double* __restrict__ val,
                                                         const i32* __restrict__ agg, double scale,
                                                         double* __restrict__ ac) {
    extern __shared__ unsigned long long acc[];                // ncp entries
    const i64 I = blockIdx.x;
    for (i64 j = threadIdx.x; j < ncp; j += blockDim.x) acc[j] = 0ull;
    __syncthreads();
    for (i64 m = mem_off[I]; m < mem_off[I + 1]; ++m) {
        const i32 r = mem[m];
        for (i64 p = outer[r] + threadIdx.x; p < outer[r + 1]; p += blockDim.x) {
            const long long q = llrint(val[p] * scale);
            atomicAdd(&acc[agg[cols[p]]], static_cast<unsigned long long>(q));
        }
    }
    __syncthreads();
    const double inv_scale = 1.0 / scale;
    for (i64 j = threadIdx.x; j < ncp; j += blockDim.x)
        ac[I * ncp + j] = double(static_cast<long long>(acc[j])) * inv_scale;
}

__global__ void k_unit_diag(i64 from, i64 ncp, double* __restrict__ a) {
    const i64 i = from + i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < ncp) a[i * ncp + i] = 1.0;
}

// In-place Gauss-Jordan inverse of the 64 x 64 diagonal tile (k, k) into dinv (no pivoting);
// GJ_DIAG_THREADS threads (1024: 256 and 64 threads were measured slower), 4096 / GJ_DIAG_THREADS
// entries each.  Per pivot p every thread reads what its entries need (the pivot, its rows' entries
// in column p, its column's entry in row p), computes the new values -- row p scaled by the pivot,
// the other rows eliminated with the scaled row, each thread scaling its own column's entry -- and
// writes them after a barrier: two barriers per pivot.  The operations are those of the textbook
// element-wise form, in the same order.
#ifndef NLF_GJ_DIAG_THREADS
#define NLF_GJ_DIAG_THREADS 1024
#endif
constexpr int GJ_DIAG_THREADS = NLF_GJ_DIAG_THREADS;
__global__ __launch_bounds__(GJ_DIAG_THREADS) void k_gj_diag(const double* __restrict__ a, i64 ncp, i64 k, double* __restrict__ dinv) {
    constexpr int PER = TB * TB / GJ_DIAG_THREADS;
    static_assert(TB * TB % GJ_DIAG_THREADS == 0 && GJ_DIAG_THREADS % TB == 0, "whole rows per pass");
    __shared__ double d[TB][TB + 1];
    const int t = threadIdx.x;
    const int j = t % TB, i0 = t / TB;                        // column j, rows i0 + u * (threads / TB)
    constexpr int RS = GJ_DIAG_THREADS / TB;
    for (int e = t; e < TB * TB; e += GJ_DIAG_THREADS) d[e / TB][e % TB] = a[(k * TB + e / TB) * ncp + k * TB + e % TB];
    __syncthreads();
    for (int p = 0; p < TB; ++p) {
        const double piv = d[p][p];
        const double rp = (j == p ? 1.0 : d[p][j]) / piv;     // the scaled row p at column j
        double nv[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = i0 + u * RS;
            nv[u] = i == p ? rp : (j == p ? 0.0 : d[i][j]) - d[i][p] * rp;
        }
        __syncthreads();                                      // every read of this pivot is done
#pragma unroll
        for (int u = 0; u < PER; ++u) d[i0 + u * RS][j] = nv[u];
        __syncthreads();
    }
    for (int e = t; e < TB * TB; e += GJ_DIAG_THREADS) dinv[e] = d[e / TB][e % TB];
}

typedef double gj_d4 __attribute__((ext_vector_type(4)));

// acc = X (64 x 64, leading dimension ldx) * Y (64 x 64, ldy) with V_MFMA_F64_16X16X4_F64.  Lane l
// of a wave holds A[l % 16][l / 16] and B[l / 16][l % 16] of a 16 x 4 by 4 x 16 product and
// D[l / 16 + 4 r][l % 16] in register r (measured on gfx950).  Wave w computes the 32 x 32 quadrant
// (w / 2, w % 2) as 2 x 2 blocks of 16 x 16.  The inner dimension is staged through LDS in two halves
// of 32, XOR-swizzled by row so that the operand reads of one instruction hit distinct banks.  Every
// read of X and Y is done when it returns (the last staging is followed by a barrier).
__device__ __forceinline__ void gj_tile_mfma(const double* __restrict__ X, i64 ldx, const double* __restrict__ Y, i64 ldy,
                                             gj_d4 (&acc)[2][2], double (*sx)[32], double (*sy)[64]) {
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int li = l & 15, lk = l >> 4;
    const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
    for (int u = 0; u < 2; ++u)
        for (int v = 0; v < 2; ++v) acc[u][v] = gj_d4{0.0, 0.0, 0.0, 0.0};
    for (int half = 0; half < 2; ++half) {
        __syncthreads();
        for (int e = t; e < 64 * 32; e += 256) {
            const int i = e >> 5, j = e & 31;
            sx[i][j ^ ((2 * i) & 31)] = X[i64(i) * ldx + 32 * half + j];
        }
        for (int e = t; e < 32 * 64; e += 256) {
            const int kk = e >> 6, j = e & 63;
            sy[kk][j ^ (16 * (kk & 3))] = Y[i64(32 * half + kk) * ldy + j];
        }
        __syncthreads();
        for (int ks = 0; ks < 32; ks += 4) {
            const int kk = ks + lk;
            double av[2], bv[2];
            for (int b = 0; b < 2; ++b) {
                const int i = r0 + 16 * b + li;
                av[b] = sx[i][kk ^ ((2 * i) & 31)];
                bv[b] = sy[kk][(c0 + 16 * b + li) ^ (16 * (kk & 3))];
            }
            for (int u = 0; u < 2; ++u)
                for (int v = 0; v < 2; ++v) acc[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[v], acc[u][v], 0, 0, 0);
        }
    }
}

// C (64 x 64, ldc) op= acc in the layout of gj_tile_mfma: SIGN -1 subtracts, +1 stores, 0 stores -acc.
template<int SIGN>
__device__ __forceinline__ void gj_tile_store(double* __restrict__ C, i64 ldc, const gj_d4 (&acc)[2][2]) {
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int li = l & 15, lk = l >> 4;
    const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
    for (int u = 0; u < 2; ++u)
        for (int v = 0; v < 2; ++v)
            for (int r = 0; r < 4; ++r) {
                double& c = C[i64(r0 + 16 * u + lk + 4 * r) * ldc + c0 + 16 * v + li];
                if constexpr (SIGN < 0) c -= acc[u][v][r];
                else if constexpr (SIGN > 0) c = acc[u][v][r];
                else c = -acc[u][v][r];
            }
}

// Row panel: A(k, j) = D^-1 A(k, j) for every tile column j != k.
__global__ __launch_bounds__(256) void k_gj_rowpanel(double* __restrict__ a, i64 ncp, i64 k, const double* __restrict__ dinv) {
    __shared__ double sx[64][32], sy[32][64];
    const i64 nb = ncp / TB;
    const i64 j = blockIdx.x < k ? blockIdx.x : blockIdx.x + 1;
    if (j >= nb) return;
    double* blk = a + (k * TB) * ncp + j * TB;
    gj_d4 acc[2][2];
    gj_tile_mfma(dinv, TB, blk, ncp, acc, sx, sy);
    __syncthreads();                                              // every read of the tile is done
    gj_tile_store<1>(blk, ncp, acc);
}

// Trailing update: A(i, j) -= A(i, k) A(k, j) for tile rows i != k and tile columns j != k.
__global__ __launch_bounds__(256) void k_gj_update_mfma(double* __restrict__ a, i64 ncp, i64 k) {
    __shared__ double sx[64][32], sy[32][64];
    const i64 bi = blockIdx.y < k ? blockIdx.y : blockIdx.y + 1;
    const i64 bj = blockIdx.x < k ? blockIdx.x : blockIdx.x + 1;
    const i64 nb = ncp / TB;
    if (bi >= nb || bj >= nb) return;
    gj_d4 acc[2][2];
    gj_tile_mfma(a + (bi * TB) * ncp + k * TB, ncp, a + (k * TB) * ncp + bj * TB, ncp, acc, sx, sy);
    gj_tile_store<-1>(a + (bi * TB) * ncp + bj * TB, ncp, acc);
}

// Column panel: A(i, k) = -A(i, k) D^-1 for tile rows i != k.
__global__ __launch_bounds__(256) void k_gj_colpanel(double* __restrict__ a, i64 ncp, i64 k, const double* __restrict__ dinv) {
    __shared__ double sx[64][32], sy[32][64];
    const i64 nb = ncp / TB;
    const i64 i = blockIdx.x < k ? blockIdx.x : blockIdx.x + 1;
    if (i >= nb) return;
    double* blk = a + (i * TB) * ncp + k * TB;
    gj_d4 acc[2][2];
    gj_tile_mfma(blk, ncp, dinv, TB, acc, sx, sy);
    __syncthreads();
    gj_tile_store<0>(blk, ncp, acc);
}

__global__ void k_gj_setdiag(double* __restrict__ a, i64 ncp, i64 k, const double* __restrict__ dinv) {
    for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) a[(k * TB + e / TB) * ncp + k * TB + e % TB] = dinv[e];
}

// rc[I] = sum of r over the fine rows of I (one wave per coarse row, fixed order).
__global__ __launch_bounds__(256) void k_coarse_restrict(i64 nc, const i64* __restrict__ mem_off,
                                                         const i32* __restrict__ mem, const double* __restrict__ r,
                                                         double* __restrict__ rc) {
    const i64 I = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (I >= nc) return;
    double s = 0.0;
    for (i64 m = mem_off[I] + lane; m < mem_off[I + 1]; m += 64) s += r[mem[m]];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) rc[I] = s;
}

// yc = A_c^-1 rc (one wave per row, fixed order).
__global__ __launch_bounds__(256) void k_coarse_gemv(i64 ncp, const double* __restrict__ inv,
                                                     const double* __restrict__ rc, double* __restrict__ yc) {
    const i64 I = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (I >= ncp) return;
    // 16-byte loads, lane l takes the column pairs l, l + 64, ... (ncp is a multiple of 64), two sums
    const double2* row = reinterpret_cast<const double2*>(inv + I * ncp);
    const double2* r2 = reinterpret_cast<const double2*>(rc);
    const i64 n2 = ncp / 2;
    double s0 = 0.0, s1 = 0.0;
#pragma unroll 4
    for (i64 j = lane; j < n2; j += 64) {
        const double2 a = row[j], b = r2[j];
        s0 += a.x * b.x;
        s1 += a.y * b.y;
    }
    double s = s0 + s1;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) yc[I] = s;
}

__global__ void k_coarse_prolong(i64 n, const i32* __restrict__ agg, const double* __restrict__ yc, double* __restrict__ z) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) z[i] += yc[agg[i]];
}

// Cell of every node on the cx x cy grid (the operations of the host form: offset, division,
// truncation, clamp to the last cell), and the used cells marked.
__global__ void k_cells(i64 N, const double* __restrict__ xy, double xmin, double ymin, double hx, double hy, i64 cx,
                        i64 cy, i64* __restrict__ cell, i64* __restrict__ used) {
    const i64 v = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const i64 fx = i64((xy[2 * v] - xmin) / hx), fy = i64((xy[2 * v + 1] - ymin) / hy);
    const i64 ix = fx < cx - 1 ? fx : cx - 1, iy = fy < cy - 1 ? fy : cy - 1;
    cell[v] = ix * cy + iy;
    used[ix * cy + iy] = 1;
}

// Coarse row of every fine row (the used cells numbered in cell order, times dof, plus the fine
// row's component), the sort key of the members list, and the per-aggregate counts.
__global__ void k_aggregate(i64 n, int dof, const i64* __restrict__ cell, const i64* __restrict__ cid,
                            i32* __restrict__ agg, uint64_t* __restrict__ key, i32* __restrict__ row,
                            unsigned long long* __restrict__ count) {
    const i64 r = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const i32 I = i32(cid[cell[r / dof]] * dof + r % dof);
    agg[r] = I;
    key[r] = uint64_t(I);
    row[r] = i32(r);
    atomicAdd(&count[I], 1ull);
}

__global__ void k_to_i64(i64 n, const unsigned long long* __restrict__ in, i64* __restrict__ out) {
    const i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i64(in[i]);
}

}  // namespace

std::unique_ptr<CoarseSpace> build_coarse(const nlf_matrix_s* a, i64 target, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    const nlf_mesh m = a->mesh;
    const int dof = a->dof;
    const i64 N = m->N, n = a->rows;
    NLF_CHECK(target > 0, NLF_ERR_ARGUMENT, "coarse target must be positive");
    NLF_CHECK(N > 0 && n > 0, NLF_ERR_ARGUMENT, "the coarse space needs a mesh with nodes");
    // at least 16 nodes per aggregate on average, at most 4096 coarse rows (dense inverse in LDS-tiled
    // Gauss-Jordan, one 32 KiB LDS row per coarse row in the Galerkin product)
    target = std::min<i64>(target, std::max<i64>(1, N / 16));
    target = std::min<i64>(target, 4096 / dof);
    const double xmin = m->bbox[0], xmax = m->bbox[1], ymin = m->bbox[2], ymax = m->bbox[3];
    const double w = std::max(xmax - xmin, 1e-300), h = std::max(ymax - ymin, 1e-300);
    const i64 cx = std::min<i64>(target, std::max<i64>(1, i64(std::lround(std::sqrt(double(target) * w / h)))));
    const i64 cy = std::max<i64>(1, target / cx);
    const double hx = w / double(cx), hy = h / double(cy);
    // on the device: the cell of every node, the non-empty cells numbered in cell order, the coarse
    // row of every fine row, and the members of every aggregate, rows ascending (stable sort by
    // coarse row)
    DevBuf<i64> cell{size_t(N)}, cid{size_t(cx * cy + 1)};
    NLF_HIP(hipMemsetAsync(cid.p, 0, cid.bytes(), s));
    hipLaunchKernelGGL(k_cells, dim3(blocks_for(N, 256)), dim3(256), 0, s, N, m->xy.p, xmin, ymin, hx, hy, cx, cy, cell.p,
                       cid.p);
    exclusive_scan_i64(cid.p, cx * cy, s);
    i64 nagg = 0;
    download(&nagg, cid.p + cx * cy, 1, s);
    auto cs = std::make_unique<CoarseSpace>();
    cs->n = n;
    cs->nc = nagg * dof;
    cs->ncp = (cs->nc + TB - 1) / TB * TB;
    cs->agg.alloc(size_t(n));
    cs->mem.alloc(size_t(n));
    cs->mem_off.alloc(size_t(cs->nc + 1));
    {
        DevBuf<uint64_t> key{size_t(n)}, key_sorted{size_t(n)};
        DevBuf<i32> row{size_t(n)};
        DevBuf<unsigned long long> count{size_t(cs->nc)};
        NLF_HIP(hipMemsetAsync(count.p, 0, count.bytes(), s));
        hipLaunchKernelGGL(k_aggregate, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, dof, cell.p, cid.p, cs->agg.p, key.p,
                           row.p, count.p);
        hipLaunchKernelGGL(k_to_i64, dim3(blocks_for(cs->nc, 256)), dim3(256), 0, s, cs->nc, count.p, cs->mem_off.p);
        NLF_HIP(hipGetLastError());
        exclusive_scan_i64(cs->mem_off.p, cs->nc, s);
        sort_pairs_u64(key.p, key_sorted.p, row.p, cs->mem.p, n, s);
    }
    std::vector<i64> cnt(size_t(cs->nc + 1));
    download(cnt.data(), cs->mem_off.p, cnt.size(), s);
    // fixed-point scale: every partial sum of a coarse entry is bounded by (rows per aggregate) x
    // (largest absolute row sum); keep it below 2^61
    i64 maxmem = 0;
    for (i64 I = 0; I < cs->nc; ++I) maxmem = std::max(maxmem, cnt[size_t(I + 1)] - cnt[size_t(I)]);
    constexpr int ABS_BLOCKS = 2048;
    DevBuf<double> d_abs{ABS_BLOCKS};
    hipLaunchKernelGGL(k_row_abs_max, dim3(ABS_BLOCKS), dim3(256), 0, s, n, a->outer.p, a->val.p, d_abs.p);
    std::vector<double> h_abs(ABS_BLOCKS);
    download(h_abs.data(), d_abs.p, h_abs.size(), s);
    const double rowabs = *std::max_element(h_abs.begin(), h_abs.end());
    const double bound = std::max(1e-300, rowabs * double(std::max<i64>(maxmem, 1)));
    const double scale = std::ldexp(1.0, std::clamp(61 - int(std::ceil(std::log2(bound))), -1000, 1000));
    cs->inv.alloc(size_t(cs->ncp * cs->ncp));
    NLF_HIP(hipMemsetAsync(cs->inv.p, 0, cs->inv.bytes(), s));
    hipLaunchKernelGGL(k_coarse_galerkin, dim3(unsigned(cs->nc)), dim3(256), size_t(cs->ncp) * sizeof(unsigned long long), s,
                       cs->ncp, cs->mem_off.p, cs->mem.p, a->outer.p, a->cols.p, a->val.p, cs->agg.p, scale, cs->inv.p);
    if (cs->ncp > cs->nc)
        hipLaunchKernelGGL(k_unit_diag, dim3(1), dim3(64), 0, s, cs->nc, cs->ncp, cs->inv.p);
    NLF_HIP(hipGetLastError());
    // blocked Gauss-Jordan inversion in place
    const i64 nb = cs->ncp / TB;
    DevBuf<double> dinv{size_t(TB * TB)};
    for (i64 k = 0; k < nb; ++k) {
        hipLaunchKernelGGL(k_gj_diag, dim3(1), dim3(GJ_DIAG_THREADS), 0, s, cs->inv.p, cs->ncp, k, dinv.p);
        if (nb > 1) {
            hipLaunchKernelGGL(k_gj_rowpanel, dim3(unsigned(nb - 1)), dim3(256), 0, s, cs->inv.p, cs->ncp, k, dinv.p);
            hipLaunchKernelGGL(k_gj_update_mfma, dim3(unsigned(nb - 1), unsigned(nb - 1)), dim3(256), 0, s, cs->inv.p, cs->ncp, k);
            hipLaunchKernelGGL(k_gj_colpanel, dim3(unsigned(nb - 1)), dim3(256), 0, s, cs->inv.p, cs->ncp, k, dinv.p);
        }
        hipLaunchKernelGGL(k_gj_setdiag, dim3(1), dim3(256), 0, s, cs->inv.p, cs->ncp, k, dinv.p);
    }
    NLF_HIP(hipGetLastError());
    cs->rc.alloc(size_t(cs->ncp));
    cs->yc.alloc(size_t(cs->ncp));
    NLF_HIP(hipMemsetAsync(cs->rc.p, 0, cs->rc.bytes(), s));
    NLF_HIP(hipStreamSynchronize(s));
    cs->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return cs;
}

void coarse_apply_add(const CoarseSpace& c, const double* r, double* z, hipStream_t s) {
    hipLaunchKernelGGL(k_coarse_restrict, dim3(blocks_for(c.nc, 4)), dim3(256), 0, s, c.nc, c.mem_off.p, c.mem.p, r, c.rc.p);
    hipLaunchKernelGGL(k_coarse_gemv, dim3(blocks_for(c.ncp, 4)), dim3(256), 0, s, c.ncp, c.inv.p, c.rc.p, c.yc.p);
    hipLaunchKernelGGL(k_coarse_prolong, dim3(unsigned(std::min<i64>(blocks_for(c.n, 256), 4096))), dim3(256), 0, s, c.n,
                       c.agg.p, c.yc.p, z);
}

}  // namespace nlf
