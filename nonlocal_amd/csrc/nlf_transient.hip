// Pieces of the nonstationary heat solver (nonstationary_heat_equation_solver_2d,
// solvers/solver_2d/thermal/nonstationary_heat_equation_solver_2d.hpp:17-106) and of the source term
// of every solver (integrate_right_part, solvers/solver_2d/base/right_part_2d.hpp:10-35):
//
//   * nlf_right_part: f_i += sum over the node's elements (ascending ids) of
//     sum_q w_q N_i(q) |J_q| s(x_q), one thread per node, so each f_i has one fixed summation order.
//     The caller gives s at the quadrature points (it evaluates the function on its side).
//   * nlf_matrix_combine: K := alpha K + beta C on K's pattern, the bound part scaled by alpha, the
//     Dirichlet rows kept as identity rows -- the "inner() *= tau; bound() *= tau; inner() += C;
//     diagonal = 1" of compute() (:76-83).  C's pattern must lie within K's (a local matrix in a
//     nonlocal one); an entry of C outside K's pattern is an argument error, found by a pass with
//     no writes before K is touched.
//
// The time step itself (right-hand side tau f + C T_n, first-kind conditions, CG from T_n) is
// composed by the host from these, nlf_spmv, nlf_apply_first_kind and nlf_cg_solve.
#include "nlf_internal.hpp"

namespace nlf {
namespace {

template<int NN, int NQ>
__global__ __launch_bounds__(256) void k_right_part(i64 N, int dof, const i32* __restrict__ ne_off,
                                                    const i32* __restrict__ ne_ids, const i32* __restrict__ el,
                                                    const double* __restrict__ jac, FeTables t,
                                                    const double* __restrict__ s, double* __restrict__ f) {
    const i64 node = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= N) return;
    double acc[2] = {0.0, 0.0};
    for (i32 k = ne_off[node]; k < ne_off[node + 1]; ++k) {
        const i64 e = ne_ids[k];
        int i = 0;
        for (int j = 0; j < NN; ++j)
            if (el[e * NN + j] == node) i = j;
        double integral[2] = {0.0, 0.0};
        for (int q = 0; q < NQ; ++q) {
            const double wq = t.w[q] * t.qN[i * NQ + q] * jac[e * NQ + q];
            for (int d = 0; d < dof; ++d) integral[d] += wq * s[(e * NQ + q) * dof + d];
        }
        for (int d = 0; d < dof; ++d) acc[d] += integral[d];
    }
    for (int d = 0; d < dof; ++d) f[node * dof + d] += acc[d];
}

__global__ void k_scale(i64 n, double alpha, double* __restrict__ v) {
    const i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] *= alpha;
}

// one wave per row: every entry of C's row r is found in K's row r by a binary search of its column;
// WRITE = false only flags an entry that K's row lacks, WRITE = true adds beta C_rc into K
template<bool WRITE>
__global__ __launch_bounds__(256) void k_add_rows(i64 rows, const i64* __restrict__ k_outer,
                                                  const i32* __restrict__ k_cols, double* __restrict__ k_val,
                                                  const i64* __restrict__ c_outer, const i32* __restrict__ c_cols,
                                                  const double* __restrict__ c_val, double beta, int* __restrict__ bad) {
    const i64 r = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    const i64 kb = k_outer[r];
    const int kn = int(k_outer[r + 1] - kb);
    for (i64 p = c_outer[r] + lane; p < c_outer[r + 1]; p += 64) {
        const i32 c = c_cols[p];
        int lo = 0, hi = kn;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (k_cols[kb + mid] < c) lo = mid + 1; else hi = mid;
        }
        if (lo < kn && k_cols[kb + lo] == c) {
            if (WRITE) k_val[kb + lo] += beta * c_val[p];
        } else {
            *bad = 1;
        }
    }
}

__global__ void k_identity_rows(i64 rows, const uint8_t* __restrict__ inner, const i64* __restrict__ diag,
                                double* __restrict__ val) {
    const i64 r = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < rows && !inner[r]) val[diag[r]] = 1.0;
}

}  // namespace
}  // namespace nlf

using namespace nlf;

extern "C" int nlf_right_part(nlf_mesh m, int dof, const double* values_dev, double* f_dev) {
    if (!m || !values_dev || !f_dev || (dof != 1 && dof != 2)) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        hipStream_t s = m->ctx->stream;
        if (m->N == 0) return;
        if (m->nn == NLF_ELEMENT_Q4)
            hipLaunchKernelGGL((k_right_part<4, 4>), dim3(blocks_for(m->N, 256)), dim3(256), 0, s, m->N, dof, m->ne_off.p,
                               m->ne_ids.p, m->el.p, m->jac.p, m->fe, values_dev, f_dev);
        else
            hipLaunchKernelGGL((k_right_part<8, 9>), dim3(blocks_for(m->N, 256)), dim3(256), 0, s, m->N, dof, m->ne_off.p,
                               m->ne_ids.p, m->el.p, m->jac.p, m->fe, values_dev, f_dev);
        NLF_HIP(hipGetLastError());
        NLF_HIP(hipStreamSynchronize(s));
    });
}

extern "C" int nlf_matrix_combine(nlf_matrix K, double alpha, nlf_matrix C, double beta) {
    if (!K || !C) return NLF_ERR_ARGUMENT;
    return guarded(K->ctx, [&] {
        NLF_CHECK(K->rows == C->rows && K->dof == C->dof, NLF_ERR_ARGUMENT, "the two matrices have different sizes");
        NLF_CHECK(K != C, NLF_ERR_ARGUMENT, "a matrix cannot be combined with itself");
        hipStream_t s = K->ctx->stream;
        const i64 rows = K->rows;
        // containment check first, with no writes, so that K is left untouched when it fails
        DevBuf<int> bad{1};
        NLF_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
        hipLaunchKernelGGL(k_add_rows<false>, dim3(blocks_for(rows, 4)), dim3(256), 0, s, rows, K->outer.p, K->cols.p,
                           K->val.p, C->outer.p, C->cols.p, C->val.p, beta, bad.p);
        NLF_HIP(hipGetLastError());
        int h_bad = 0;
        download(&h_bad, bad.p, 1, s);
        NLF_CHECK(!h_bad, NLF_ERR_ARGUMENT, "the second matrix has entries outside the first one's pattern");
        if (K->nnz_full) {
            hipLaunchKernelGGL(k_scale, dim3(blocks_for(K->nnz_full, 256)), dim3(256), 0, s, K->nnz_full, alpha, K->val.p);
            NLF_HIP(hipGetLastError());
        }
        if (K->nnz_bound) {
            hipLaunchKernelGGL(k_scale, dim3(blocks_for(K->nnz_bound, 256)), dim3(256), 0, s, K->nnz_bound, alpha,
                               K->b_val.p);
            NLF_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_add_rows<true>, dim3(blocks_for(rows, 4)), dim3(256), 0, s, rows, K->outer.p, K->cols.p,
                           K->val.p, C->outer.p, C->cols.p, C->val.p, beta, bad.p);
        hipLaunchKernelGGL(k_identity_rows, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows, K->inner.p, K->diag.p,
                           K->val.p);
        NLF_HIP(hipGetLastError());
        NLF_HIP(hipStreamSynchronize(s));
        // derived values no longer match: drop them; the FP32 value pass cannot repeat a combination
        K->val32.release();
        K->val_lo.release();
        K->val32_assembled = false;
        ++K->val_version;
        K->split_ready = false;
        K->combined = true;
    });
}
