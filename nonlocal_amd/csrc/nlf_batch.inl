// Batched solves: many independent instances stored as the diagonal blocks of one matrix (instance
// i owns rows [row_begin[i], row_begin[i+1])), solved by lockstep preconditioned CG with
// per-instance scalars: Jacobi, or an IC(0) (ILLT) preconditioner of the batch's local matrix, whose
// factor is block diagonal like the matrix, so one level-scheduled sweep applies every instance's.  Compiled as part of nlf_solver.hip (shares block_sum / shfl helpers).
//
// Every launch covers all instances: a block never straddles two instances, partial dot products
// are written per block and summed per instance in a fixed order (bitwise reproducible), and a
// converged instance's blocks skip their work.  The SpMV groups instances by lanes-per-row class
// (one launch per class); the class of an instance comes from its nonzeros per row and a target
// number of nonzeros per lane, which the autotuner picks by timing the batched SpMV (and caches
// per batch signature).  G iterations are captured once into a hipGraph and the graph is
// replayed until every instance has converged; the captured region holds kernels only.
//
// An instance stops when its FP64 true residual ||b - A x|| / ||b|| meets the tolerance: when the
// recursive criterion sqrt(r'z)/||b|| is met inside the graph, the instance is re-judged on its
// true residual at the replay boundary and continues from the refreshed residual if needed.
// Mixed mode reads FP32 matrix values in the Krylov SpMV; vectors, dot products, the stopping
// test and x stay FP64; at a replay boundary a live instance's recursive residual is replaced by
// the FP64 true one when |r| has dropped tenfold since its last replacement (reliable update).

namespace nlf {

struct BatchState {
    double rz, pq, alpha, beta, bnorm, true_res;
    double rz_upd;             // r'z at the last residual replacement
    int done, conv;            // done: skipped by every kernel; conv: recursive criterion met
    int replace, restart;      // replace the recursive residual at this refresh; restart p = z after it
    i64 it;
};

constexpr int BATCH_VROWS = 256;       // rows per block of the vector kernels
constexpr int BATCH_GRAPH_ITERS = 16;  // CG iterations per graph replay

struct BatchClass {
    int lpr = 64;
    unsigned blocks = 0;
    DevBuf<i32> inst;          // instance of each block
    DevBuf<i64> row0;          // first row of each block
    DevBuf<i32> slot;          // partial slot of each block (instance-ordered)
};

struct BatchPlan {
    int count = 0;
    std::vector<i64> row_begin_h;
    DevBuf<i64> row_begin;
    double target = 4.0;                   // nonzeros per lane that set the lanes-per-row class
    std::vector<BatchClass> classes;
    std::vector<int> inst_lpr;
    DevBuf<i32> sp_first, sp_count;        // SpMV partial slots of each instance
    unsigned sp_slots = 0;
    DevBuf<i32> v_inst;                    // vector blocks
    DevBuf<i64> v_row0;
    DevBuf<i32> v_first, v_count;
    unsigned v_blocks = 0;
};

// ---- kernels ---------------------------------------------------------------------------------------
template<int LPR, class V, bool DOT>
__global__ __launch_bounds__(SPMV_THREADS) void k_bspmv(const i32* __restrict__ binst, const i64* __restrict__ brow0,
                                                        const i32* __restrict__ bslot, const i64* __restrict__ row_begin,
                                                        const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                        const int16_t* __restrict__ off16, const i64* __restrict__ pstart,
                                                        const V* __restrict__ val,
                                                        const double* __restrict__ x, double* __restrict__ y,
                                                        double* __restrict__ partial, const BatchState* __restrict__ st) {
    __shared__ double sh[SPMV_THREADS / 64];
    const int inst = binst[blockIdx.x];
    const bool skip = st && st[inst].done;          // uniform over the block
    const i64 r = brow0[blockIdx.x] + threadIdx.x / LPR;
    const int sub = threadIdx.x % LPR;
    double acc = 0.0;
    if (!skip && r < row_begin[inst + 1]) {
        const i64 b = outer[r], e = outer[r + 1];
        const i64 ob = pstart ? pstart[r] : 0;        // row-pattern dictionary (shared offsets)
        for (i64 p = b + sub; p < e; p += LPR) {
            const i64 c = off16 ? r + i64(off16[ob + p]) : i64(cols[p]);
            acc += double(val[p]) * x[c];
        }
    }
    for (int o = LPR / 2; o > 0; o >>= 1) acc += shfl_down_t(acc, o);
    double d = 0.0;
    if (!skip && sub == 0 && r < row_begin[inst + 1]) {
        y[r] = acc;
        if (DOT) d = x[r] * acc;
    }
    if (DOT) {
        d = block_sum<double>(d, sh);
        if (threadIdx.x == 0) partial[bslot[blockIdx.x]] = d;
    }
}

// one block per instance: s = sum of its partial slots in slot order
__device__ __forceinline__ double seg_sum(const double* __restrict__ partial, int first, int count, double* sh) {
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) s += partial[first + i];
    return block_sum<double>(s, sh);
}

__global__ __launch_bounds__(256) void k_balpha(const double* __restrict__ partial, const i32* __restrict__ first,
                                                const i32* __restrict__ cnt, BatchState* __restrict__ st) {
    __shared__ double sh[256 / 64];
    const int i = blockIdx.x;
    if (st[i].done) return;
    const double pq = seg_sum(partial, first[i], cnt[i], sh);
    if (threadIdx.x == 0) {
        st[i].pq = pq;
        const double a = st[i].rz / pq;
        st[i].alpha = (pq > 0.0 && isfinite(a)) ? a : 0.0;
        if (!(pq > 0.0) || !isfinite(a)) st[i].done = 1;
    }
}

// x += alpha p; r -= alpha q; z = invd r; partial r.z
__global__ __launch_bounds__(BATCH_VROWS) void k_bupdate(const i32* __restrict__ vinst, const i64* __restrict__ vrow0,
                                                         const i64* __restrict__ row_begin, double* __restrict__ x,
                                                         double* __restrict__ r, const double* __restrict__ p,
                                                         const double* __restrict__ q, double* __restrict__ z,
                                                         const double* __restrict__ invd, double* __restrict__ partial,
                                                         const BatchState* __restrict__ st) {
    __shared__ double sh[BATCH_VROWS / 64];
    const int inst = vinst[blockIdx.x];
    const i64 row = vrow0[blockIdx.x] + threadIdx.x;
    double d = 0.0;
    if (!st[inst].done && row < row_begin[inst + 1]) {
        const double a = st[inst].alpha;
        x[row] += a * p[row];
        const double ri = r[row] - a * q[row];
        r[row] = ri;
        if (invd) {                               // Jacobi; else z = M r and r.z follow (k_bdot)
            const double zi = invd[row] * ri;
            z[row] = zi;
            d = ri * zi;
        }
    }
    d = block_sum<double>(d, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = d;
}

// partial r.z per vector block of the live instances
__global__ __launch_bounds__(BATCH_VROWS) void k_bdot(const i32* __restrict__ vinst, const i64* __restrict__ vrow0,
                                                      const i64* __restrict__ row_begin, const double* __restrict__ r,
                                                      const double* __restrict__ z, double* __restrict__ partial,
                                                      const BatchState* __restrict__ st) {
    __shared__ double sh[BATCH_VROWS / 64];
    const int inst = vinst[blockIdx.x];
    const i64 row = vrow0[blockIdx.x] + threadIdx.x;
    double d = 0.0;
    if (!st[inst].done && row < row_begin[inst + 1]) d = r[row] * z[row];
    d = block_sum<double>(d, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = d;
}

// rt = b - q over all rows (the true residual the preconditioner is applied to at a refresh)
__global__ void k_bsub(i64 n, const double* __restrict__ b, const double* __restrict__ q, double* __restrict__ rt) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) rt[i] = b[i] - q[i];
}

// rz_new -> beta, iteration count and the reference criterion sqrt(r'z)/||b|| <= tol
__global__ __launch_bounds__(256) void k_bbeta(const double* __restrict__ partial, const i32* __restrict__ first,
                                               const i32* __restrict__ cnt, BatchState* __restrict__ st, double tol,
                                               i64 max_it) {
    __shared__ double sh[256 / 64];
    const int i = blockIdx.x;
    if (st[i].done) return;
    const double rz = seg_sum(partial, first[i], cnt[i], sh);
    if (threadIdx.x == 0) {
        st[i].beta = st[i].rz != 0.0 ? rz / st[i].rz : 0.0;
        st[i].rz = rz;
        st[i].it += 1;
        const double res = sqrt(fabs(rz)) / st[i].bnorm;
        if (!isfinite(res) || res <= tol || st[i].it >= max_it) {
            st[i].conv = 1;
            st[i].done = 1;
        }
    }
}

__global__ __launch_bounds__(BATCH_VROWS) void k_bp(const i32* __restrict__ vinst, const i64* __restrict__ vrow0,
                                                    const i64* __restrict__ row_begin, double* __restrict__ p,
                                                    const double* __restrict__ z, const BatchState* __restrict__ st) {
    const int inst = vinst[blockIdx.x];
    const i64 row = vrow0[blockIdx.x] + threadIdx.x;
    if (!st[inst].done && row < row_begin[inst + 1]) p[row] = z[row] + st[inst].beta * p[row];
}

// residual refresh: r_true = b - q (q = A x), z = invd r_true (or zt), partials r_true.z, b.b,
// r_true.r_true, |r_true - r|^2 and |r|^2 (deviation of the recursive residual r).  r and z are
// overwritten for instances being replaced (all of them at the start, where p = z as well).
__global__ __launch_bounds__(BATCH_VROWS) void k_bresidual(const i32* __restrict__ vinst, const i64* __restrict__ vrow0,
                                                           const i64* __restrict__ row_begin, const double* __restrict__ b,
                                                           const double* __restrict__ q, double* __restrict__ r,
                                                           double* __restrict__ z, double* __restrict__ p,
                                                           const double* __restrict__ invd, const double* __restrict__ zt,
                                                           double* __restrict__ prz, double* __restrict__ pbb,
                                                           double* __restrict__ prr, double* __restrict__ pdev,
                                                           double* __restrict__ prec, const BatchState* __restrict__ st,
                                                           int start) {
    __shared__ double sh[BATCH_VROWS / 64];
    const int inst = vinst[blockIdx.x];
    const i64 row = vrow0[blockIdx.x] + threadIdx.x;
    double a = 0.0, c = 0.0, e = 0.0, dv = 0.0, rc = 0.0;
    if ((start || !st[inst].done) && row < row_begin[inst + 1]) {
        const double ri = b[row] - q[row];
        const double zi = zt ? zt[row] : invd[row] * ri;       // zt = M (b - q) when M is not Jacobi
        if (!start) {
            const double ro = r[row];
            dv = (ri - ro) * (ri - ro);
            rc = ro * ro;
        }
        if (start || st[inst].replace) {
            r[row] = ri;
            z[row] = zi;
            if (start) p[row] = zi;
        }
        a = ri * zi;
        c = b[row] * b[row];
        e = ri * ri;
    }
    a = block_sum<double>(a, sh);
    c = block_sum<double>(c, sh);
    e = block_sum<double>(e, sh);
    dv = block_sum<double>(dv, sh);
    rc = block_sum<double>(rc, sh);
    if (threadIdx.x == 0) {
        prz[blockIdx.x] = a; pbb[blockIdx.x] = c; prr[blockIdx.x] = e; pdev[blockIdx.x] = dv; prec[blockIdx.x] = rc;
    }
}

// p = z for the instances restarted by the last refresh
__global__ __launch_bounds__(BATCH_VROWS) void k_brestart(const i32* __restrict__ vinst, const i64* __restrict__ vrow0,
                                                          const i64* __restrict__ row_begin, double* __restrict__ p,
                                                          const double* __restrict__ z, const BatchState* __restrict__ st) {
    const int inst = vinst[blockIdx.x];
    const i64 row = vrow0[blockIdx.x] + threadIdx.x;
    if (!st[inst].done && st[inst].restart && row < row_begin[inst + 1]) p[row] = z[row];
}

// after a residual refresh: rz, bnorm (start), true residual, done.  A replaced instance keeps its
// direction p (reliable update) while the recursive residual deviated < 10 % from the true one;
// otherwise (or after meeting the recursive criterion) CG restarts from p = z.
__global__ __launch_bounds__(256) void k_brefresh(const double* __restrict__ prz, const double* __restrict__ pbb,
                                                  const double* __restrict__ prr, const double* __restrict__ pdev,
                                                  const double* __restrict__ prec, const i32* __restrict__ first,
                                                  const i32* __restrict__ cnt, BatchState* __restrict__ st, double tol,
                                                  i64 max_it, int start) {
    __shared__ double sh[256 / 64];
    const int i = blockIdx.x;
    if (!start && st[i].done) return;                       // broke down: leave it
    const double rz = seg_sum(prz, first[i], cnt[i], sh);
    const double bb = seg_sum(pbb, first[i], cnt[i], sh);
    const double rr = seg_sum(prr, first[i], cnt[i], sh);
    const double dv = seg_sum(pdev, first[i], cnt[i], sh);
    const double rc = seg_sum(prec, first[i], cnt[i], sh);
    if (threadIdx.x == 0) {
        if (start) {
            st[i].bnorm = bb > 0.0 ? sqrt(bb) : 1.0;
            st[i].it = 0;
            st[i].beta = 0.0;
        }
        st[i].restart = 0;
        if (start || st[i].replace) {
            st[i].rz = rz;
            st[i].rz_upd = rz;
            st[i].restart = (!start && (st[i].conv || dv > 0.01 * rc)) ? 1 : 0;
        }
        st[i].true_res = sqrt(rr) / st[i].bnorm;
        const bool met = st[i].true_res <= tol || st[i].it >= max_it || !isfinite(st[i].true_res);
        st[i].done = met ? 1 : 0;
        st[i].conv = 0;
        st[i].replace = 0;
    }
}

__global__ void k_bfill(i64 n, double* __restrict__ x, double v) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) x[i] = v;
}

// Before a refresh: instances that met the recursive criterion are reopened for the FP64
// true-residual check and get their residual replaced (with a restart); in mixed mode a live
// instance also gets it replaced once r'z has dropped 100x (|r| tenfold) since its last replacement.
__global__ void k_breopen(BatchState* __restrict__ st, int count, int mixed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    if (st[i].conv) {
        st[i].done = 0;
        st[i].replace = 1;
    } else {
        st[i].replace = (mixed && !st[i].done && st[i].rz < 0.01 * st[i].rz_upd) ? 1 : 0;
    }
}

__global__ void k_bstate_init(BatchState* __restrict__ st, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    BatchState s{};
    st[i] = s;
}

// ---- host side ---------------------------------------------------------------------------------------
int lpr_for(double nnz_per_row, double target) {
    int l = 4;
    while (l < 64 && double(l) * target < nnz_per_row) l *= 2;
    return l;
}

void batch_build_plan(BatchPlan& P, nlf_matrix A, double target, hipStream_t s) {
    P.target = target;
    const int count = P.count;
    std::vector<i64> outer_b(size_t(count) + 1);
    // nonzeros per instance from the row pointers at the instance boundaries
    for (int i = 0; i <= count; ++i) {
        i64 v = 0;
        download(&v, A->outer.p + P.row_begin_h[size_t(i)], 1, s);
        outer_b[size_t(i)] = v;
    }
    P.inst_lpr.assign(size_t(count), 4);
    for (int i = 0; i < count; ++i) {
        const i64 rows = P.row_begin_h[size_t(i) + 1] - P.row_begin_h[size_t(i)];
        const double avg = rows > 0 ? double(outer_b[size_t(i) + 1] - outer_b[size_t(i)]) / double(rows) : 0.0;
        P.inst_lpr[size_t(i)] = lpr_for(avg, target);
    }
    // SpMV blocks in instance order (slots), then split by class
    std::vector<i32> sp_first(static_cast<size_t>(count)), sp_count(static_cast<size_t>(count));
    std::map<int, std::vector<std::array<i64, 3>>> by_class;       // lpr -> (inst, row0, slot)
    i64 slot = 0;
    for (int i = 0; i < count; ++i) {
        const int lpr = P.inst_lpr[size_t(i)], rpb = SPMV_THREADS / lpr;
        const i64 r0 = P.row_begin_h[size_t(i)], r1 = P.row_begin_h[size_t(i) + 1];
        sp_first[size_t(i)] = i32(slot);
        i32 c = 0;
        for (i64 r = r0; r < r1; r += rpb, ++c, ++slot) by_class[lpr].push_back({i64(i), r, slot});
        sp_count[size_t(i)] = c;
    }
    P.sp_slots = unsigned(std::max<i64>(slot, 1));
    P.classes.clear();
    for (auto& kv : by_class) {
        BatchClass c;
        c.lpr = kv.first;
        c.blocks = unsigned(kv.second.size());
        std::vector<i32> inst(kv.second.size()), sl(kv.second.size());
        std::vector<i64> row0(kv.second.size());
        for (size_t k = 0; k < kv.second.size(); ++k) {
            inst[k] = i32(kv.second[k][0]);
            row0[k] = kv.second[k][1];
            sl[k] = i32(kv.second[k][2]);
        }
        upload(c.inst, inst.data(), inst.size(), s);
        upload(c.row0, row0.data(), row0.size(), s);
        upload(c.slot, sl.data(), sl.size(), s);
        P.classes.push_back(std::move(c));
    }
    upload(P.sp_first, sp_first.data(), sp_first.size(), s);
    upload(P.sp_count, sp_count.data(), sp_count.size(), s);
    // vector blocks
    std::vector<i32> vinst, v_first(static_cast<size_t>(count)), v_count(static_cast<size_t>(count));
    std::vector<i64> vrow0;
    for (int i = 0; i < count; ++i) {
        const i64 r0 = P.row_begin_h[size_t(i)], r1 = P.row_begin_h[size_t(i) + 1];
        v_first[size_t(i)] = i32(vinst.size());
        for (i64 r = r0; r < r1; r += BATCH_VROWS) { vinst.push_back(i32(i)); vrow0.push_back(r); }
        v_count[size_t(i)] = i32(vinst.size()) - v_first[size_t(i)];
    }
    P.v_blocks = unsigned(vinst.size());
    upload(P.v_inst, vinst.data(), vinst.size(), s);
    upload(P.v_row0, vrow0.data(), vrow0.size(), s);
    upload(P.v_first, v_first.data(), v_first.size(), s);
    upload(P.v_count, v_count.data(), v_count.size(), s);
}

template<class V, bool DOT>
void batch_spmv(const BatchPlan& P, nlf_matrix A, const V* val, const double* x, double* y, double* partial,
                const BatchState* st, hipStream_t s) {
    const SpmvCols cs = spmv_cols(A);
    for (const auto& c : P.classes) {
        if (!c.blocks) continue;
#define NLF_BSPMV_CASE(L)                                                                                            \
    case L:                                                                                                          \
        hipLaunchKernelGGL((k_bspmv<L, V, DOT>), dim3(c.blocks), dim3(SPMV_THREADS), 0, s, c.inst.p, c.row0.p, c.slot.p, \
                           P.row_begin.p, A->outer.p, cs.cols, cs.off, cs.pstart, val, x, y, partial, st);            \
        break;
        switch (c.lpr) {
            NLF_BSPMV_CASE(4)
            NLF_BSPMV_CASE(8)
            NLF_BSPMV_CASE(16)
            NLF_BSPMV_CASE(32)
            default: NLF_BSPMV_CASE(64)
        }
#undef NLF_BSPMV_CASE
    }
}

}  // namespace nlf

struct nlf_batch_s {
    nlf_ctx ctx = nullptr;
    nlf_matrix A = nullptr;
    nlf::BatchPlan plan;
    nlf::DevBuf<double> invd;
    nlf::DevBuf<float> invd32;
    double tune_seconds = 0.0;
    nlf_precond precond = nullptr;       // ILLT of the batch's local matrix (nlf_batch_set_preconditioner), or Jacobi
};

namespace nlf {

// The lockstep CG; returns the number of graph replays.
template<class V>
i64 batch_cg(nlf_batch B, const V* val, const double* b, double* x, double tol, i64 max_it, bool use_graph,
             std::vector<BatchState>& out) {
    BatchPlan& P = B->plan;
    nlf_matrix A = B->A;
    hipStream_t s = B->ctx->stream;
    const i64 n = A->rows;
    const int count = P.count;
    DevBuf<double> r{size_t(n)}, z{size_t(n)}, p{size_t(n)}, q{size_t(n)};
    const size_t np = std::max<size_t>(P.sp_slots, P.v_blocks);
    DevBuf<double> pa{np}, pb{np}, pc{np}, pd{np}, pe{np};
    DevBuf<BatchState> st{size_t(count)};
    const nlf_precond_s* M = B->precond && B->precond->kind == NLF_PRECOND_ILLT ? B->precond : nullptr;
    DevBuf<double> rt, zt;
    if (M) { rt.alloc(size_t(n)); zt.alloc(size_t(n)); }
    hipLaunchKernelGGL(k_bstate_init, dim3(blocks_for(count, 64)), dim3(64), 0, s, st.p, count);
    auto refresh = [&](int start) {
        // FP64 true residual of every live instance (and the start values): q = A x with FP64 values
        batch_spmv<double, false>(P, A, A->val.p, x, q.p, nullptr, start ? nullptr : st.p, s);
        if (M) {
            hipLaunchKernelGGL(k_bsub, dim3(grid_stride_blocks(n)), dim3(256), 0, s, n, b, q.p, rt.p);
            precond_apply<double>(M, n, rt.p, zt.p, nullptr, s);
        }
        hipLaunchKernelGGL(k_bresidual, dim3(P.v_blocks), dim3(BATCH_VROWS), 0, s, P.v_inst.p, P.v_row0.p, P.row_begin.p, b,
                           q.p, r.p, z.p, p.p, B->invd.p, M ? (const double*)zt.p : nullptr, pa.p, pb.p, pc.p, pd.p, pe.p,
                           (const BatchState*)st.p, start);
        hipLaunchKernelGGL(k_brefresh, dim3(count), dim3(256), 0, s, pa.p, pb.p, pc.p, pd.p, pe.p, P.v_first.p, P.v_count.p,
                           st.p, tol, max_it, start);
        if (!start)
            hipLaunchKernelGGL(k_brestart, dim3(P.v_blocks), dim3(BATCH_VROWS), 0, s, P.v_inst.p, P.v_row0.p, P.row_begin.p,
                               p.p, z.p, (const BatchState*)st.p);
    };
    auto iteration = [&]() {
        batch_spmv<V, true>(P, A, val, p.p, q.p, pa.p, st.p, s);
        hipLaunchKernelGGL(k_balpha, dim3(count), dim3(256), 0, s, pa.p, P.sp_first.p, P.sp_count.p, st.p);
        hipLaunchKernelGGL(k_bupdate, dim3(P.v_blocks), dim3(BATCH_VROWS), 0, s, P.v_inst.p, P.v_row0.p, P.row_begin.p, x,
                           r.p, p.p, q.p, z.p, M ? (const double*)nullptr : B->invd.p, pb.p, st.p);
        if (M) {
            precond_apply<double>(M, n, r.p, z.p, nullptr, s);
            hipLaunchKernelGGL(k_bdot, dim3(P.v_blocks), dim3(BATCH_VROWS), 0, s, P.v_inst.p, P.v_row0.p, P.row_begin.p, r.p,
                               z.p, pb.p, (const BatchState*)st.p);
        }
        hipLaunchKernelGGL(k_bbeta, dim3(count), dim3(256), 0, s, pb.p, P.v_first.p, P.v_count.p, st.p, tol, max_it);
        hipLaunchKernelGGL(k_bp, dim3(P.v_blocks), dim3(BATCH_VROWS), 0, s, P.v_inst.p, P.v_row0.p, P.row_begin.p, p.p, z.p,
                           st.p);
    };
    refresh(1);
    NLF_HIP(hipGetLastError());
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
        NLF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < BATCH_GRAPH_ITERS; ++k) iteration();
        NLF_HIP(hipStreamEndCapture(s, &graph));
        NLF_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    std::vector<BatchState> h(static_cast<size_t>(count));
    i64 replays = 0;
    const i64 max_replays = max_it / BATCH_GRAPH_ITERS + 2;
    try {
        for (;;) {
            if (use_graph) NLF_HIP(hipGraphLaunch(exec, s));
            else for (int k = 0; k < BATCH_GRAPH_ITERS; ++k) iteration();
            ++replays;
            download(h.data(), st.p, size_t(count), s);
            bool any_conv = false, all_done = true;
            for (const auto& e : h) {
                any_conv = any_conv || e.conv;
                all_done = all_done && e.done;
            }
            if (any_conv || std::is_same<V, float>::value) {
                // true residual: converged ones are re-judged, mixed ones reliably updated
                hipLaunchKernelGGL(k_breopen, dim3(blocks_for(count, 64)), dim3(64), 0, s, st.p, count,
                                   int(std::is_same<V, float>::value));
                refresh(0);
                download(h.data(), st.p, size_t(count), s);
                all_done = true;
                for (const auto& e : h) all_done = all_done && e.done;
            }
            if (all_done || replays >= max_replays) break;
        }
    } catch (...) {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    if (exec) NLF_HIP(hipGraphExecDestroy(exec));
    if (graph) NLF_HIP(hipGraphDestroy(graph));
    out = h;
    return replays;
}

}  // namespace nlf

extern "C" {

// A batch over the diagonal blocks of `a`: instance i owns rows [row_begin[i], row_begin[i+1]).
int nlf_batch_create(nlf_matrix a, int count, const int64_t* row_begin, nlf_batch* out) {
    if (!a || !row_begin || !out || count <= 0) return NLF_ERR_ARGUMENT;
    return guarded(a->ctx, [&] {
        NLF_CHECK(row_begin[0] == 0 && row_begin[count] == a->rows, NLF_ERR_ARGUMENT, "row ranges must cover the matrix");
        NLF_CHECK(!a->neumann, NLF_ERR_UNSUPPORTED, "batched solves of bordered Neumann systems are not implemented");
        for (int i = 0; i < count; ++i) NLF_CHECK(row_begin[i] <= row_begin[i + 1], NLF_ERR_ARGUMENT, "row ranges not ascending");
        hipStream_t s = a->ctx->stream;
        auto B = std::make_unique<nlf_batch_s>();
        B->ctx = a->ctx;
        B->A = a;
        B->plan.count = count;
        B->plan.row_begin_h.assign(row_begin, row_begin + count + 1);
        upload(B->plan.row_begin, row_begin, size_t(count) + 1, s);
        B->invd.alloc(size_t(a->rows));
        B->invd32.alloc(size_t(a->rows));
        hipLaunchKernelGGL(k_inv_diag, dim3(blocks_for(a->rows, 256)), dim3(256), 0, s, a->rows, a->diag.p, a->val.p,
                           B->invd.p, B->invd32.p);
        batch_build_plan(B->plan, a, 4.0, s);
        NLF_HIP(hipGetLastError());
        NLF_HIP(hipStreamSynchronize(s));
        *out = B.release();
    });
}

int nlf_batch_destroy(nlf_batch b) {
    if (!b) return NLF_OK;
    CtxScope scope(live_ctx(b->ctx));
    delete b;
    return NLF_OK;
}

// Autotuner: times the batched SpMV for each target of nonzeros per lane (which sets every
// instance's lanes-per-row class and rows-per-block tile) and keeps the fastest.  With a cache
// file, a batch whose signature (instance sizes and nonzeros) is listed there reuses the stored
// choice; a new measurement is appended.  chosen (optional) receives the target.
int nlf_batch_autotune(nlf_batch b, const char* cache_path, double* chosen) {
    if (!b) return NLF_ERR_ARGUMENT;
    return guarded(b->ctx, [&] {
        BatchPlan& P = b->plan;
        nlf_matrix A = b->A;
        hipStream_t s = b->ctx->stream;
        // signature: FNV-1a over row ranges and the nonzero count
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](uint64_t v) { for (int k = 0; k < 8; ++k) { h ^= (v >> (8 * k)) & 0xff; h *= 1099511628211ull; } };
        for (i64 v : P.row_begin_h) mix(uint64_t(v));
        mix(uint64_t(A->nnz_full));
        char key[32];
        std::snprintf(key, sizeof key, "%016llx", (unsigned long long)h);
        double best_t = 0.0;
        bool found = false;
        if (cache_path) {
            if (FILE* f = std::fopen(cache_path, "r")) {
                char k[64];
                double t;
                while (std::fscanf(f, "%63s %lf", k, &t) == 2)
                    if (std::strcmp(k, key) == 0) { best_t = t; found = true; }
                std::fclose(f);
            }
        }
        const auto t0 = std::chrono::steady_clock::now();
        if (!found) {
            DevBuf<double> x{size_t(A->rows)}, y{size_t(A->rows)};
            hipLaunchKernelGGL(k_bfill, dim3(grid_stride_blocks(A->rows)), dim3(256), 0, s, A->rows, x.p, 1.0);
            hipEvent_t e0, e1;
            NLF_HIP(hipEventCreate(&e0));
            NLF_HIP(hipEventCreate(&e1));
            double best_ms = 1e300;
            for (double t : {1.0, 2.0, 4.0, 8.0, 16.0}) {
                batch_build_plan(P, A, t, s);
                batch_spmv<double, false>(P, A, A->val.p, x.p, y.p, nullptr, nullptr, s);   // warm-up
                NLF_HIP(hipEventRecord(e0, s));
                for (int k = 0; k < 5; ++k) batch_spmv<double, false>(P, A, A->val.p, x.p, y.p, nullptr, nullptr, s);
                NLF_HIP(hipEventRecord(e1, s));
                NLF_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                NLF_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best_ms) { best_ms = ms; best_t = t; }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (cache_path)
                if (FILE* f = std::fopen(cache_path, "a")) {
                    std::fprintf(f, "%s %g\n", key, best_t);
                    std::fclose(f);
                }
        }
        batch_build_plan(P, A, best_t, s);
        NLF_HIP(hipStreamSynchronize(s));
        b->tune_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (chosen) *chosen = best_t;
    });
}

// No reference counterpart.  The preconditioner of the batch's solves: an ILLT built on the batch's local
// matrix (block diagonal like the batch; same rows), or NULL for Jacobi of the matrix (the default).  An
// ILLT of a block-diagonal matrix factors and sweeps every instance on a workgroup of its own.
int nlf_batch_set_preconditioner(nlf_batch bt, nlf_precond p) {
    if (!bt) return NLF_ERR_ARGUMENT;
    return guarded(bt->ctx, [&] {
        NLF_CHECK(!p || p->kind == NLF_PRECOND_ILLT || p->kind == NLF_PRECOND_JACOBI, NLF_ERR_UNSUPPORTED,
                  "batched solves take Jacobi or ILLT");
        NLF_CHECK(!p || p->rows == bt->A->rows, NLF_ERR_ARGUMENT, "the preconditioner has another size than the batch");
        NLF_CHECK(!p || p->ok, NLF_ERR_NUMERIC, "the preconditioner's factorisation failed");
        bt->precond = p && p->kind == NLF_PRECOND_ILLT ? p : nullptr;
    });
}

// Lockstep preconditioned CG of every instance.  x holds the initial guesses.  mode NLF_MODE_MIXED reads
// FP32 values in the Krylov SpMV (nlf_matrix_make_fp32 first) with FP64 reliable updates.
// use_graph: capture the iterations into a hipGraph (1) or launch them eagerly (0).  iterations[i]
// and true_residual[i] per instance (optional); stats->iterations = the largest count.
int nlf_batch_solve(nlf_batch bt, const double* b, double* x, double tol, int64_t max_it, int mode, int use_graph,
                    int64_t* iterations, double* true_residual, nlf_solve_stats* stats) {
    if (!bt || !b || !x) return NLF_ERR_ARGUMENT;
    return guarded(bt->ctx, [&] {
        const auto t0 = std::chrono::steady_clock::now();
        NLF_CHECK(mode == NLF_MODE_FP64 || mode == NLF_MODE_MIXED, NLF_ERR_UNSUPPORTED,
                  "batched solves run in NLF_MODE_FP64 or NLF_MODE_MIXED");
        std::vector<BatchState> h;
        i64 replays = 0;
        if (mode == NLF_MODE_MIXED) {
            NLF_CHECK(bt->A->val32.p, NLF_ERR_STATE, "mixed mode needs nlf_matrix_make_fp32 first");
            replays = batch_cg<float>(bt, bt->A->val32.p, b, x, tol, max_it, use_graph != 0, h);
        } else {
            replays = batch_cg<double>(bt, bt->A->val.p, b, x, tol, max_it, use_graph != 0, h);
        }
        nlf_solve_stats out{};
        out.preconditioner = bt->precond && bt->precond->kind == NLF_PRECOND_ILLT ? NLF_PRECOND_ILLT : NLF_PRECOND_JACOBI;
        out.outer_iterations = replays;
        for (size_t i = 0; i < h.size(); ++i) {
            if (iterations) iterations[i] = h[i].it;
            if (true_residual) true_residual[i] = h[i].true_res;
            out.iterations = std::max<int64_t>(out.iterations, h[i].it);
            out.true_residual = std::max(out.true_residual, h[i].true_res);
        }
        out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = out;
    });
}

}  // extern "C"
