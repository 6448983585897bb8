// Learned preconditioner: a message-passing network over mesh nodes, its training kernels, and
// flexible CG.  Compiled as part of nlf_solver.hip (shares its Krylov kernels).
//
// Graph: the nodes of the thermal problem; edges = the pattern of the LOCAL matrix (element
// adjacency, self loops included).  Edge features (symmetric in u, v):
//   f0 = K_L(u,v) s_u s_v, f1 = K(u,v) s_u s_v, f2 = (1-p1) phi(x_u, x_v) / phi(0), f3 = [u == v],
// with s = 1/sqrt(diag K).  Network N (H = 32 channels, three message-passing layers, FP32):
//   x_u = s_u r_u / sigma,  sigma = rms(S r)
//   h0  = a_in x_u
//   m_u = sum_{v in N(u)} (W_e f_uv) (.) h_v ;  h' = W_s h_u + W_m m_u           (x3)
//   y_u = w_out . h3_u ;  N(r)_u = s_u sigma (x_u + y_u) = s_u^2 r_u + s_u sigma y_u  (first term in FP64)
// The message path is linear in the residual (edge gates W_e f carry the mesh, the influence
// function and the coefficients), so N = S^2 + S Y S is a fixed linear operator with a three-hop
// stencil and flexible CG keeps the convergence of PCG.  The preconditioner applied in the Krylov
// loop is its symmetric part  M r = S^2 r + S (Y + Y^T)/2 S r : Y^T runs the layers backwards
// (transposed weights, same gates over the symmetric graph).  With zero output weights M is Jacobi.
// Training minimises ||N(r) - e||_A^2 / ||e||_A^2 over (residual, error) pairs (A e = r).
// Gradients are reduced per block in a fixed order (deterministic) and applied with Adam.
// Parameter layout keeps slots for biases (P_BIN, P_B, P_BOUT); they are held at zero so that the
// operator stays linear.

namespace nlf {

constexpr int MP_H = 32, MP_F = 4, MP_L = 3;
constexpr int P_AIN = 0, P_BIN = MP_H, P_LAYER0 = 2 * MP_H;
constexpr int P_WE = 0, P_WS = MP_H * MP_F, P_WM = P_WS + MP_H * MP_H, P_B = P_WM + MP_H * MP_H;
constexpr int P_LAYER_SIZE = P_B + MP_H;
constexpr int P_WOUT = P_LAYER0 + MP_L * P_LAYER_SIZE, P_BOUT = P_WOUT + MP_H;
constexpr int NPARAM = P_BOUT + 1;
constexpr int MP_THREADS = 256;                 // 8 nodes x 32 channels per block
constexpr int MP_NODES = MP_THREADS / MP_H;
constexpr int MP_GRAD_BLOCKS = 256;             // blocks of the gradient kernels (partial rows)

struct LearnedGraph {
    i64 n = 0, nnz = 0;
    DevBuf<i64> outer;
    DevBuf<i32> cols;
    DevBuf<float> feat;                         // nnz x 4
    DevBuf<float> s;                            // 1/sqrt(K_uu)
    DevBuf<float> w;                            // NPARAM
    DevBuf<float> x;                            // scaled residual
    DevBuf<float> y;                            // forward output y = w_out . h3
    DevBuf<float> h[MP_L + 1], m[MP_L];         // training: all activations; inference: h[0], h[1]
    DevBuf<float> d;                            // transposed pass: own-node part (in place)
    int ell_d = 0;                              // inference graph: ELL, k-major, D = longest row
    DevBuf<i32> ell_col;
    DevBuf<float4> ell_f;
    DevBuf<float> v;                            // 64 w_out-contracted weights of the last layer
    i64 band = 0;                               // max |col - row| of the local pattern
    bool m_ready = false;                       // compiled operator M (three-hop pattern, FP32 values)
    i64 m_nnz = 0;
    DevBuf<i64> m_outer;
    DevBuf<i32> m_cols;
    DevBuf<int16_t> m_off16;
    DevBuf<int16_t> m_pat;                      // row-pattern dictionary of M (when it pays)
    DevBuf<i64> m_pstart;
    i64 m_pat_size = 0, m_pat_groups = 0;
    DevBuf<float> m_val;
    double compile_seconds = 0.0;
    bool network_apply = false;                 // run the network kernels even when M is compiled
    bool jacobi_apply = false;                  // apply only the network's fixed term s^2 r (Jacobi)
    std::unique_ptr<CoarseSpace> coarse;        // two-level form: + P A_c^-1 P^T r (nlf_learned_set_coarse)
    DevBuf<float> g0, g1, gdir, gm;             // backward work
    DevBuf<double> part, sigma, zbuf, az;
    DevBuf<float> gpart, grad;
    bool train_ready = false;
};

// ---- feature construction ----------------------------------------------------------------------
__global__ void k_lf_scale(i64 n, const i64* __restrict__ kdiag, const double* __restrict__ kval, float* __restrict__ s) {
    const i64 u = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const double d = kval[kdiag[u]];
    s[u] = d > 0.0 ? float(1.0 / sqrt(d)) : 1.0f;
}

// group of a node = group of the lowest-numbered element containing it (0 for a node in no element)
__device__ __forceinline__ int node_group(i64 u, const i32* __restrict__ ne_off, const i32* __restrict__ ne_ids,
                                          const i32* __restrict__ group_of) {
    return ne_off[u] < ne_off[u + 1] ? group_of[ne_ids[ne_off[u]]] : 0;
}

__global__ void k_lf_features(i64 n, const i64* __restrict__ lo, const i32* __restrict__ lc, const double* __restrict__ lv,
                              const i64* __restrict__ ko, const i32* __restrict__ kc, const double* __restrict__ kv,
                              const float* __restrict__ s, const double* __restrict__ xy, const i32* __restrict__ ne_off,
                              const i32* __restrict__ ne_ids, const i32* __restrict__ group_of,
                              const GroupParams* __restrict__ groups, float* __restrict__ feat) {
    const i64 u = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const GroupParams gu = groups[node_group(u, ne_off, ne_ids, group_of)];
    const double xu = xy[2 * u], yu = xy[2 * u + 1];
    // phi(x, x) is the influence's norm for every kind and distance (the rotated distance is NaN at the
    // origin, so it is not evaluated); a NaN weight there counts as no influence
    const double phi0u = gu.nonlocal ? gu.inf.norm : 1.0;
    for (i64 p = lo[u]; p < lo[u + 1]; ++p) {
        const i32 v = lc[p];
        // K(u, v) by binary search in K's row u
        i64 b = ko[u], e = ko[u + 1];
        while (b < e) {
            const i64 mid = (b + e) >> 1;
            if (kc[mid] < v) b = mid + 1; else e = mid;
        }
        const double kuv = (b < ko[u + 1] && kc[b] == v) ? kv[b] : 0.0;
        const GroupParams gv = groups[node_group(v, ne_off, ne_ids, group_of)];
        const double xv = xy[2 * v], yv = xy[2 * v + 1];
        double wu = gu.nonlocal ? (1.0 - gu.p1) * influence(gu.inf, xu, yu, xv, yv) / phi0u : 0.0;
        const double phi0v = gv.nonlocal ? gv.inf.norm : 1.0;
        double wv = gv.nonlocal ? (1.0 - gv.p1) * influence(gv.inf, xv, yv, xu, yu) / phi0v : 0.0;
        wu = isfinite(wu) ? wu : 0.0;
        wv = isfinite(wv) ? wv : 0.0;
        const double ss = double(s[u]) * double(s[v]);
        feat[4 * p + 0] = float(lv[p] * ss);
        feat[4 * p + 1] = float(kuv * ss);
        feat[4 * p + 2] = float(0.5 * (wu + wv));
        feat[4 * p + 3] = u == v ? 1.0f : 0.0f;
    }
}

// ---- forward -------------------------------------------------------------------------------------
// partial[b] = sum (s_u r_u)^2 over the block's rows
template<class R>
__global__ __launch_bounds__(256) void k_lf_rnorm(i64 n, const R* __restrict__ r, const float* __restrict__ s,
                                                  double* __restrict__ partial) {
    __shared__ double sh[256 / 64];
    double acc = 0.0;
    for (i64 u = i64(blockIdx.x) * blockDim.x + threadIdx.x; u < n; u += i64(gridDim.x) * blockDim.x) {
        const double v = double(s[u]) * double(r[u]);
        acc += v * v;
    }
    acc = block_sum<double>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void k_lf_sigma(const double* __restrict__ partial, int nb, i64 n, double* __restrict__ sigma) {
    __shared__ double sh[1024 / 64];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partial[i];
    acc = block_sum<double>(acc, sh);
    if (threadIdx.x == 0) {
        const double sg = sqrt(acc / double(n > 0 ? n : 1));
        *sigma = (sg > 0.0 && isfinite(sg)) ? sg : 1.0;
    }
}

template<class R>
__global__ __launch_bounds__(MP_THREADS) void k_lf_input(i64 n, const R* __restrict__ r, const float* __restrict__ s,
                                                         const double* __restrict__ sigma, const float* __restrict__ w,
                                                         float* __restrict__ x, float* __restrict__ h0) {
    const int c = threadIdx.x % MP_H;
    const float inv = float(1.0 / *sigma);
    const float a = w[P_AIN + c];
    for (i64 u = i64(blockIdx.x) * MP_NODES + threadIdx.x / MP_H; u < n; u += i64(gridDim.x) * MP_NODES) {
        const float xu = float(double(s[u]) * double(r[u])) * inv;
        h0[u * MP_H + c] = a * xu;
        if (c == 0) x[u] = xu;
    }
}

// ---- message-passing layers ----------------------------------------------------------------------
// Block = 256 threads = 8 slots x 32 channels; each slot carries MP_NPL nodes per iteration, so
// every lane keeps MP_NPL independent gather streams in flight (the gathers of one node form a
// dependent chain: row offsets -> columns -> neighbour vectors).  Edge columns of the 8*MP_NPL
// nodes are staged in LDS (local pattern rows are at most MP_MAXDEG long); edge features are read
// directly (one 16-byte broadcast per edge and node).  The 32 x 32 products read the weights and
// the node vectors from LDS as float4, each weight float4 serving MP_NPL nodes.
constexpr int MP_MAXDEG = 32;
constexpr int MP_NPL = 4;                       // nodes per lane and iteration
constexpr int MP_TILE = MP_NODES * MP_NPL;      // nodes per block iteration
constexpr int WPAD = MP_H + 4;                  // LDS row stride of a weight matrix (float4 rows)

struct __align__(16) LayerLds {
    float A[MP_H][WPAD], B[MP_H][WPAD];         // two 32 x 32 matrices (as stored or transposed)
    float va[MP_TILE][MP_H], vb[MP_TILE][MP_H]; // node vectors for the products
    int ecol[MP_TILE][MP_MAXDEG];               // staged edge columns
};

// W row-major [out][in] from global into LDS, as stored (dst[o][i]) or transposed (dst[i][o]).
template<bool T>
__device__ __forceinline__ void lds_load_w(float (*dst)[WPAD], const float* __restrict__ src) {
    for (int i = threadIdx.x; i < MP_H * MP_H; i += blockDim.x) {
        const int o = i / MP_H, k = i % MP_H;
        if (T) dst[k][o] = src[i]; else dst[o][k] = src[i];
    }
}

// sum_k W[c][k] v[k] (v broadcast from LDS)
__device__ __forceinline__ float lds_dot(const float (*W)[WPAD], const float* v, int c) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < MP_H; k += 4) {
        const float4 w = *reinterpret_cast<const float4*>(&W[c][k]);
        const float4 x = *reinterpret_cast<const float4*>(&v[k]);
        acc += w.x * x.x + w.y * x.y + w.z * x.z + w.w * x.w;
    }
    return acc;
}

// out[j] = sum_k W[c][k] v[slot + 8 j][k] for the MP_NPL nodes of a slot; accumulates into out
__device__ __forceinline__ void lds_dot_tile(const float (*W)[WPAD], const float (*v)[MP_H], int slot, int c,
                                             float (&out)[MP_NPL]) {
#pragma unroll 2
    for (int k = 0; k < MP_H; k += 4) {
        const float4 w = *reinterpret_cast<const float4*>(&W[c][k]);
#pragma unroll
        for (int j = 0; j < MP_NPL; ++j) {
            const float4 x = *reinterpret_cast<const float4*>(&v[slot + MP_NODES * j][k]);
            out[j] += w.x * x.x + w.y * x.y + w.z * x.z + w.w * x.w;
        }
    }
}

// Row offsets of the slot's nodes and their columns into LDS (lane c stages edge c of each node).
__device__ __forceinline__ void stage_tile(LayerLds& L, i64 u0, int slot, int c, i64 n, const i64* __restrict__ outer,
                                           const i32* __restrict__ cols, i64 (&p0)[MP_NPL], int (&deg)[MP_NPL]) {
#pragma unroll
    for (int j = 0; j < MP_NPL; ++j) {
        const i64 u = u0 + slot + MP_NODES * j;
        p0[j] = 0;
        deg[j] = 0;
        if (u < n) {
            p0[j] = outer[u];
            deg[j] = int(outer[u + 1] - p0[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < MP_NPL; ++j)
        if (c < deg[j]) L.ecol[slot + MP_NODES * j][c] = cols[p0[j] + c];
}

// m[j] = sum over node j's edges of (W_e f)_c * src[col * stride + c * cstep]: a vector gather
// (stride 32, cstep 1) or a scalar one (stride 1, cstep 0).  The edge loop runs to the slot's
// longest row so the MP_NPL streams advance together.
__device__ __forceinline__ void gather_tile(const LayerLds& L, int slot, int c, const i64 (&p0)[MP_NPL],
                                            const int (&deg)[MP_NPL], const float4* __restrict__ feat, float e0, float e1,
                                            float e2, float e3, const float* __restrict__ src, int stride, int cstep,
                                            float (&m)[MP_NPL]) {
    int kmax = 0;
#pragma unroll
    for (int j = 0; j < MP_NPL; ++j) { m[j] = 0.0f; kmax = max(kmax, deg[j]); }
#pragma unroll 1
    for (int k = 0; k < kmax; ++k) {
#pragma unroll
        for (int j = 0; j < MP_NPL; ++j) {
            if (k < deg[j]) {
                const float4 f = feat[p0[j] + k];
                const i64 v = L.ecol[slot + MP_NODES * j][k];
                m[j] += (e0 * f.x + e1 * f.y + e2 * f.z + e3 * f.w) * src[v * stride + c * cstep];
            }
        }
    }
}

// One message-passing layer h' = W_s h + W_m m, m_u = sum_p (W_e f_p) (.) h_{cols p}; the messages
// are stored when mout is given (training).  With wout given, y_u = w_out . h'_u is written instead
// of h' (last layer of inference).
__global__ __launch_bounds__(MP_THREADS) void k_lf_layer(i64 n, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                         const float4* __restrict__ feat, const float* __restrict__ wl,
                                                         const float* __restrict__ hin, float* __restrict__ hout,
                                                         float* __restrict__ mout, const float* __restrict__ wout,
                                                         float* __restrict__ y) {
    __shared__ LayerLds L;
    lds_load_w<false>(L.A, wl + P_WS);
    lds_load_w<false>(L.B, wl + P_WM);
    const int c = threadIdx.x % MP_H, slot = threadIdx.x / MP_H;
    const float e0 = wl[P_WE + c * MP_F + 0], e1 = wl[P_WE + c * MP_F + 1], e2 = wl[P_WE + c * MP_F + 2],
                e3 = wl[P_WE + c * MP_F + 3];
    const float wo = wout ? wout[c] : 0.0f;
    // every thread runs the same number of iterations (barriers; uniform width-32 shuffles)
    const i64 nn = (n + MP_TILE - 1) / MP_TILE * MP_TILE;
    for (i64 u0 = i64(blockIdx.x) * MP_TILE; u0 < nn; u0 += i64(gridDim.x) * MP_TILE) {
        i64 p0[MP_NPL];
        int deg[MP_NPL];
        float hu[MP_NPL], m[MP_NPL], acc[MP_NPL];
        stage_tile(L, u0, slot, c, n, outer, cols, p0, deg);
#pragma unroll
        for (int j = 0; j < MP_NPL; ++j) {
            const i64 u = u0 + slot + MP_NODES * j;
            hu[j] = u < n ? hin[u * MP_H + c] : 0.0f;
        }
        __syncthreads();
        gather_tile(L, slot, c, p0, deg, feat, e0, e1, e2, e3, hin, MP_H, 1, m);
#pragma unroll
        for (int j = 0; j < MP_NPL; ++j) {
            L.va[slot + MP_NODES * j][c] = hu[j];
            L.vb[slot + MP_NODES * j][c] = m[j];
            acc[j] = 0.0f;
        }
        __syncthreads();
        lds_dot_tile(L.A, L.va, slot, c, acc);
        lds_dot_tile(L.B, L.vb, slot, c, acc);
#pragma unroll
        for (int j = 0; j < MP_NPL; ++j) {
            const i64 u = u0 + slot + MP_NODES * j;
            const bool live = u < n;
            if (wout) {
                float t = wo * acc[j];
                for (int o = MP_H / 2; o > 0; o >>= 1) t += __shfl_down(t, o, MP_H);
                if (live && c == 0) y[u] = t;
            } else if (live) {
                hout[u * MP_H + c] = acc[j];
                if (mout) mout[u * MP_H + c] = m[j];
            }
        }
        __syncthreads();
    }
}

// ---- transposed pass Y^T x (layers backwards) and the symmetric output ----------------------------
// Last layer, transposed, from the scalar input: g_v = x_v W_s^T w_out + sum_p G_p (.) x_{cols p} W_m^T w_out;
// then d = W_s'^T g, q = W_m'^T g with the previous layer's weights (wprev).
__global__ __launch_bounds__(MP_THREADS) void k_lft_first(i64 n, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                          const float4* __restrict__ feat, const float* __restrict__ wl,
                                                          const float* __restrict__ wout, const float* __restrict__ wprev,
                                                          const float* __restrict__ x, float* __restrict__ d,
                                                          float* __restrict__ q) {
    __shared__ LayerLds L;
    __shared__ float vd[MP_H], vq[MP_H];
    lds_load_w<true>(L.A, wl + P_WS);
    lds_load_w<true>(L.B, wl + P_WM);
    if (threadIdx.x < MP_H) L.va[0][threadIdx.x] = wout[threadIdx.x];
    __syncthreads();
    if (threadIdx.x < MP_H) {
        vd[threadIdx.x] = lds_dot(L.A, L.va[0], threadIdx.x);      // (W_s^T w_out)_c
        vq[threadIdx.x] = lds_dot(L.B, L.va[0], threadIdx.x);      // (W_m^T w_out)_c
    }
    __syncthreads();
    lds_load_w<true>(L.A, wprev + P_WS);
    lds_load_w<true>(L.B, wprev + P_WM);
    const int c = threadIdx.x % MP_H, slot = threadIdx.x / MP_H;
    const float e0 = wl[P_WE + c * MP_F + 0], e1 = wl[P_WE + c * MP_F + 1], e2 = wl[P_WE + c * MP_F + 2],
                e3 = wl[P_WE + c * MP_F + 3];
    const float vdc = vd[c], vqc = vq[c];
    const i64 nn = (n + MP_TILE - 1) / MP_TILE * MP_TILE;
    for (i64 u0 = i64(blockIdx.x) * MP_TILE; u0 < nn; u0 += i64(gridDim.x) * MP_TILE) {
        i64 p0[MP_NPL];
        int deg[MP_NPL];
        float xv[MP_NPL], m[MP_NPL], dd[MP_NPL], qq[MP_NPL];
        stage_tile(L, u0, slot, c, n, outer, cols, p0, deg);
#pragma unroll
        for (int j = 0; j < MP_NPL; ++j) {
            const i64 v = u0 + slot + MP_NODES * j;
            xv[j] = v < n ? x[v] : 0.0f;
        }
        __syncthreads();
        gather_tile(L, slot, c, p0, deg, feat, e0, e1, e2, e3, x, 1, 0, m);
#pragma unroll
        for (int j = 0; j < MP_NPL; ++j) {
            L.va[slot + MP_NODES * j][c] = xv[j] * vdc + m[j] * vqc;
            dd[j] = 0.0f;
            qq[j] = 0.0f;
        }
        __syncthreads();
        lds_dot_tile(L.A, L.va, slot, c, dd);
        lds_dot_tile(L.B, L.va, slot, c, qq);
#pragma unroll
        for (int j = 0; j < MP_NPL; ++j) {
            const i64 v = u0 + slot + MP_NODES * j;
            if (v < n) {
                d[v * MP_H + c] = dd[j];
                q[v * MP_H + c] = qq[j];
            }
        }
        __syncthreads();
    }
}

// A middle layer, transposed: g_v = d_v + sum_p G_p (.) q_{cols p}; then d' = W_s'^T g (in place),
// q' = W_m'^T g (new buffer).  For the first layer (wprev null) the output is
// M r = s^2 r + s sigma (y + a_in . g) / 2.
template<class Z>
__global__ __launch_bounds__(MP_THREADS) void k_lft_layer(i64 n, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                          const float4* __restrict__ feat, const float* __restrict__ wl,
                                                          const float* __restrict__ wprev, float* __restrict__ d,
                                                          const float* __restrict__ q, float* __restrict__ qout,
                                                          const float* __restrict__ ain, const float* __restrict__ x,
                                                          const float* __restrict__ y, const float* __restrict__ s,
                                                          const double* __restrict__ sigma, const double* __restrict__ r,
                                                          Z* __restrict__ z) {
    __shared__ LayerLds L;
    if (wprev) {
        lds_load_w<true>(L.A, wprev + P_WS);
        lds_load_w<true>(L.B, wprev + P_WM);
    }
    const int c = threadIdx.x % MP_H, slot = threadIdx.x / MP_H;
    const float e0 = wl[P_WE + c * MP_F + 0], e1 = wl[P_WE + c * MP_F + 1], e2 = wl[P_WE + c * MP_F + 2],
                e3 = wl[P_WE + c * MP_F + 3];
    const float a = ain ? ain[c] : 0.0f;
    const double sg = sigma ? *sigma : 1.0;
    const i64 nn = (n + MP_TILE - 1) / MP_TILE * MP_TILE;
    for (i64 u0 = i64(blockIdx.x) * MP_TILE; u0 < nn; u0 += i64(gridDim.x) * MP_TILE) {
        i64 p0[MP_NPL];
        int deg[MP_NPL];
        float g[MP_NPL], m[MP_NPL], dd[MP_NPL], qq[MP_NPL];
        stage_tile(L, u0, slot, c, n, outer, cols, p0, deg);
#pragma unroll
        for (int j = 0; j < MP_NPL; ++j) {
            const i64 v = u0 + slot + MP_NODES * j;
            g[j] = v < n ? d[v * MP_H + c] : 0.0f;
        }
        __syncthreads();
        gather_tile(L, slot, c, p0, deg, feat, e0, e1, e2, e3, q, MP_H, 1, m);
        if (wprev) {
#pragma unroll
            for (int j = 0; j < MP_NPL; ++j) {
                L.va[slot + MP_NODES * j][c] = g[j] + m[j];
                dd[j] = 0.0f;
                qq[j] = 0.0f;
            }
            __syncthreads();
            lds_dot_tile(L.A, L.va, slot, c, dd);
            lds_dot_tile(L.B, L.va, slot, c, qq);
#pragma unroll
            for (int j = 0; j < MP_NPL; ++j) {
                const i64 v = u0 + slot + MP_NODES * j;
                if (v < n) {
                    d[v * MP_H + c] = dd[j];
                    qout[v * MP_H + c] = qq[j];
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < MP_NPL; ++j) {
                const i64 v = u0 + slot + MP_NODES * j;
                float t = a * (g[j] + m[j]);
                for (int o = MP_H / 2; o > 0; o >>= 1) t += __shfl_down(t, o, MP_H);
                if (v < n && c == 0) {
                    // the Jacobi term s^2 r in FP64 (exactly linear); the network correction in FP32
                    const double sv = double(s[v]);
                    z[v] = Z(sv * sv * r[v] + sv * sg * 0.5 * (double(y[v]) + double(t)));
                }
            }
        }
        __syncthreads();
    }
}

// N(r) (training forward, not symmetrised): z_u = s_u^2 r_u + s_u sigma w_out . h_u
template<class Z>
__global__ __launch_bounds__(MP_THREADS) void k_lf_output(i64 n, const float* __restrict__ h, const float* __restrict__ w,
                                                          const double* __restrict__ r, const float* __restrict__ s,
                                                          const double* __restrict__ sigma, Z* __restrict__ z) {
    const int c = threadIdx.x % MP_H;
    const double sg = *sigma;
    const i64 nn = (n + MP_NODES - 1) / MP_NODES * MP_NODES;
    for (i64 u0 = i64(blockIdx.x) * MP_NODES; u0 < nn; u0 += i64(gridDim.x) * MP_NODES) {
        const i64 u = u0 + threadIdx.x / MP_H;
        float y = u < n ? w[P_WOUT + c] * h[u * MP_H + c] : 0.0f;
        for (int o = MP_H / 2; o > 0; o >>= 1) y += __shfl_down(y, o, MP_H);
        if (c == 0 && u < n) {
            const double su = double(s[u]);
            z[u] = Z(su * su * r[u] + su * sg * double(y));
        }
    }
}

// ---- backward ------------------------------------------------------------------------------------
// dL/dz (double) -> dL/dh3; gradient of w_out into the block partial row (the b_out slot stays 0)
__global__ __launch_bounds__(MP_THREADS) void k_lf_out_back(i64 n, const double* __restrict__ gz, const float* __restrict__ s,
                                                            const double* __restrict__ sigma, const float* __restrict__ h,
                                                            const float* __restrict__ w, float* __restrict__ gh,
                                                            float* __restrict__ gpart) {
    __shared__ float red[MP_NODES][MP_H + 1];
    const int c = threadIdx.x % MP_H, slot = threadIdx.x / MP_H;
    const double sg = *sigma;
    float gw = 0.0f;
    for (i64 u = i64(blockIdx.x) * MP_NODES + slot; u < n; u += i64(gridDim.x) * MP_NODES) {
        const float dy = float(double(s[u]) * sg * gz[u]);
        gh[u * MP_H + c] = dy * w[P_WOUT + c];
        gw += dy * h[u * MP_H + c];
    }
    float* row = gpart + i64(blockIdx.x) * NPARAM;
    red[slot][c] = gw;
    __syncthreads();
    if (slot == 0) {
        float t = 0.0f;
        for (int k = 0; k < MP_NODES; ++k) t += red[k][c];
        row[P_WOUT + c] = t;
    }
}

// gradients of W_s, W_m; dL/dh_in (direct part) = W_s^T gh, dL/dm = W_m^T gh.
__global__ __launch_bounds__(MP_THREADS) void k_lf_layer_back1(i64 n, const float* __restrict__ gh, const float* __restrict__ hin,
                                                               const float* __restrict__ m, const float* __restrict__ wl,
                                                               float* __restrict__ gdir, float* __restrict__ gm,
                                                               float* __restrict__ gpart_layer) {
    __shared__ float Ws[MP_H][MP_H + 1], Wm[MP_H][MP_H + 1];
    __shared__ float red[MP_NODES][MP_H + 1];
    for (int i = threadIdx.x; i < MP_H * MP_H; i += blockDim.x) {
        Ws[i / MP_H][i % MP_H] = wl[P_WS + i];
        Wm[i / MP_H][i % MP_H] = wl[P_WM + i];
    }
    __syncthreads();
    const int c = threadIdx.x % MP_H, slot = threadIdx.x / MP_H;
    float gws[MP_H], gwm[MP_H];
#pragma unroll
    for (int k = 0; k < MP_H; ++k) { gws[k] = 0.0f; gwm[k] = 0.0f; }
    const i64 nn = (n + MP_NODES - 1) / MP_NODES * MP_NODES;
    for (i64 u0 = i64(blockIdx.x) * MP_NODES; u0 < nn; u0 += i64(gridDim.x) * MP_NODES) {
        const i64 u = u0 + slot;
        const bool live = u < n;
        float da = 0.0f, hu = 0.0f, mu = 0.0f;
        if (live) {
            da = gh[u * MP_H + c];
            hu = hin[u * MP_H + c];
            mu = m[u * MP_H + c];
        }
        float dh = 0.0f, dm = 0.0f;
#pragma unroll
        for (int k = 0; k < MP_H; ++k) {
            const float hk = __shfl(hu, k, MP_H), mk = __shfl(mu, k, MP_H), dk = __shfl(da, k, MP_H);
            gws[k] += da * hk;                  // dW_s[c][k]
            gwm[k] += da * mk;                  // dW_m[c][k]
            dh += Ws[k][c] * dk;                // (W_s^T da)_c
            dm += Wm[k][c] * dk;                // (W_m^T da)_c
        }
        if (live) {
            gdir[u * MP_H + c] = dh;
            gm[u * MP_H + c] = dm;
        }
    }
    float* row = gpart_layer + i64(blockIdx.x) * NPARAM;
    for (int k = 0; k < MP_H; ++k) {
        red[slot][c] = gws[k];
        __syncthreads();
        if (slot == 0) {
            float t = 0.0f;
            for (int q = 0; q < MP_NODES; ++q) t += red[q][c];
            row[P_WS + c * MP_H + k] = t;
        }
        __syncthreads();
        red[slot][c] = gwm[k];
        __syncthreads();
        if (slot == 0) {
            float t = 0.0f;
            for (int q = 0; q < MP_NODES; ++q) t += red[q][c];
            row[P_WM + c * MP_H + k] = t;
        }
        __syncthreads();
    }
}

// dL/dh_in[v] = direct + sum_{u in N(v)} g_uv (.) dm_u ;  dW_e[c][f] += sum dm_u[c] h_v[c] f_uv[f]
__global__ __launch_bounds__(MP_THREADS) void k_lf_layer_back2(i64 n, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                               const float4* __restrict__ feat, const float* __restrict__ wl,
                                                               const float* __restrict__ gm, const float* __restrict__ hin,
                                                               const float* __restrict__ gdir, float* __restrict__ ghin,
                                                               float* __restrict__ gpart_layer) {
    __shared__ float red[MP_NODES][MP_H + 1];
    const int c = threadIdx.x % MP_H, slot = threadIdx.x / MP_H;
    const float e0 = wl[P_WE + c * MP_F + 0], e1 = wl[P_WE + c * MP_F + 1], e2 = wl[P_WE + c * MP_F + 2],
                e3 = wl[P_WE + c * MP_F + 3];
    float ge[MP_F] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (i64 v = i64(blockIdx.x) * MP_NODES + slot; v < n; v += i64(gridDim.x) * MP_NODES) {
        float acc = gdir[v * MP_H + c];
        const float hv = hin[v * MP_H + c];
        for (i64 p = outer[v]; p < outer[v + 1]; ++p) {
            const float4 f = feat[p];
            const float dmu = gm[i64(cols[p]) * MP_H + c];
            acc += (e0 * f.x + e1 * f.y + e2 * f.z + e3 * f.w) * dmu;
            const float t = dmu * hv;
            ge[0] += t * f.x; ge[1] += t * f.y; ge[2] += t * f.z; ge[3] += t * f.w;
        }
        ghin[v * MP_H + c] = acc;
    }
    float* row = gpart_layer + i64(blockIdx.x) * NPARAM;
    for (int f = 0; f < MP_F; ++f) {
        red[slot][c] = ge[f];
        __syncthreads();
        if (slot == 0) {
            float t = 0.0f;
            for (int q = 0; q < MP_NODES; ++q) t += red[q][c];
            row[P_WE + c * MP_F + f] = t;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(MP_THREADS) void k_lf_input_back(i64 n, const float* __restrict__ gh, const float* __restrict__ x,
                                                              float* __restrict__ gpart) {
    __shared__ float red[MP_NODES][MP_H + 1];
    const int c = threadIdx.x % MP_H, slot = threadIdx.x / MP_H;
    float ga = 0.0f;
    for (i64 u = i64(blockIdx.x) * MP_NODES + slot; u < n; u += i64(gridDim.x) * MP_NODES) ga += gh[u * MP_H + c] * x[u];
    float* row = gpart + i64(blockIdx.x) * NPARAM;
    red[slot][c] = ga;
    __syncthreads();
    if (slot == 0) {
        float t = 0.0f;
        for (int q = 0; q < MP_NODES; ++q) t += red[q][c];
        row[P_AIN + c] = t;
    }
}

__global__ void k_lf_reduce(const float* __restrict__ gpart, int nb, float* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += gpart[i64(b) * NPARAM + i];
    grad[i] = float(t);
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       float lr, float b1, float b2, float eps, int t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    const float gi = isfinite(g[i]) ? g[i] : 0.0f;
    m1[i] = b1 * m1[i] + (1.0f - b1) * gi;
    m2[i] = b2 * m2[i] + (1.0f - b2) * gi * gi;
    const float mh = m1[i] / (1.0f - powf(b1, float(t)));
    const float vh = m2[i] / (1.0f - powf(b2, float(t)));
    w[i] -= lr * mh / (sqrtf(vh) + eps);
}

// loss pieces: d = z - e ; partial sums of d.(A z - r) and e.r  (A e = r, so A d = A z - r)
__global__ __launch_bounds__(256) void k_lf_loss(i64 n, const double* __restrict__ z, const double* __restrict__ e,
                                                 const double* __restrict__ az, const double* __restrict__ r,
                                                 double* __restrict__ gz, double* __restrict__ p1, double* __restrict__ p2) {
    __shared__ double sh[256 / 64];
    double a = 0.0, b = 0.0;
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) {
        const double ad = az[i] - r[i];
        a += (z[i] - e[i]) * ad;
        b += e[i] * r[i];
        gz[i] = ad;                 // scaled by 2 / e.r afterwards
    }
    a = block_sum<double>(a, sh);
    b = block_sum<double>(b, sh);
    if (threadIdx.x == 0) { p1[blockIdx.x] = a; p2[blockIdx.x] = b; }
}

__global__ void k_scale_vec(i64 n, double* __restrict__ v, const double* __restrict__ sums) {
    const double f = 2.0 / sums[1];
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) v[i] *= f;
}

__global__ void k_fill_f32(i64 n, float* __restrict__ x, float v) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) x[i] = v;
}

// blocks of the node-loop kernels; the layer kernels take MP_TILE nodes per iteration, the others
// MP_NODES, and all of them stride over the grid
unsigned mp_blocks(i64 n) { return unsigned(std::min<i64>(std::max<i64>((n + MP_NODES - 1) / MP_NODES, 1), 8192)); }
unsigned mp_tile_blocks(i64 n) { return unsigned(std::min<i64>(std::max<i64>((n + MP_TILE - 1) / MP_TILE, 1), 4096)); }

void learned_alloc_train(LearnedGraph& g) {
    if (g.train_ready) return;
    const size_t nh = size_t(g.n) * MP_H;
    for (int l = 2; l <= MP_L; ++l) g.h[l].alloc(nh);        // h[0], h[1] exist for inference
    for (int l = 0; l < MP_L; ++l) g.m[l].alloc(nh);
    g.g0.alloc(nh); g.g1.alloc(nh); g.gdir.alloc(nh); g.gm.alloc(nh);
    g.gpart.alloc(size_t(MP_GRAD_BLOCKS) * NPARAM);
    g.grad.alloc(NPARAM);
    g.zbuf.alloc(size_t(g.n));
    g.az.alloc(size_t(g.n));
    g.train_ready = true;
}

void learned_scaled_input(LearnedGraph& g, const float* w, const double* r, hipStream_t s) {
    const i64 n = g.n;
    const unsigned vb = grid_stride_blocks(n), mb = mp_blocks(n);
    hipLaunchKernelGGL(k_lf_rnorm<double>, dim3(vb), dim3(256), 0, s, n, r, g.s.p, g.part.p);
    hipLaunchKernelGGL(k_lf_sigma, dim3(1), dim3(1024), 0, s, g.part.p, int(vb), n, g.sigma.p);
    hipLaunchKernelGGL(k_lf_input<double>, dim3(mb), dim3(MP_THREADS), 0, s, n, r, g.s.p, g.sigma.p, w, g.x.p, g.h[0].p);
}

// Training forward: z = N(r) with every activation and message kept (h[0..3], m[0..2]).
void learned_forward_train(LearnedGraph& g, const float* w, const double* r, double* z, hipStream_t s) {
    const i64 n = g.n;
    const unsigned mb = mp_blocks(n), tb = mp_tile_blocks(n);
    learned_scaled_input(g, w, r, s);
    const float4* feat = reinterpret_cast<const float4*>(g.feat.p);
    for (int l = 0; l < MP_L; ++l)
        hipLaunchKernelGGL(k_lf_layer, dim3(tb), dim3(MP_THREADS), 0, s, n, g.outer.p, g.cols.p, feat,
                           w + P_LAYER0 + l * P_LAYER_SIZE, g.h[l].p, g.h[l + 1].p, g.m[l].p, (const float*)nullptr,
                           (float*)nullptr);
    hipLaunchKernelGGL(k_lf_output<double>, dim3(mb), dim3(MP_THREADS), 0, s, n, g.h[MP_L].p, w, r, g.s.p, g.sigma.p, z);
}

// ---- inference: one node per lane, channel-major activations --------------------------------------
// The applied operator runs with lane = node: activations are stored channel-blocked (float4 of
// channels 4b..4b+3 at a4[b * n + u]),
// the graph as ELL with DM slots per node (k-major; rows of the local pattern padded with the node
// itself and zero features), and the weights are wave-uniform (scalar loads).  A lane first loads
// its DM columns and features, then gathers four channels at a time: the DM float4 loads of one
// block are independent and coalesced across neighbouring nodes.  Offsets are 32-bit (32 n < 2^31).  The last
// forward layer and the first transposed layer only need w_out-contracted weights:
// v_s = W_s2^T w_out, v_m = W_m2^T w_out (64 floats, computed per application by k_soa_prep).
constexpr int SOA_THREADS = 256;

__global__ void k_soa_ell(i64 n, int D, const i64* __restrict__ outer, const i32* __restrict__ cols,
                          const float4* __restrict__ feat, i32* __restrict__ ecol, float4* __restrict__ ef) {
    const i64 u = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const i64 p0 = outer[u];
    const int deg = int(outer[u + 1] - p0);
    for (int k = 0; k < D; ++k) {
        const bool e = k < deg;
        ecol[i64(k) * n + u] = e ? cols[p0 + k] : i32(u);
        ef[i64(k) * n + u] = e ? feat[p0 + k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// v[0:32] = W_s^T w_out, v[32:64] = W_m^T w_out for the last layer
__global__ void k_soa_prep(const float* __restrict__ wl, const float* __restrict__ wout, float* __restrict__ v) {
    const int c = threadIdx.x % MP_H, which = threadIdx.x / MP_H;
    if (which > 1) return;
    const float* W = wl + (which ? P_WM : P_WS);
    float acc = 0.0f;
    for (int o = 0; o < MP_H; ++o) acc += W[o * MP_H + c] * wout[o];
    v[which * MP_H + c] = acc;
}

template<int DM>
struct EllRow {
    uint32_t col[DM];
    float4 f[DM];
};

template<int DM>
__device__ __forceinline__ void ell_row(uint32_t n, uint32_t u, const i32* __restrict__ ecol, const float4* __restrict__ ef,
                                        EllRow<DM>& e) {
#pragma unroll
    for (int k = 0; k < DM; ++k) {
        e.col[k] = uint32_t(ecol[k * n + u]);
        e.f[k] = ef[k * n + u];
    }
}

// Activations are channel-blocked float4: a4[b * n + u] holds channels 4b..4b+3 of node u.
__device__ __forceinline__ float comp(const float4& v, int j) { return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w; }

// m[c] = sum_k (W_e f_k)_c * x[col_k]  (scalar input)
template<int DM>
__device__ __forceinline__ void soa_gather_x(const EllRow<DM>& e, const float* __restrict__ wl, const float* __restrict__ x,
                                             float (&m)[MP_H]) {
    float v[DM];
#pragma unroll
    for (int k = 0; k < DM; ++k) v[k] = x[e.col[k]];
#pragma unroll
    for (int c = 0; c < MP_H; ++c) {
        const float w0 = wl[P_WE + c * MP_F + 0], w1 = wl[P_WE + c * MP_F + 1], w2 = wl[P_WE + c * MP_F + 2],
                    w3 = wl[P_WE + c * MP_F + 3];
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < DM; ++k) acc += (w0 * e.f[k].x + w1 * e.f[k].y + w2 * e.f[k].z + w3 * e.f[k].w) * v[k];
        m[c] = acc;
    }
}

// m[c] = sum_k (W_e f_k)_c * a_{col_k}[c]  (channel-blocked vector input: DM float4 loads per block)
template<int DM>
__device__ __forceinline__ void soa_gather_v(uint32_t n, const EllRow<DM>& e, const float* __restrict__ wl,
                                             const float4* __restrict__ a4, float (&m)[MP_H]) {
#pragma unroll
    for (int b = 0; b < MP_H / 4; ++b) {
        float4 v[DM];
#pragma unroll
        for (int k = 0; k < DM; ++k) v[k] = a4[uint32_t(b) * n + e.col[k]];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = 4 * b + j;
            const float w0 = wl[P_WE + c * MP_F + 0], w1 = wl[P_WE + c * MP_F + 1], w2 = wl[P_WE + c * MP_F + 2],
                        w3 = wl[P_WE + c * MP_F + 3];
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < DM; ++k)
                acc += (w0 * e.f[k].x + w1 * e.f[k].y + w2 * e.f[k].z + w3 * e.f[k].w) * comp(v[k], j);
            m[c] = acc;
        }
    }
}

__device__ __forceinline__ void load_node(uint32_t n, uint32_t u, const float4* __restrict__ a4, float (&h)[MP_H]) {
#pragma unroll
    for (int b = 0; b < MP_H / 4; ++b) {
        const float4 v = a4[uint32_t(b) * n + u];
        h[4 * b] = v.x; h[4 * b + 1] = v.y; h[4 * b + 2] = v.z; h[4 * b + 3] = v.w;
    }
}

// out = A a (+ init) for a row-major 32 x 32 A, or A^T a when T; stored channel-blocked
template<bool T>
__device__ __forceinline__ void product_store(uint32_t n, uint32_t u, const float* __restrict__ A, const float (&a)[MP_H],
                                              const float* init, float4* __restrict__ out4) {
    for (int b = 0; b < MP_H / 4; ++b) {
        float o4[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = 4 * b + j;
            float t = init ? init[o] : 0.0f;
#pragma unroll
            for (int k = 0; k < MP_H; ++k) t += A[T ? k * MP_H + o : o * MP_H + k] * a[k];
            o4[j] = t;
        }
        out4[uint32_t(b) * n + u] = make_float4(o4[0], o4[1], o4[2], o4[3]);
    }
}

// Forward layers 0 and 1: h' = W_s h + W_m m.  Layer 0 reads the scalar input (h0 = a_in x).
template<int DM, bool FIRST>
__global__ __launch_bounds__(SOA_THREADS) void k_soa_layer(uint32_t n, const i32* __restrict__ ecol,
                                                           const float4* __restrict__ ef, const float* __restrict__ wl,
                                                           const float* __restrict__ ain, const float* __restrict__ x,
                                                           const float4* __restrict__ hin, float4* __restrict__ hout) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    // the own-node part W_s h first, so that only 32 partial outputs stay live through the gather
    float acc[MP_H];
    {
        float h[MP_H];
        if (FIRST) {
            const float xu = x[u];
#pragma unroll
            for (int c = 0; c < MP_H; ++c) h[c] = ain[c] * xu;
        } else {
            load_node(n, u, hin, h);
        }
        for (int o = 0; o < MP_H; ++o) {
            float t = 0.0f;
#pragma unroll
            for (int k = 0; k < MP_H; ++k) t += wl[P_WS + o * MP_H + k] * h[k];
            acc[o] = t;
        }
    }
    EllRow<DM> e;
    ell_row<DM>(n, u, ecol, ef, e);
    float m[MP_H];
    if (FIRST) {
        soa_gather_x<DM>(e, wl, x, m);
#pragma unroll
        for (int c = 0; c < MP_H; ++c) m[c] *= ain[c];
    } else {
        soa_gather_v<DM>(n, e, wl, hin, m);
    }
    product_store<false>(n, u, wl + P_WM, m, acc, hout);
}

// Forward layer 2 contracted with w_out: y_u = v_s . h_u + v_m . m_u
template<int DM>
__global__ __launch_bounds__(SOA_THREADS) void k_soa_last(uint32_t n, const i32* __restrict__ ecol, const float4* __restrict__ ef,
                                                          const float* __restrict__ wl, const float* __restrict__ v,
                                                          const float4* __restrict__ hin, float* __restrict__ y) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    EllRow<DM> e;
    ell_row<DM>(n, u, ecol, ef, e);
    float m[MP_H], h[MP_H];
    soa_gather_v<DM>(n, e, wl, hin, m);
    load_node(n, u, hin, h);
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < MP_H; ++c) acc += v[c] * h[c] + v[MP_H + c] * m[c];
    y[u] = acc;
}

// Transposed layer 2 from the scalar input: g = x_u v_s + v_m (.) sum_k G_k x_{col k};
// d = W_s1^T g, q = W_m1^T g.
template<int DM>
__global__ __launch_bounds__(SOA_THREADS) void k_soa_tfirst(uint32_t n, const i32* __restrict__ ecol,
                                                            const float4* __restrict__ ef, const float* __restrict__ wl,
                                                            const float* __restrict__ v, const float* __restrict__ wprev,
                                                            const float* __restrict__ x, float4* __restrict__ d,
                                                            float4* __restrict__ q) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    EllRow<DM> e;
    ell_row<DM>(n, u, ecol, ef, e);
    float m[MP_H];
    soa_gather_x<DM>(e, wl, x, m);
    const float xu = x[u];
#pragma unroll
    for (int c = 0; c < MP_H; ++c) m[c] = xu * v[c] + v[MP_H + c] * m[c];
    product_store<true>(n, u, wprev + P_WS, m, nullptr, d);
    product_store<true>(n, u, wprev + P_WM, m, nullptr, q);
}

// Transposed middle layer: g = d_u + sum_k G_k (.) q_{col k}; d = W_s'^T g (in place), q' = W_m'^T g.
template<int DM>
__global__ __launch_bounds__(SOA_THREADS) void k_soa_tmid(uint32_t n, const i32* __restrict__ ecol, const float4* __restrict__ ef,
                                                          const float* __restrict__ wl, const float* __restrict__ wprev,
                                                          float4* __restrict__ d, const float4* __restrict__ q,
                                                          float4* __restrict__ qout) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    EllRow<DM> e;
    ell_row<DM>(n, u, ecol, ef, e);
    float g[MP_H], dv[MP_H];
    soa_gather_v<DM>(n, e, wl, q, g);
    load_node(n, u, d, dv);
#pragma unroll
    for (int c = 0; c < MP_H; ++c) g[c] += dv[c];
    product_store<true>(n, u, wprev + P_WS, g, nullptr, d);
    product_store<true>(n, u, wprev + P_WM, g, nullptr, qout);
}

// Transposed layer 0 and the output: t = a_in . (d_u + sum_k G_k (.) q_{col k});
// M r = s^2 r (FP64) + s sigma (y + t) / 2.
template<int DM, class Z>
__global__ __launch_bounds__(SOA_THREADS) void k_soa_tlast(uint32_t n, const i32* __restrict__ ecol, const float4* __restrict__ ef,
                                                           const float* __restrict__ wl, const float* __restrict__ ain,
                                                           const float4* __restrict__ d, const float4* __restrict__ q,
                                                           const float* __restrict__ y, const float* __restrict__ s,
                                                           const double* __restrict__ sigma, const double* __restrict__ r,
                                                           Z* __restrict__ z) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    EllRow<DM> e;
    ell_row<DM>(n, u, ecol, ef, e);
    float g[MP_H], dv[MP_H];
    soa_gather_v<DM>(n, e, wl, q, g);
    load_node(n, u, d, dv);
    float t = 0.0f;
#pragma unroll
    for (int c = 0; c < MP_H; ++c) t += ain[c] * (g[c] + dv[c]);
    const double su = double(s[u]);
    z[u] = Z(su * su * r[u] + su * (*sigma) * 0.5 * (double(y[u]) + double(t)));
}

// ---- the learned operator as an explicit sparse matrix --------------------------------------------
// N is linear, so M = S^2 + S (Y + Y^T)/2 S is a fixed matrix on the three-hop pattern P3 of the local
// graph.  Row u of Y is the transposed network applied to a unit impulse at u (Y^T e_u), and the
// layer structure collapses it to small sums (all vectors have 32 channels):
//   g2(v) = [v = u] v_s + sum_f f_uv[f] nu_f,          nu_f = W_e2[:, f] (.) v_m,  v_s/m = W_s2/m2^T w_out
//   d1(v) = W_s1^T g2(v),  t1(v) = W_m1^T g2(v)        (linear in the 5 coefficients: 10 basis vectors)
//   g1(w) = [w in P1] d1(w) + sum_{v in row(w), v in P1} (W_e1 f_wv) (.) t1(v)            (w in P2)
//   Y_ux  = [x in P2] alpha . g1(x) + sum_{w in row(x), w in P2} f_xw . (Lambda^T g1(w))   (x in P3)
// with alpha = W_s0 a_in and Lambda[c][f] = sum_k W_m0[c][k] a_in[k] W_e0[k][f].  Per node this is a
// few thousand flops instead of the network's six gather passes per application; the apply becomes
// one SpMV (FP32 values, FP64 vectors).  Node sets are built in LDS bitmaps over the id window
// [u - 3b, u + 3b] (b = bandwidth of the local pattern); a row that does not fit leaves the
// operator uncompiled and the network kernels are used instead.
constexpr int CP_THREADS = 64;
constexpr int CP_P1 = 32, CP_P2 = 160, CP_P3 = 512;     // set capacities
constexpr int CP_WORDS = 2048;                          // bitmap words: id window of 65536

struct CompileBasis {
    float d1[5][MP_H], t1[5][MP_H];     // basis of d1 / t1: index 0 = the impulse term, 1 + f = feature f
    float alpha[MP_H];
    float lam[MP_H][MP_F];
};

// one block of 32 threads
__global__ void k_lf_basis(const float* __restrict__ w, CompileBasis* __restrict__ B) {
    const int c = threadIdx.x;
    const float* L0 = w + P_LAYER0;
    const float* L1 = w + P_LAYER0 + P_LAYER_SIZE;
    const float* L2 = w + P_LAYER0 + 2 * P_LAYER_SIZE;
    const float* wout = w + P_WOUT;
    const float* ain = w + P_AIN;
    __shared__ float vs[MP_H], vm[MP_H], nu[MP_F][MP_H];
    float a = 0.0f, b = 0.0f;
    for (int o = 0; o < MP_H; ++o) { a += L2[P_WS + o * MP_H + c] * wout[o]; b += L2[P_WM + o * MP_H + c] * wout[o]; }
    vs[c] = a;
    vm[c] = b;
    __syncthreads();
    for (int f = 0; f < MP_F; ++f) nu[f][c] = L2[P_WE + c * MP_F + f] * vm[c];
    __syncthreads();
    for (int m = 0; m < 5; ++m) {
        const float* g = m == 0 ? vs : nu[m - 1];
        float d = 0.0f, t = 0.0f;
        for (int k = 0; k < MP_H; ++k) { d += L1[P_WS + k * MP_H + c] * g[k]; t += L1[P_WM + k * MP_H + c] * g[k]; }
        B->d1[m][c] = d;
        B->t1[m][c] = t;
    }
    float al = 0.0f;
    for (int k = 0; k < MP_H; ++k) al += L0[P_WS + c * MP_H + k] * ain[k];
    B->alpha[c] = al;
    for (int f = 0; f < MP_F; ++f) {
        float l = 0.0f;
        for (int k = 0; k < MP_H; ++k) l += L0[P_WM + c * MP_H + k] * ain[k] * L0[P_WE + k * MP_F + f];
        B->lam[c][f] = l;
    }
}

__device__ __forceinline__ int find_sorted(const i32* a, int n, i32 key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return (lo < n && a[lo] == key) ? lo : -1;
}

// Union of the rows of `src` (ns nodes) into a sorted list `dst` through the bitmap of `words` words
// (a multiple of CP_THREADS) starting at id `base`; returns its size, or -1 when it exceeds cap or
// the id window.
__device__ int union_rows(const i64* __restrict__ outer, const i32* __restrict__ cols, const i32* src, int ns, i64 base,
                          unsigned* bm, int words, i32* dst, int cap, i64* rs, int* pre) {
    for (int i = threadIdx.x; i < words; i += blockDim.x) bm[i] = 0u;
    __shared__ int bad;
    // the rows of src flattened (rs[k]: start of row k in cols, rs[CP_P2], pre[CP_P2 + 1]): row k's
    // entries are [pre[k], pre[k + 1]) of the concatenation, so every entry is loaded in parallel
    // instead of row after row
    static_assert(CP_P2 <= 3 * CP_THREADS && CP_THREADS == 64, "three rows per lane in the scan, one wave");
    if (threadIdx.x == 0) bad = 0;
    {
        const int l = threadIdx.x;
        int len[3] = {0, 0, 0}, sum = 0;
        for (int q = 0; q < 3; ++q) {
            const int k = 3 * l + q;
            if (k < ns) {
                const i64 v = src[k];
                rs[k] = outer[v];
                len[q] = int(outer[v + 1] - rs[k]);
            }
            sum += len[q];
        }
        int incl = sum;                                  // inclusive wave scan of the lane sums
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (l >= o) incl += y;
        }
        int run = incl - sum;
        for (int q = 0; q < 3; ++q) {
            const int k = 3 * l + q;
            if (k < ns) pre[k] = run;
            run += len[q];
        }
        if (l == 63) pre[ns] = incl;
    }
    __syncthreads();
    const int entries = pre[ns];
    for (int e = threadIdx.x; e < entries; e += blockDim.x) {
        int lo = 0, hi = ns - 1;                         // the row k with pre[k] <= e < pre[k + 1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= e) lo = mid; else hi = mid - 1;
        }
        const i64 p = rs[lo] + (e - pre[lo]);
        const i64 off = i64(cols[p]) - base;
        if (off < 0 || off >= i64(words) * 32) bad = 1;
        else atomicOr(&bm[off >> 5], 1u << (off & 31));
    }
    __syncthreads();
    if (bad) return -1;
    // ordered compaction: each thread owns a contiguous run of words
    const int PER = words / CP_THREADS;          // words is a multiple of CP_THREADS
    __shared__ int cnt[CP_THREADS];
    int c = 0;
    for (int i = 0; i < PER; ++i) c += __popc(bm[threadIdx.x * PER + i]);
    cnt[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < CP_THREADS; ++t) { const int v = cnt[t]; cnt[t] = acc; acc += v; }
        bad = acc;
    }
    __syncthreads();
    const int total = bad;
    if (total > cap) return -1;
    int pos = cnt[threadIdx.x];
    for (int i = 0; i < PER; ++i) {
        unsigned word = bm[threadIdx.x * PER + i];
        while (word) {
            const int b = __ffs(word) - 1;
            word &= word - 1;
            dst[pos++] = i32(base + i64(threadIdx.x * PER + i) * 32 + b);
        }
    }
    __syncthreads();
    return total;
}

// PASS 0: |P3(u)| into cnt[u] (-1 on overflow).  PASS 1: columns and Y row values of u at m_outer[u].
// PASS 2: both in one pass, |P3(u)| into cnt[u] and the row at u * CP_P3 (a padded layout that
// k_lf_compact then packs by the scanned counts).
template<int PASS>
__global__ __launch_bounds__(CP_THREADS) void k_lf_compile(i64 n, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                           const float4* __restrict__ feat, i64 band, int words,
                                                           const CompileBasis* __restrict__ B,
                                                           const float* __restrict__ w, i64* __restrict__ cnt,
                                                           const i64* __restrict__ m_outer, i32* __restrict__ m_cols,
                                                           float* __restrict__ yrow, int* __restrict__ overflow,
                                                           int cdrows) {
    // bitmap of `words` words, reused after the unions for the contracted basis of cdrows 1-hop
    // entries (cdrows >= the longest row of the graph); sized by the launch
    extern __shared__ unsigned bm[];
    __shared__ i32 p1[CP_P1], p2[CP_P2], p3[CP_P3];
    __shared__ float cf[CP_P1][5];
    __shared__ float aw[CP_P2], kw[CP_P2][MP_F];
    const i64 u = blockIdx.x;
    if (u >= n) return;
    const i64 r0 = outer[u];
    const int n1 = int(outer[u + 1] - r0);
    if (n1 > CP_P1 || n1 > cdrows) {
        if (threadIdx.x == 0) { if (PASS != 1) cnt[u] = 0; *overflow = 1; }
        return;
    }
    for (int k = threadIdx.x; k < n1; k += blockDim.x) {
        p1[k] = cols[r0 + k];
        const float4 f = feat[r0 + k];
        cf[k][0] = cols[r0 + k] == i32(u) ? 1.0f : 0.0f;
        cf[k][1] = f.x; cf[k][2] = f.y; cf[k][3] = f.z; cf[k][4] = f.w;
    }
    __syncthreads();
    const i64 base2 = std::max<i64>(u - 2 * band, 0), base3 = std::max<i64>(u - 3 * band, 0);
    __shared__ i64 rs[CP_P2];
    __shared__ int pre[CP_P2 + 1];
    const int n2 = union_rows(outer, cols, p1, n1, base2, bm, words, p2, CP_P2, rs, pre);
    if (n2 < 0) {
        if (threadIdx.x == 0) { if (PASS != 1) cnt[u] = 0; *overflow = 1; }
        return;
    }
    const int n3 = union_rows(outer, cols, p2, n2, base3, bm, words, p3, CP_P3, rs, pre);
    if (n3 < 0) {
        if (threadIdx.x == 0) { if (PASS != 1) cnt[u] = 0; *overflow = 1; }
        return;
    }
    if (PASS == 2 && threadIdx.x == 0) cnt[u] = n3;
    if (PASS == 0) {
        if (threadIdx.x == 0) cnt[u] = n3;
        return;
    }
    const float* L1 = w + P_LAYER0 + P_LAYER_SIZE;
    // the basis contracted with the coefficients of every j in P1, once per node (the bitmap's
    // LDS is free after the unions): cd[j][c] = sum_m cf[j][m] d1[m][c], ct[j][c] likewise with t1
    float (*cd)[MP_H] = reinterpret_cast<float (*)[MP_H]>(bm);
    float (*ct)[MP_H] = cd + cdrows;
    for (int idx = threadIdx.x; idx < n1 * MP_H; idx += blockDim.x) {
        const int j = idx / MP_H, cc = idx - j * MP_H;
        float a = 0.0f, b = 0.0f;
        for (int m = 0; m < 5; ++m) {
            a += cf[j][m] * B->d1[m][cc];
            b += cf[j][m] * B->t1[m][cc];
        }
        cd[j][cc] = a;
        ct[j][cc] = b;
    }
    __syncthreads();
    // g1(w) for w in P2: lanes 0..31 channels of one w, lanes 32..63 the next w
    const int c = threadIdx.x & 31, half = threadIdx.x >> 5;
    const float e0 = L1[P_WE + c * MP_F + 0], e1 = L1[P_WE + c * MP_F + 1], e2 = L1[P_WE + c * MP_F + 2],
                e3 = L1[P_WE + c * MP_F + 3];
    const float alpha_c = B->alpha[c];
    float lam_c[MP_F];
    for (int f = 0; f < MP_F; ++f) lam_c[f] = B->lam[c][f];
    // row wn of each half is staged in LDS 32 entries at a time, one entry per lane, so its
    // column, P1 index and features are loaded in parallel rather than one after another
    __shared__ float4 rf[2][32];
    __shared__ int rj[2][32];
    for (int i0 = 0; i0 < n2; i0 += 2) {
        const int i = i0 + half;
        float g = 0.0f;
        i64 rb = 0, re = 0;
        if (i < n2) {
            const i32 wn = p2[i];
            const int j = find_sorted(p1, n1, wn);
            if (j >= 0) g += cd[j][c];
            rb = outer[wn];
            re = outer[wn + 1];
        }
        // both halves step through their rows together (the block is one wave)
        i64 len = re - rb, other = __shfl_xor(len, 32);
        const i64 steps = len > other ? len : other;
        for (i64 k0 = 0; k0 < steps; k0 += 32) {
            if (k0 + c < len) {
                const i64 p = rb + k0 + c;
                rj[half][c] = find_sorted(p1, n1, cols[p]);
                rf[half][c] = feat[p];
            }
            __syncthreads();
            const int cnt = int(len - k0 < 32 ? len - k0 : 32);
            for (int k = 0; k < cnt; ++k) {
                const int jv = rj[half][k];
                if (jv < 0) continue;
                const float4 f = rf[half][k];
                g += (e0 * f.x + e1 * f.y + e2 * f.z + e3 * f.w) * ct[jv][c];
            }
            __syncthreads();
        }
        // alpha . g and Lambda^T g over the 32 channels of each half
        float a = alpha_c * g;
        float kf[MP_F];
        for (int f = 0; f < MP_F; ++f) kf[f] = lam_c[f] * g;
        for (int o = 16; o > 0; o >>= 1) {
            a += __shfl_down(a, o, 32);
            for (int f = 0; f < MP_F; ++f) kf[f] += __shfl_down(kf[f], o, 32);
        }
        if (c == 0 && i < n2) {
            aw[i] = a;
            for (int f = 0; f < MP_F; ++f) kw[i][f] = kf[f];
        }
    }
    __syncthreads();
    const i64 o = PASS == 2 ? u * CP_P3 : m_outer[u];
    for (int k = threadIdx.x; k < n3; k += blockDim.x) {
        const i32 x = p3[k];
        const int ix = find_sorted(p2, n2, x);
        float y = ix >= 0 ? aw[ix] : 0.0f;
        // four entries of row x loaded together, then used in order
        const i64 xe = outer[x + 1];
        for (i64 p = outer[x]; p < xe; p += 4) {
            i32 cc[4];
            float4 ff[4];
            for (int t = 0; t < 4; ++t)
                if (p + t < xe) { cc[t] = cols[p + t]; ff[t] = feat[p + t]; }
            for (int t = 0; t < 4; ++t) {
                if (p + t >= xe) break;
                const int iw = find_sorted(p2, n2, cc[t]);
                if (iw < 0) continue;
                const float4 f = ff[t];
                y += f.x * kw[iw][0] + f.y * kw[iw][1] + f.z * kw[iw][2] + f.w * kw[iw][3];
            }
        }
        m_cols[o + k] = x;
        yrow[o + k] = y;
    }
}

// M_ux = [u = x] s_u^2 + s_u s_x (Y_ux + Y_xu) / 2 (Y_xu found in row x by binary search)
__global__ void k_lf_symmetrize(i64 n, const i64* __restrict__ m_outer, const i32* __restrict__ m_cols,
                                const float* __restrict__ yrow, const float* __restrict__ s, float* __restrict__ mval) {
    const i64 u = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (u >= n) return;
    const double su = s[u];
    for (i64 p = m_outer[u] + (threadIdx.x & 63); p < m_outer[u + 1]; p += 64) {
        const i32 x = m_cols[p];
        const i64 b = m_outer[x], e = m_outer[x + 1];
        const int q = find_sorted(m_cols + b, int(e - b), i32(u));
        const double yxu = q >= 0 ? double(yrow[b + q]) : 0.0;
        double v = su * double(s[x]) * 0.5 * (double(yrow[p]) + yxu);
        if (x == i32(u)) v += su * su;
        mval[p] = float(v);
    }
}

__global__ void k_lf_off16(i64 n, const i64* __restrict__ outer, const i32* __restrict__ cols, int16_t* __restrict__ off) {
    const i64 r = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (r >= n) return;
    for (i64 p = outer[r] + (threadIdx.x & 63); p < outer[r + 1]; p += 64) off[p] = int16_t(i64(cols[p]) - r);
}

// Packs the padded rows of k_lf_compile<2> (row u at u * CP_P3) by the scanned counts.
__global__ void k_lf_compact(i64 n, const i64* __restrict__ m_outer, const i32* __restrict__ pcols,
                             const float* __restrict__ py, i32* __restrict__ m_cols, float* __restrict__ yrow) {
    const i64 u = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (u >= n) return;
    const i64 o = m_outer[u], len = m_outer[u + 1] - o;
    for (i64 k = threadIdx.x & 63; k < len; k += 64) {
        m_cols[o + k] = pcols[u * CP_P3 + k];
        yrow[o + k] = py[u * CP_P3 + k];
    }
}

// Builds M for the weights w; leaves m_ready false when a node's sets exceed the LDS capacities.
void learned_compile(LearnedGraph& g, const float* w, hipStream_t s, bool ctx_row_patterns) {
    const auto t0 = std::chrono::steady_clock::now();
    g.m_ready = false;
    const i64 n = g.n;
    const i64 need_words = (6 * g.band + 1 + 31) / 32 + 1;
    if (need_words > CP_WORDS) return;
    const int words = int((need_words + CP_THREADS - 1) / CP_THREADS * CP_THREADS);
    DevBuf<CompileBasis> basis{1};
    DevBuf<i64> cnt{size_t(n + 1)};
    DevBuf<int> overflow{1};
    hipLaunchKernelGGL(k_lf_basis, dim3(1), dim3(MP_H), 0, s, w, basis.p);
    NLF_HIP(hipMemsetAsync(overflow.p, 0, sizeof(int), s));
    NLF_HIP(hipMemsetAsync(cnt.p, 0, size_t(n + 1) * sizeof(i64), s));
    const float4* feat = reinterpret_cast<const float4*>(g.feat.p);
    // dynamic LDS: the bitmap, or the contracted basis of up to ell_d 1-hop entries after it
    const int cdrows = std::min(g.ell_d, CP_P1);
    const size_t lds = size_t(std::max(words, 2 * cdrows * MP_H)) * sizeof(unsigned);
    // one pass into rows padded to CP_P3 entries (the capacity of a row's three-hop set), then the
    // rows packed by the scanned lengths: the unions are formed once instead of once to count and once
    // to fill
    DevBuf<i32> pcols{size_t(n) * CP_P3};
    DevBuf<float> py{size_t(n) * CP_P3};
    hipLaunchKernelGGL(k_lf_compile<2>, dim3(unsigned(n)), dim3(CP_THREADS), lds, s, n, g.outer.p, g.cols.p, feat, g.band,
                       words, basis.p, w, cnt.p, (const i64*)nullptr, pcols.p, py.p, overflow.p, cdrows);
    NLF_HIP(hipGetLastError());
    int ovf = 0;
    download(&ovf, overflow.p, 1, s);
    if (ovf) return;
    exclusive_scan_i64(cnt.p, n, s);
    i64 nnz = 0;
    download(&nnz, cnt.p + n, 1, s);
    g.m_outer = std::move(cnt);
    g.m_nnz = nnz;
    g.m_cols.alloc(size_t(std::max<i64>(nnz, 1)));
    g.m_val.alloc(size_t(std::max<i64>(nnz, 1)));
    DevBuf<float> yrow{size_t(std::max<i64>(nnz, 1))};
    hipLaunchKernelGGL(k_lf_compact, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, g.m_outer.p, pcols.p, py.p, g.m_cols.p,
                       yrow.p);
    hipLaunchKernelGGL(k_lf_symmetrize, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, g.m_outer.p, g.m_cols.p, yrow.p, g.s.p,
                       g.m_val.p);
    g.m_off16.release();
    if (3 * g.band < 32768) {
        g.m_off16.alloc(size_t(std::max<i64>(nnz, 1)));
        hipLaunchKernelGGL(k_lf_off16, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, g.m_outer.p, g.m_cols.p, g.m_off16.p);
        NLF_HIP(hipGetLastError());
        if (ctx_row_patterns) build_row_patterns(n, nnz, g.m_outer.p, g.m_off16.p, g.m_pat, g.m_pstart, g.m_pat_size,
                                                 g.m_pat_groups, s);
    }
    NLF_HIP(hipGetLastError());
    NLF_HIP(hipStreamSynchronize(s));
    g.m_ready = true;
    g.compile_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ELL width for a longest row D: the smallest instantiated DM >= D
int ell_width(i64 D) { return D <= 9 ? 9 : D <= 16 ? 16 : D <= 24 ? 24 : 32; }

template<int DM, class Z>
void learned_apply_ell(LearnedGraph& g, const float* w, const double* r, Z* z, hipStream_t s) {
    const uint32_t n = uint32_t(g.n);
    const unsigned nb = blocks_for(g.n, SOA_THREADS);
    auto layer = [&](int l) { return w + P_LAYER0 + l * P_LAYER_SIZE; };
    static_assert(MP_L == 3, "the inference schedule below is written for three layers");
    float4* A = reinterpret_cast<float4*>(g.h[0].p);
    float4* B = reinterpret_cast<float4*>(g.h[1].p);
    float4* Dv = reinterpret_cast<float4*>(g.d.p);
    const i32* ec = g.ell_col.p;
    const float4* ef = g.ell_f.p;
    hipLaunchKernelGGL(k_soa_prep, dim3(1), dim3(64), 0, s, layer(2), w + P_WOUT, g.v.p);
    hipLaunchKernelGGL((k_soa_layer<DM, true>), dim3(nb), dim3(SOA_THREADS), 0, s, n, ec, ef, layer(0), w + P_AIN, g.x.p,
                       (const float4*)nullptr, A);
    hipLaunchKernelGGL((k_soa_layer<DM, false>), dim3(nb), dim3(SOA_THREADS), 0, s, n, ec, ef, layer(1), w + P_AIN, g.x.p,
                       A, B);
    hipLaunchKernelGGL(k_soa_last<DM>, dim3(nb), dim3(SOA_THREADS), 0, s, n, ec, ef, layer(2), g.v.p, B, g.y.p);
    hipLaunchKernelGGL(k_soa_tfirst<DM>, dim3(nb), dim3(SOA_THREADS), 0, s, n, ec, ef, layer(2), g.v.p, layer(1), g.x.p,
                       Dv, A);
    hipLaunchKernelGGL(k_soa_tmid<DM>, dim3(nb), dim3(SOA_THREADS), 0, s, n, ec, ef, layer(1), layer(0), Dv, A, B);
    hipLaunchKernelGGL((k_soa_tlast<DM, Z>), dim3(nb), dim3(SOA_THREADS), 0, s, n, ec, ef, layer(0), w + P_AIN, Dv, B,
                       g.y.p, g.s.p, g.sigma.p, r, z);
}

// z = s^2 r: the fixed first term of the network's output alone, i.e. Jacobi with the learned
// correction switched off (the comparison that shows what the trained network adds)
template<class Z>
__global__ void k_lf_jacobi(i64 n, const float* __restrict__ s, const double* __restrict__ r, Z* __restrict__ z) {
    for (i64 v = i64(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += i64(gridDim.x) * blockDim.x) {
        const double sv = double(s[v]);
        z[v] = Z(sv * sv * r[v]);
    }
}

// Lanes per row of the compiled operator's SpMV (0: the same rule as for the matrix, which gives 4
// for its 37 nonzeros a row; measured per apply at 1000 x 1000: 1 lane 0.152 ms, 2 lanes 0.133,
// 4 lanes 0.134, 8 lanes 0.171, 16 lanes 0.234, the coarse correction included).
#ifndef NLF_M_LPR
#define NLF_M_LPR 0
#endif

// z = M r, the symmetric part of N, on the lane-per-node kernels: forward layers (A, B channel-major
// ping-pong), y; transposed layers (d in place, q through A then B) ending in the output.
template<class Z>
void learned_apply_one(LearnedGraph& g, const float* w, const double* r, Z* z, hipStream_t s) {
    if (g.jacobi_apply) {
        hipLaunchKernelGGL(k_lf_jacobi<Z>, dim3(grid_stride_blocks(g.n)), dim3(256), 0, s, g.n, g.s.p, r, z);
        return;
    }
    if (g.m_ready && !g.network_apply) {
        const int lpr = NLF_M_LPR ? NLF_M_LPR : pick_lpr(double(g.m_nnz) / double(std::max<i64>(g.n, 1)), g.m_off16.p != nullptr);
        const SpmvCols cs = g.m_pstart.p ? SpmvCols{g.m_cols.p, g.m_pat.p, g.m_pstart.p}
                                         : SpmvCols{g.m_cols.p, g.m_off16.p, nullptr};
        launch_spmv<float, double, double, false>(lpr, g.n, g.m_outer.p, cs, g.m_val.p, r, z, nullptr,
                                                  nullptr, s);
        return;
    }
    learned_scaled_input(g, w, r, s);
    switch (g.ell_d) {
        case 9: learned_apply_ell<9, Z>(g, w, r, z, s); break;
        case 16: learned_apply_ell<16, Z>(g, w, r, z, s); break;
        case 24: learned_apply_ell<24, Z>(g, w, r, z, s); break;
        default: learned_apply_ell<32, Z>(g, w, r, z, s); break;
    }
}

// z = B r: M r, plus the coarse correction P A_c^-1 P^T r when the two-level form is on.
template<class Z>
void learned_apply(LearnedGraph& g, const float* w, const double* r, Z* z, hipStream_t s) {
    learned_apply_one<Z>(g, w, r, z, s);
    if constexpr (sizeof(Z) == sizeof(double))
        if (g.coarse) coarse_apply_add(*g.coarse, r, z, s);
}

// Loss ||N(r) - e||_A^2 / ||e||_A^2 and its gradient w.r.t. the weights (into g.grad).
double learned_loss_grad(LearnedGraph& g, const float* w, nlf_matrix A, const double* r, const double* e, hipStream_t s) {
    learned_alloc_train(g);
    const i64 n = g.n;
    learned_forward_train(g, w, r, g.zbuf.p, s);
    const int lpr = pick_lpr(A);
    launch_spmv<double, double, double, false>(lpr, n, A->outer.p, spmv_cols(A), A->val.p, g.zbuf.p, g.az.p, nullptr,
                                               nullptr, s);
    const unsigned vb = grid_stride_blocks(n);
    DevBuf<double> p1{vb}, p2{vb}, sums{2}, gz{size_t(n)};
    hipLaunchKernelGGL(k_lf_loss, dim3(vb), dim3(256), 0, s, n, g.zbuf.p, e, g.az.p, r, gz.p, p1.p, p2.p);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, p1.p, int(vb), sums.p);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, p2.p, int(vb), sums.p + 1);
    hipLaunchKernelGGL(k_scale_vec, dim3(vb), dim3(256), 0, s, n, gz.p, sums.p);
    // backward (bias slots stay zero: the partial rows are cleared and never written there)
    hipLaunchKernelGGL(k_fill_f32, dim3(256), dim3(256), 0, s, i64(MP_GRAD_BLOCKS) * NPARAM, g.gpart.p, 0.0f);
    const unsigned gb = MP_GRAD_BLOCKS;
    hipLaunchKernelGGL(k_lf_out_back, dim3(gb), dim3(MP_THREADS), 0, s, n, gz.p, g.s.p, g.sigma.p, g.h[MP_L].p, w, g.g0.p,
                       g.gpart.p);
    const float4* feat = reinterpret_cast<const float4*>(g.feat.p);
    float* gcur = g.g0.p;
    float* gnext = g.g1.p;
    for (int l = MP_L - 1; l >= 0; --l) {
        const float* wl = w + P_LAYER0 + l * P_LAYER_SIZE;
        float* gp = g.gpart.p + P_LAYER0 + l * P_LAYER_SIZE;
        hipLaunchKernelGGL(k_lf_layer_back1, dim3(gb), dim3(MP_THREADS), 0, s, n, gcur, g.h[l].p, g.m[l].p, wl, g.gdir.p,
                           g.gm.p, gp);
        hipLaunchKernelGGL(k_lf_layer_back2, dim3(gb), dim3(MP_THREADS), 0, s, n, g.outer.p, g.cols.p, feat, wl, g.gm.p,
                           g.h[l].p, g.gdir.p, gnext, gp);
        std::swap(gcur, gnext);
    }
    hipLaunchKernelGGL(k_lf_input_back, dim3(gb), dim3(MP_THREADS), 0, s, n, gcur, g.x.p, g.gpart.p);
    hipLaunchKernelGGL(k_lf_reduce, dim3((NPARAM + 255) / 256), dim3(256), 0, s, g.gpart.p, int(gb), g.grad.p);
    NLF_HIP(hipGetLastError());
    double h[2];
    download(h, sums.p, 2, s);
    return h[0] / h[1];
}

}  // namespace nlf

struct nlf_trainer_s {
    nlf_ctx ctx = nullptr;
    nlf::DevBuf<float> w, m1, m2;
    int t = 0;
};

namespace nlf {

LearnedGraph* learned_of(nlf_precond p) {
    NLF_CHECK(p && p->kind == NLF_PRECOND_LEARNED && p->learned, NLF_ERR_ARGUMENT, "not a learned preconditioner");
    return static_cast<LearnedGraph*>(p->learned);
}

// Flexible CG (Notay 2000, FCG(1)): beta = -(z.q_prev)/(p_prev.q_prev) keeps the directions
// A-orthogonal for a preconditioner that changes with every application (the network is
// nonlinear).  FP64 vectors; the SpMV reads FP64 or FP32 matrix values.  The FP64 true residual is
// recomputed at every tenfold drop of the recursive residual (reliable update) and is the only
// stopping test.  Stagnation (no 10% decrease of the residual over `window` iterations, or a
// non-finite value) returns -1 so the caller can fall back.
// Stagnation window: the 2-norm of a CG residual is not monotone, so a plateau must be long before
// it counts as stagnation.
constexpr int FCG_WINDOW = 1000;

struct FCGState {
    double pr, pq, zq, rr, alpha, beta;
};

__global__ void k_fcg_alpha(const double* __restrict__ ppr, const double* __restrict__ ppq, int n, FCGState* __restrict__ st) {
    __shared__ double sh[1024 / 64];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) { a += ppr[i]; b += ppq[i]; }
    a = block_sum<double>(a, sh);
    b = block_sum<double>(b, sh);
    if (threadIdx.x == 0) {
        // a non-finite step is never applied, so x keeps the last finite iterate for the fallback
        const double al = a / b;
        st->pr = a;
        st->pq = b;
        st->alpha = (b != 0.0 && isfinite(al)) ? al : 0.0;
    }
}

__global__ __launch_bounds__(256) void k_fcg_update(i64 n, double* __restrict__ x, double* __restrict__ r,
                                                    const double* __restrict__ p, const double* __restrict__ q,
                                                    const FCGState* __restrict__ st, double* __restrict__ prr) {
    __shared__ double sh[256 / 64];
    const double a = st->alpha;
    double acc = 0.0;
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) {
        double ri = r[i];
        if (a != 0.0) {             // alpha = 0 marks a rejected (non-finite) step: p may hold NaN
            x[i] += a * p[i];
            ri -= a * q[i];
            r[i] = ri;
        }
        acc += ri * ri;
    }
    acc = block_sum<double>(acc, sh);
    if (threadIdx.x == 0) prr[blockIdx.x] = acc;
}

__global__ void k_fcg_beta(const double* __restrict__ pzq, int n, const double* __restrict__ prr, int nr,
                           FCGState* __restrict__ st) {
    __shared__ double sh[1024 / 64];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a += pzq[i];
    for (int i = threadIdx.x; i < nr; i += blockDim.x) b += prr[i];
    a = block_sum<double>(a, sh);
    b = block_sum<double>(b, sh);
    if (threadIdx.x == 0) { st->zq = a; st->rr = b; st->beta = st->pq != 0.0 ? -a / st->pq : 0.0; }
}

__global__ void k_fcg_p(i64 n, double* __restrict__ p, const double* __restrict__ z, const FCGState* __restrict__ st) {
    const double b = st->beta;
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x) p[i] = z[i] + b * p[i];
}

// Per-block longest row and bandwidth max |col - row| of a CSR pattern (partials: 2 per block).
__global__ __launch_bounds__(256) void k_graph_stats(i64 n, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                     i64* __restrict__ part) {
    __shared__ i64 sd[256], sb[256];
    i64 deg = 0, band = 0;
    for (i64 u = i64(blockIdx.x) * blockDim.x + threadIdx.x; u < n; u += i64(gridDim.x) * blockDim.x) {
        const i64 b = outer[u], e = outer[u + 1];
        deg = std::max(deg, e - b);
        if (e > b) band = std::max(band, std::max(u - i64(cols[b]), i64(cols[e - 1]) - u));
    }
    sd[threadIdx.x] = deg;
    sb[threadIdx.x] = band;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            sd[threadIdx.x] = std::max(sd[threadIdx.x], sd[threadIdx.x + w]);
            sb[threadIdx.x] = std::max(sb[threadIdx.x], sb[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = sd[0]; part[2 * blockIdx.x + 1] = sb[0]; }
}

template<class V>
i64 fcg(hipStream_t s, nlf_matrix A, const V* val, nlf_precond M, const double* b, double* x, double tol, i64 max_it,
        double* true_res, i64* updates, int window, KernelTimer& timer, const int16_t* lo = nullptr) {
    const i64 n = A->rows;
    LearnedGraph& g = *learned_of(M);
    // FP64 values: the products read the upper triangle once (the symmetric storage), when it applies
    bool sym = false;
    if constexpr (std::is_same_v<V, double>) sym = !lo && symv_prepare(A, s);
    // the lane count (and with it the partial-sum grid) the full-storage layout would take; the symmetric
    // storage applies only where every offset is below 2^15, as the 16-bit one does, so it is known
    // without building that layout
    const int lpr = sym ? pick_lpr(double(A->nnz_full) / double(std::max<i64>(n, 1)), true) : pick_lpr(A);
    const unsigned sb = spmv_blocks(lpr, n), vb = grid_stride_blocks(n), rb = spmv_blocks(64, n);
    const size_t np = std::max<size_t>(std::max<size_t>(sb, vb), rb);
    DevBuf<double> r{size_t(n)}, z{size_t(n)}, p{size_t(n)}, q{size_t(n)}, pa{np}, pb{np}, sums{2};
    DevBuf<FCGState> st{1};
    auto true_residual = [&]() {
        residual64(A, x, b, r.p, pa.p, pb.p, rb, s);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, pa.p, int(rb), sums.p);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, pb.p, int(rb), sums.p + 1);
        double h[2];
        download(h, sums.p, 2, s);
        return std::make_pair(std::sqrt(h[0]), std::sqrt(h[1]));
    };
    auto spmv_pq = [&](bool sample) {             // q = A p, partial p.q into pb; partial p.r into pa
        if (sample) timer.start(0, s);
        if (sym) launch_symv<true>(A, p.p, q.p, pb.p, nullptr, s, sb);
        else launch_spmv<V, double, double, true>(lpr, n, A->outer.p, spmv_cols(A), val, p.p, q.p, pb.p, nullptr, s, false, lo);
        if (sample) timer.stop(0, s);
        hipLaunchKernelGGL(k_dot<double>, dim3(sb), dim3(SPMV_THREADS), 0, s, n, p.p, r.p, pa.p, (const CGState*)nullptr);
    };
    auto apply_m = [&](bool sample) {
        if (sample) timer.start(1, s);
        learned_apply<double>(g, g.w.p, r.p, z.p, s);
        if (sample) timer.stop(1, s);
    };
    auto [rn, bn] = true_residual();
    if (bn == 0.0) bn = 1.0;
    double rel = rn / bn;
    i64 it = 0, nupd = 0;
    double last_update = rn, best = rn, best_at = 0;
    if (rel > tol && max_it > 0) {
        apply_m(false);
        NLF_HIP(hipMemcpyAsync(p.p, z.p, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
        spmv_pq(false);
        // the k_dot for p.r wrote sb partials into pa; the SpMV wrote sb partials of p.q into pb
        FCGState hs{};
        for (;;) {
            const bool sample = (it % 16) == 0;
            hipLaunchKernelGGL(k_fcg_alpha, dim3(1), dim3(1024), 0, s, pa.p, pb.p, int(sb), st.p);
            hipLaunchKernelGGL(k_fcg_update, dim3(vb), dim3(256), 0, s, n, x, r.p, p.p, q.p, st.p, pa.p);
            apply_m(sample);
            hipLaunchKernelGGL(k_dot<double>, dim3(vb), dim3(SPMV_THREADS), 0, s, n, z.p, q.p, pb.p, (const CGState*)nullptr);
            hipLaunchKernelGGL(k_fcg_beta, dim3(1), dim3(1024), 0, s, pb.p, int(vb), pa.p, int(vb), st.p);
            hipLaunchKernelGGL(k_fcg_p, dim3(vb), dim3(256), 0, s, n, p.p, z.p, st.p);
            ++it;
            bool check = (it % 8) == 0 || it >= max_it;
            if (check) {
                NLF_HIP(hipMemcpyAsync(&hs, st.p, sizeof(FCGState), hipMemcpyDeviceToHost, s));
                NLF_HIP(hipStreamSynchronize(s));
                const double est = std::sqrt(std::fabs(hs.rr));
                if (!std::isfinite(est) || !std::isfinite(hs.beta) || !std::isfinite(hs.pq)) { it = -it; break; }
                if (est < 0.9 * best) { best = est; best_at = double(it); }
                if (double(it) - best_at > window) { it = -it; break; }         // stagnation
                // the recursive residual is replaced by the FP64 true residual at every tenfold drop
                // when the SpMV reads reduced-precision values, every thousandfold drop with FP64
                // values (where it drifts far less), and always before stopping
                const double refresh = sizeof(V) == sizeof(double) ? 1e-3 : 0.1;
                if (est < refresh * last_update || est <= tol * bn || it >= max_it) {
                    auto tr = true_residual();
                    ++nupd;
                    rel = tr.first / bn;
                    last_update = tr.first;
                    if (rel <= tol || it >= max_it) break;
                }
            }
            spmv_pq(sample);
        }
    }
    if (true_res) *true_res = rel;
    if (updates) *updates = nupd;
    timer.collect();
    return it;
}

}  // namespace nlf

using namespace nlf;

namespace {

// splitmix64 (Steele, Lea, Flood 2014): deterministic weight initialisation from a seed
uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

float uniform(uint64_t& st, float a) { return a * (2.0f * float(splitmix(st) >> 40) / float(1ull << 24) - 1.0f); }

void destroy_graph(void* p) { delete static_cast<LearnedGraph*>(p); }

}  // namespace

extern "C" {

int nlf_learned_num_params(void) { return NPARAM; }

// Deterministic initial weights (splitmix64); output weights and the unused bias slots start at zero,
// so the initial preconditioner is Jacobi.
int nlf_learned_init_weights(uint64_t seed, float* w) {
    if (!w) return NLF_ERR_ARGUMENT;
    uint64_t st = seed * 0x2545f4914f6cdd1dull + 1;
    for (int i = 0; i < NPARAM; ++i) w[i] = 0.0f;
    for (int c = 0; c < MP_H; ++c) w[P_AIN + c] = uniform(st, 1.0f);
    const float a = 0.5f * std::sqrt(6.0f / float(2 * MP_H));   // half the Glorot bound: linear layers compound
    for (int l = 0; l < MP_L; ++l) {
        float* L = w + P_LAYER0 + l * P_LAYER_SIZE;
        for (int i = 0; i < MP_H * MP_F; ++i) L[P_WE + i] = uniform(st, 0.5f);
        for (int i = 0; i < MP_H * MP_H; ++i) L[P_WS + i] = uniform(st, a);
        for (int i = 0; i < MP_H * MP_H; ++i) L[P_WM + i] = uniform(st, a);
    }
    return NLF_OK;
}

// Learned preconditioner on the graph of the local matrix A_local, edge features from A_local, A
// and the influence functions of `models` (one per element group).  Weights start at zero
// (Jacobi); load trained ones with nlf_learned_set_weights.
int nlf_learned_create(nlf_matrix A, nlf_matrix A_local, const nlf_model* models, nlf_precond* out) {
    if (!A || !A_local || !out || !models) return NLF_ERR_ARGUMENT;
    return guarded(A->ctx, [&] {
        NLF_CHECK(A->dof == 1 && A_local->dof == 1 && A->rows == A_local->rows, NLF_ERR_UNSUPPORTED,
                  "the learned preconditioner is implemented for the thermal (1 DoF) problem");
        NLF_CHECK(!A->neumann && !A_local->neumann, NLF_ERR_UNSUPPORTED,
                  "the learned preconditioner works on mesh nodes: not on the bordered Neumann system");
        hipStream_t s = A->ctx->stream;
        const nlf_mesh m = A->mesh;
        auto P = std::make_unique<nlf_precond_s>();
        P->ctx = A->ctx;
        P->kind = NLF_PRECOND_LEARNED;
        P->rows = A->rows;
        auto g = std::make_unique<LearnedGraph>();
        g->n = A->rows;
        g->nnz = A_local->nnz_full;
        // the graph is the local pattern, copied on the device; its longest row and bandwidth are
        // reduced on the device
        g->outer.alloc(size_t(g->n + 1));
        g->cols.alloc(size_t(std::max<i64>(g->nnz, 1)));
        NLF_HIP(hipMemcpyAsync(g->outer.p, A_local->outer.p, g->outer.bytes(), hipMemcpyDeviceToDevice, s));
        NLF_HIP(hipMemcpyAsync(g->cols.p, A_local->cols.p, size_t(g->nnz) * sizeof(i32), hipMemcpyDeviceToDevice, s));
        constexpr int GS_BLOCKS = 1024;
        DevBuf<i64> d_stats{size_t(2 * GS_BLOCKS)};
        hipLaunchKernelGGL(k_graph_stats, dim3(GS_BLOCKS), dim3(256), 0, s, g->n, g->outer.p, g->cols.p, d_stats.p);
        std::vector<i64> h_stats(size_t(2 * GS_BLOCKS));
        download(h_stats.data(), d_stats.p, h_stats.size(), s);
        i64 maxdeg = 0;
        for (int b = 0; b < GS_BLOCKS; ++b) {
            maxdeg = std::max(maxdeg, h_stats[size_t(2 * b)]);
            g->band = std::max(g->band, h_stats[size_t(2 * b + 1)]);
        }
        NLF_CHECK(maxdeg <= MP_MAXDEG, NLF_ERR_UNSUPPORTED, "local pattern has a row longer than the staged edge list");
        g->s.alloc(size_t(g->n));
        g->feat.alloc(size_t(4 * g->nnz));
        g->w.alloc(NPARAM);
        g->x.alloc(size_t(g->n));
        g->y.alloc(size_t(g->n));
        g->h[0].alloc(size_t(g->n) * MP_H);
        g->h[1].alloc(size_t(g->n) * MP_H);
        g->d.alloc(size_t(g->n) * MP_H);
        g->part.alloc(grid_stride_blocks(g->n));
        g->sigma.alloc(1);
        std::vector<GroupParams> gps(m->G);
        for (int k = 0; k < m->G; ++k) gps[k] = model_group_params(models[k], false);
        DevBuf<GroupParams> d_groups;
        upload(d_groups, gps.data(), gps.size(), s);
        const i64 n = g->n;
        hipLaunchKernelGGL(k_lf_scale, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, A->diag.p, A->val.p, g->s.p);
        hipLaunchKernelGGL(k_lf_features, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, g->outer.p, g->cols.p,
                           A_local->val.p, A->outer.p, A->cols.p, A->val.p, g->s.p, m->xy.p, m->ne_off.p, m->ne_ids.p,
                           m->group_of.p, d_groups.p, g->feat.p);
        hipLaunchKernelGGL(k_fill_f32, dim3(64), dim3(256), 0, s, i64(NPARAM), g->w.p, 0.0f);
        NLF_CHECK(n * MP_H < (i64(1) << 31), NLF_ERR_UNSUPPORTED, "too many nodes for 32-bit activation offsets");
        g->ell_d = ell_width(maxdeg);
        g->ell_col.alloc(size_t(g->ell_d) * size_t(n));
        g->ell_f.alloc(size_t(g->ell_d) * size_t(n));
        g->v.alloc(2 * MP_H);
        hipLaunchKernelGGL(k_soa_ell, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, g->ell_d, g->outer.p, g->cols.p,
                           reinterpret_cast<const float4*>(g->feat.p), g->ell_col.p, g->ell_f.p);
        NLF_HIP(hipGetLastError());
        NLF_HIP(hipStreamSynchronize(s));
        P->learned = g.release();
        P->learned_destroy = destroy_graph;
        *out = P.release();
    });
}

int nlf_learned_set_weights(nlf_precond p, const float* w) {
    if (!p || !w) return NLF_ERR_ARGUMENT;
    return guarded(p->ctx, [&] {
        LearnedGraph& g = *learned_of(p);
        NLF_HIP(hipMemcpyAsync(g.w.p, w, NPARAM * sizeof(float), hipMemcpyHostToDevice, p->ctx->stream));
        learned_compile(g, g.w.p, p->ctx->stream, p->ctx->row_patterns != 0);
        NLF_HIP(hipStreamSynchronize(p->ctx->stream));
    });
}

// Whether the operator is compiled to its explicit sparse form, its nonzeros and compile time.
int nlf_learned_info(nlf_precond p, int* compiled, int64_t* nnz, double* compile_seconds) {
    if (!p) return NLF_ERR_ARGUMENT;
    return guarded(p->ctx, [&] {
        LearnedGraph& g = *learned_of(p);
        if (compiled) *compiled = g.m_ready ? 1 : 0;
        if (nnz) *nnz = g.m_ready ? g.m_nnz : 0;
        if (compile_seconds) *compile_seconds = g.compile_seconds;
    });
}

// Two-level form: aggregates > 0 builds the coarse correction P A_c^-1 P^T from the system matrix
// a (about that many aggregates: a grid of cells over the nodes), 0 removes it.
int nlf_learned_set_coarse(nlf_precond p, nlf_matrix a, int64_t aggregates) {
    if (!p) return NLF_ERR_ARGUMENT;
    return guarded(p->ctx, [&] {
        LearnedGraph& g = *learned_of(p);
        g.coarse.reset();
        if (aggregates <= 0) return;
        NLF_CHECK(a && a->rows == g.n, NLF_ERR_ARGUMENT, "the coarse space needs the system matrix of the preconditioner");
        g.coarse = build_coarse(a, aggregates, p->ctx->stream);
    });
}

int nlf_learned_coarse_info(nlf_precond p, int64_t* coarse_rows, double* setup_seconds) {
    if (!p) return NLF_ERR_ARGUMENT;
    return guarded(p->ctx, [&] {
        LearnedGraph& g = *learned_of(p);
        if (coarse_rows) *coarse_rows = g.coarse ? g.coarse->nc : 0;
        if (setup_seconds) *setup_seconds = g.coarse ? g.coarse->setup_seconds : 0.0;
    });
}

// Application path: 0 = the compiled sparse operator when available (default), 1 = the network
// kernels, 2 = the network's Jacobi term s^2 r alone (learned correction off).
int nlf_learned_set_apply(nlf_precond p, int mode) {
    if (!p || mode < 0 || mode > 2) return NLF_ERR_ARGUMENT;
    return guarded(p->ctx, [&] {
        LearnedGraph& g = *learned_of(p);
        g.network_apply = mode == 1;
        g.jacobi_apply = mode == 2;
    });
}

int nlf_learned_get_weights(nlf_precond p, float* w) {
    if (!p || !w) return NLF_ERR_ARGUMENT;
    return guarded(p->ctx, [&] { download(w, learned_of(p)->w.p, NPARAM, p->ctx->stream); });
}

// Host export of the graph (local pattern, edge features nnz x 4, node scaling s).
int nlf_learned_export_graph(nlf_precond p, int64_t* outer, int32_t* cols, float* feat, float* s) {
    if (!p) return NLF_ERR_ARGUMENT;
    return guarded(p->ctx, [&] {
        LearnedGraph& g = *learned_of(p);
        hipStream_t st = p->ctx->stream;
        if (outer) download(outer, g.outer.p, size_t(g.n + 1), st);
        if (cols) download(cols, g.cols.p, size_t(g.nnz), st);
        if (feat) download(feat, g.feat.p, size_t(4 * g.nnz), st);
        if (s) download(s, g.s.p, size_t(g.n), st);
    });
}

// Loss ||M(r) - e||_A^2 / ||e||_A^2 and its gradient (host, NPARAM floats) for the current weights.
int nlf_learned_loss_grad(nlf_precond p, nlf_matrix A, const double* r, const double* e, double* loss, float* grad) {
    if (!p || !A) return NLF_ERR_ARGUMENT;
    return guarded(p->ctx, [&] {
        LearnedGraph& g = *learned_of(p);
        const double L = learned_loss_grad(g, g.w.p, A, r, e, p->ctx->stream);
        if (loss) *loss = L;
        if (grad) download(grad, g.grad.p, NPARAM, p->ctx->stream);
    });
}

int nlf_trainer_create(nlf_ctx ctx, const float* w, nlf_trainer* out) {
    if (!ctx || !w || !out) return NLF_ERR_ARGUMENT;
    return guarded(ctx, [&] {
        auto t = std::make_unique<nlf_trainer_s>();
        t->ctx = ctx;
        upload(t->w, w, NPARAM, ctx->stream);
        t->m1.alloc(NPARAM);
        t->m2.alloc(NPARAM);
        hipLaunchKernelGGL(k_fill_f32, dim3(64), dim3(256), 0, ctx->stream, i64(NPARAM), t->m1.p, 0.0f);
        hipLaunchKernelGGL(k_fill_f32, dim3(64), dim3(256), 0, ctx->stream, i64(NPARAM), t->m2.p, 0.0f);
        NLF_HIP(hipStreamSynchronize(ctx->stream));
        *out = t.release();
    });
}

int nlf_trainer_destroy(nlf_trainer t) {
    if (!t) return NLF_OK;
    CtxScope scope(live_ctx(t->ctx));
    delete t;
    return NLF_OK;
}

// One Adam step on one (residual, error) pair of the problem whose graph is `graph` (matrix A).
int nlf_trainer_step(nlf_trainer t, nlf_precond graph, nlf_matrix A, const double* r, const double* e, float lr,
                     double* loss) {
    if (!t || !graph || !A) return NLF_ERR_ARGUMENT;
    return guarded(t->ctx, [&] {
        LearnedGraph& g = *learned_of(graph);
        hipStream_t s = t->ctx->stream;
        const double L = learned_loss_grad(g, t->w.p, A, r, e, s);
        t->t += 1;
        hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, s, t->w.p, g.grad.p, t->m1.p, t->m2.p, lr, 0.9f,
                           0.999f, 1e-8f, t->t);
        NLF_HIP(hipGetLastError());
        if (loss) *loss = L;
    });
}

int nlf_trainer_get_weights(nlf_trainer t, float* w) {
    if (!t || !w) return NLF_ERR_ARGUMENT;
    return guarded(t->ctx, [&] { download(w, t->w.p, NPARAM, t->ctx->stream); });
}

// Flexible CG with the learned preconditioner M (taken first).  If FCG stagnates or produces a
// non-finite value, the solve continues from the current iterate with standard PCG using
// `fallback` (ILLT or Jacobi; NULL = Jacobi-free identity).  mode NLF_MODE_MIXED reads FP32
// matrix values in the Krylov SpMV (nlf_matrix_make_fp32 first), NLF_MODE_SPLIT FP32 + 16-bit
// remainders (nlf_matrix_make_split first); the FP64 true residual decides in every mode.
int nlf_fcg_solve(nlf_matrix A, nlf_precond M, nlf_precond fallback, const double* b, double* x, double tol,
                  int64_t max_it, int mode, nlf_solve_stats* stats) {
    if (!A || !M) return NLF_ERR_ARGUMENT;
    return guarded(A->ctx, [&] {
        NLF_CHECK(b && x, NLF_ERR_ARGUMENT, "null vector");
        hipStream_t s = A->ctx->stream;
        const auto t0 = std::chrono::steady_clock::now();
        nlf_solve_stats out{};
        out.preconditioner = NLF_PRECOND_LEARNED;
        KernelTimer timer;
        i64 it = 0;
        if (mode == NLF_MODE_MIXED) {
            NLF_CHECK(A->val32.p, NLF_ERR_STATE, "mixed mode needs nlf_matrix_make_fp32 first");
            it = fcg<float>(s, A, A->val32.p, M, b, x, tol, max_it, &out.true_residual, &out.outer_iterations, FCG_WINDOW, timer);
        } else if (mode == NLF_MODE_SPLIT) {
            NLF_CHECK(A->split_ready, NLF_ERR_STATE, "split mode needs nlf_matrix_make_split first");
            it = fcg<float>(s, A, A->val32.p, M, b, x, tol, max_it, &out.true_residual, &out.outer_iterations, FCG_WINDOW, timer,
                            A->val_lo.p);
        } else {
            NLF_CHECK(mode == NLF_MODE_FP64, NLF_ERR_ARGUMENT, "unknown mode");
            it = fcg<double>(s, A, A->val.p, M, b, x, tol, max_it, &out.true_residual, &out.outer_iterations, FCG_WINDOW, timer);
        }
        out.spmv_ms = timer.avg(0);
        out.precond_ms = timer.avg(1);
        out.spmv_samples = timer.count[0];
        out.spmv_bytes = spmv_stream_bytes(A, mode == NLF_MODE_MIXED ? 4 : mode == NLF_MODE_SPLIT ? 6 : 8, 8);
        if (it < 0) {
            // stagnation: fall back to PCG with the fallback preconditioner from the current iterate
            out.fallbacks = 1;
            out.preconditioner = fallback ? fallback->kind : NLF_PRECOND_IDENTITY;
            // PCG stops on sqrt(r'z)/||b|| (conjugate_gradient.hpp:48-79); the true residual decides, so
            // the PCG tolerance is tightened tenfold (at most three times) until the true residual meets tol
            int breakdown = 0;
            CG<double, double> cg(s, A->rows, A->nnz_full, A->outer.p, A->cols.p, A->val.p, fallback);
            cg.set_cols(A);
            const unsigned rb = spmv_blocks(64, A->rows);
            DevBuf<double> r{size_t(A->rows)}, prr{rb}, pbb{rb}, sums{2};
            i64 left = max_it + it, total = -it;
            double pcg_tol = tol;
            for (int round = 0; round < 4; ++round) {
                const i64 more = cg.solve(b, x, pcg_tol, left, &out.residual, &breakdown);
                total += more;
                left -= more;
                residual64(A, x, b, r.p, prr.p, pbb.p, rb, s);
                hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, prr.p, int(rb), sums.p);
                hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, pbb.p, int(rb), sums.p + 1);
                double h[2];
                download(h, sums.p, 2, s);
                out.true_residual = h[1] > 0 ? std::sqrt(h[0] / h[1]) : 0.0;
                if (out.true_residual <= tol || left <= 0 || breakdown) break;
                pcg_tol *= 0.1;
            }
            out.iterations = total;
        } else {
            out.iterations = it;
            out.residual = out.true_residual;
        }
        out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = out;
    });
}

}  // extern "C"
