// Learned preconditioner: a message-passing network over mesh nodes, its training kernels, and
// flexible CG.  Compiled as part of nlf_solver.hip (shares its Krylov kernels).
//
// Graph: the nodes of the thermal problem; edges = the pattern of the LOCAL matrix (element
// adjacency, self loops included).  Edge features (symmetric in u, v):
//   f0 = K_L(u,v) s_u s_v, f1 = K(u,v) s_u s_v, f2 = (1-p1) phi(x_u, x_v) / phi(0), f3 = [u == v],
// with s = 1/sqrt(diag K).  Network (H = 32, three layers, FP32):
//   x_u = s_u r_u / sigma,  sigma = rms(S r)
//   h0  = relu(a_in x + b_in)
//   m_u = sum_{v in N(u)} (W_e f_uv) (.) h_v ;  h' = relu(W_s h_u + W_m m_u + b)      (x3)
//   y_u = w_out . h3_u + b_out ;  z_u = s_u sigma (x_u + y_u)
// so z = D^-1 r (Jacobi) when the output weights are zero; the network learns the correction.
// Training minimises ||z(r) - e||_A^2 / ||e||_A^2 over (residual, error) pairs (A e = r).
// Gradients are reduced per block in a fixed order (deterministic) and applied with Adam.

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
    DevBuf<float> h[MP_L + 1], a[MP_L + 1], m[MP_L];
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
    const double phi0u = gu.nonlocal ? influence(gu.inf, xu, yu, xu, yu) : 1.0;
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
        const double wu = gu.nonlocal ? (1.0 - gu.p1) * influence(gu.inf, xu, yu, xv, yv) / phi0u : 0.0;
        const double phi0v = gv.nonlocal ? influence(gv.inf, xv, yv, xv, yv) : 1.0;
        const double wv = gv.nonlocal ? (1.0 - gv.p1) * influence(gv.inf, xv, yv, xu, yu) / phi0v : 0.0;
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
                                                         float* __restrict__ x, float* __restrict__ h0, float* __restrict__ a0) {
    const int c = threadIdx.x % MP_H;
    const float inv = float(1.0 / *sigma);
    for (i64 u = i64(blockIdx.x) * MP_NODES + threadIdx.x / MP_H; u < n; u += i64(gridDim.x) * MP_NODES) {
        const float xu = float(double(s[u]) * double(r[u])) * inv;
        const float av = w[P_AIN + c] * xu + w[P_BIN + c];
        h0[u * MP_H + c] = fmaxf(av, 0.0f);
        if (a0) a0[u * MP_H + c] = av;
        if (c == 0) x[u] = xu;
    }
}

// One message-passing layer; a (pre-activation) and m (messages) are stored when given (training).
__global__ __launch_bounds__(MP_THREADS) void k_lf_layer(i64 n, const i64* __restrict__ outer, const i32* __restrict__ cols,
                                                         const float4* __restrict__ feat, const float* __restrict__ wl,
                                                         const float* __restrict__ hin, float* __restrict__ hout,
                                                         float* __restrict__ aout, float* __restrict__ mout) {
    __shared__ float Ws[MP_H][MP_H + 1], Wm[MP_H][MP_H + 1], We[MP_H][MP_F], bb[MP_H];
    for (int i = threadIdx.x; i < MP_H * MP_H; i += blockDim.x) {
        Ws[i / MP_H][i % MP_H] = wl[P_WS + i];
        Wm[i / MP_H][i % MP_H] = wl[P_WM + i];
    }
    for (int i = threadIdx.x; i < MP_H * MP_F; i += blockDim.x) We[i / MP_F][i % MP_F] = wl[P_WE + i];
    if (threadIdx.x < MP_H) bb[threadIdx.x] = wl[P_B + threadIdx.x];
    __syncthreads();
    const int c = threadIdx.x % MP_H;
    const float e0 = We[c][0], e1 = We[c][1], e2 = We[c][2], e3 = We[c][3];
    // all lanes of a wave stay in the loop (both 32-lane halves) so the width-32 shuffles are uniform
    const i64 nn = (n + MP_NODES - 1) / MP_NODES * MP_NODES;
    for (i64 u0 = i64(blockIdx.x) * MP_NODES; u0 < nn; u0 += i64(gridDim.x) * MP_NODES) {
        const i64 u = u0 + threadIdx.x / MP_H;
        const bool live = u < n;
        float hu = 0.0f, m = 0.0f;
        if (live) {
            hu = hin[u * MP_H + c];
            for (i64 p = outer[u]; p < outer[u + 1]; ++p) {
                const float4 f = feat[p];
                const float g = e0 * f.x + e1 * f.y + e2 * f.z + e3 * f.w;
                m += g * hin[i64(cols[p]) * MP_H + c];
            }
        }
        float acc = bb[c];
#pragma unroll
        for (int k = 0; k < MP_H; ++k) {
            acc += Ws[c][k] * __shfl(hu, k, MP_H);
            acc += Wm[c][k] * __shfl(m, k, MP_H);
        }
        if (live) {
            hout[u * MP_H + c] = fmaxf(acc, 0.0f);
            if (aout) aout[u * MP_H + c] = acc;
            if (mout) mout[u * MP_H + c] = m;
        }
    }
}

// z_u = s_u sigma (x_u + w_out . h_u + b_out)
template<class Z>
__global__ __launch_bounds__(MP_THREADS) void k_lf_output(i64 n, const float* __restrict__ h, const float* __restrict__ w,
                                                          const float* __restrict__ x, const float* __restrict__ s,
                                                          const double* __restrict__ sigma, Z* __restrict__ z) {
    const int c = threadIdx.x % MP_H;
    const double sg = *sigma;
    const i64 nn = (n + MP_NODES - 1) / MP_NODES * MP_NODES;
    for (i64 u0 = i64(blockIdx.x) * MP_NODES; u0 < nn; u0 += i64(gridDim.x) * MP_NODES) {
        const i64 u = u0 + threadIdx.x / MP_H;
        float y = u < n ? w[P_WOUT + c] * h[u * MP_H + c] : 0.0f;
        for (int o = MP_H / 2; o > 0; o >>= 1) y += __shfl_down(y, o, MP_H);
        if (c == 0 && u < n) z[u] = Z(double(s[u]) * sg * (double(x[u]) + double(y) + double(w[P_BOUT])));
    }
}

// ---- backward ------------------------------------------------------------------------------------
// dL/dz (double) -> dL/dh3; gradients of w_out, b_out into the block partial row
__global__ __launch_bounds__(MP_THREADS) void k_lf_out_back(i64 n, const double* __restrict__ gz, const float* __restrict__ s,
                                                            const double* __restrict__ sigma, const float* __restrict__ h,
                                                            const float* __restrict__ w, float* __restrict__ gh,
                                                            float* __restrict__ gpart) {
    __shared__ float red[MP_NODES][MP_H + 1];
    const int c = threadIdx.x % MP_H, slot = threadIdx.x / MP_H;
    const double sg = *sigma;
    float gw = 0.0f, gb = 0.0f;
    for (i64 u = i64(blockIdx.x) * MP_NODES + slot; u < n; u += i64(gridDim.x) * MP_NODES) {
        const float dy = float(double(s[u]) * sg * gz[u]);
        gh[u * MP_H + c] = dy * w[P_WOUT + c];
        gw += dy * h[u * MP_H + c];
        gb += c == 0 ? dy : 0.0f;
    }
    float* row = gpart + i64(blockIdx.x) * NPARAM;
    red[slot][c] = gw;
    __syncthreads();
    if (slot == 0) {
        float t = 0.0f;
        for (int k = 0; k < MP_NODES; ++k) t += red[k][c];
        row[P_WOUT + c] = t;
    }
    __syncthreads();
    red[slot][c] = gb;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int k = 0; k < MP_NODES; ++k) t += red[k][0];
        row[P_BOUT] = t;
    }
}

// da = gh (.) [a > 0]; gradients of W_s, W_m, b; dL/dh_in (direct part) = W_s^T da, dL/dm = W_m^T da.
__global__ __launch_bounds__(MP_THREADS) void k_lf_layer_back1(i64 n, const float* __restrict__ gh, const float* __restrict__ a,
                                                               const float* __restrict__ hin, const float* __restrict__ m,
                                                               const float* __restrict__ wl, float* __restrict__ gdir,
                                                               float* __restrict__ gm, float* __restrict__ gpart_layer) {
    __shared__ float Ws[MP_H][MP_H + 1], Wm[MP_H][MP_H + 1];
    __shared__ float red[MP_NODES][MP_H + 1];
    for (int i = threadIdx.x; i < MP_H * MP_H; i += blockDim.x) {
        Ws[i / MP_H][i % MP_H] = wl[P_WS + i];
        Wm[i / MP_H][i % MP_H] = wl[P_WM + i];
    }
    __syncthreads();
    const int c = threadIdx.x % MP_H, slot = threadIdx.x / MP_H;
    float gws[MP_H], gwm[MP_H], gbias = 0.0f;
#pragma unroll
    for (int k = 0; k < MP_H; ++k) { gws[k] = 0.0f; gwm[k] = 0.0f; }
    const i64 nn = (n + MP_NODES - 1) / MP_NODES * MP_NODES;
    for (i64 u0 = i64(blockIdx.x) * MP_NODES; u0 < nn; u0 += i64(gridDim.x) * MP_NODES) {
        const i64 u = u0 + slot;
        const bool live = u < n;
        float da = 0.0f, hu = 0.0f, mu = 0.0f;
        if (live) {
            da = a[u * MP_H + c] > 0.0f ? gh[u * MP_H + c] : 0.0f;
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
        gbias += da;
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
    red[slot][c] = gbias;
    __syncthreads();
    if (slot == 0) {
        float t = 0.0f;
        for (int q = 0; q < MP_NODES; ++q) t += red[q][c];
        row[P_B + c] = t;
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

__global__ __launch_bounds__(MP_THREADS) void k_lf_input_back(i64 n, const float* __restrict__ gh, const float* __restrict__ a0,
                                                              const float* __restrict__ x, float* __restrict__ gpart) {
    __shared__ float red[MP_NODES][MP_H + 1];
    const int c = threadIdx.x % MP_H, slot = threadIdx.x / MP_H;
    float ga = 0.0f, gb = 0.0f;
    for (i64 u = i64(blockIdx.x) * MP_NODES + slot; u < n; u += i64(gridDim.x) * MP_NODES) {
        const float da = a0[u * MP_H + c] > 0.0f ? gh[u * MP_H + c] : 0.0f;
        ga += da * x[u];
        gb += da;
    }
    float* row = gpart + i64(blockIdx.x) * NPARAM;
    red[slot][c] = ga;
    __syncthreads();
    if (slot == 0) {
        float t = 0.0f;
        for (int q = 0; q < MP_NODES; ++q) t += red[q][c];
        row[P_AIN + c] = t;
    }
    __syncthreads();
    red[slot][c] = gb;
    __syncthreads();
    if (slot == 0) {
        float t = 0.0f;
        for (int q = 0; q < MP_NODES; ++q) t += red[q][c];
        row[P_BIN + c] = t;
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

unsigned mp_blocks(i64 n) { return unsigned(std::min<i64>(std::max<i64>((n + MP_NODES - 1) / MP_NODES, 1), 8192)); }

void learned_alloc_train(LearnedGraph& g) {
    if (g.train_ready) return;
    const size_t nh = size_t(g.n) * MP_H;
    for (int l = 0; l <= MP_L; ++l) { g.h[l].alloc(nh); g.a[l].alloc(nh); }
    for (int l = 0; l < MP_L; ++l) g.m[l].alloc(nh);
    g.g0.alloc(nh); g.g1.alloc(nh); g.gdir.alloc(nh); g.gm.alloc(nh);
    g.gpart.alloc(size_t(MP_GRAD_BLOCKS) * NPARAM);
    g.grad.alloc(NPARAM);
    g.zbuf.alloc(size_t(g.n));
    g.az.alloc(size_t(g.n));
    g.train_ready = true;
}

// z = M(r): inference (two ping-pong activation buffers) or training (all activations kept).
template<class R, class Z>
void learned_forward(LearnedGraph& g, const float* w, const R* r, Z* z, bool train, hipStream_t s) {
    const i64 n = g.n;
    const unsigned vb = grid_stride_blocks(n), mb = mp_blocks(n);
    hipLaunchKernelGGL(k_lf_rnorm<R>, dim3(vb), dim3(256), 0, s, n, r, g.s.p, g.part.p);
    hipLaunchKernelGGL(k_lf_sigma, dim3(1), dim3(1024), 0, s, g.part.p, int(vb), n, g.sigma.p);
    hipLaunchKernelGGL(k_lf_input<R>, dim3(mb), dim3(MP_THREADS), 0, s, n, r, g.s.p, g.sigma.p, w, g.x.p, g.h[0].p,
                       train ? g.a[0].p : (float*)nullptr);
    const float4* feat = reinterpret_cast<const float4*>(g.feat.p);
    for (int l = 0; l < MP_L; ++l) {
        const float* hin = train ? g.h[l].p : g.h[l % 2].p;
        float* hout = train ? g.h[l + 1].p : g.h[(l + 1) % 2].p;
        hipLaunchKernelGGL(k_lf_layer, dim3(mb), dim3(MP_THREADS), 0, s, n, g.outer.p, g.cols.p, feat,
                           w + P_LAYER0 + l * P_LAYER_SIZE, hin, hout, train ? g.a[l + 1].p : (float*)nullptr,
                           train ? g.m[l].p : (float*)nullptr);
    }
    const float* hl = train ? g.h[MP_L].p : g.h[MP_L % 2].p;
    hipLaunchKernelGGL(k_lf_output<Z>, dim3(mb), dim3(MP_THREADS), 0, s, n, hl, w, g.x.p, g.s.p, g.sigma.p, z);
}

// Loss ||z - e||_A^2 / ||e||_A^2 and its gradient w.r.t. the weights (into g.grad).
double learned_loss_grad(LearnedGraph& g, const float* w, nlf_matrix A, const double* r, const double* e, hipStream_t s) {
    learned_alloc_train(g);
    const i64 n = g.n;
    learned_forward<double, double>(g, w, r, g.zbuf.p, true, s);
    const int lpr = pick_lpr(double(A->nnz_full) / double(std::max<i64>(A->rows, 1)));
    launch_spmv<double, double, double, false>(lpr, n, A->outer.p, A->cols.p, A->val.p, g.zbuf.p, g.az.p, nullptr, nullptr, s);
    const unsigned vb = grid_stride_blocks(n);
    DevBuf<double> p1{vb}, p2{vb}, sums{2}, gz{size_t(n)};
    hipLaunchKernelGGL(k_lf_loss, dim3(vb), dim3(256), 0, s, n, g.zbuf.p, e, g.az.p, r, gz.p, p1.p, p2.p);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, p1.p, int(vb), sums.p);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, p2.p, int(vb), sums.p + 1);
    hipLaunchKernelGGL(k_scale_vec, dim3(vb), dim3(256), 0, s, n, gz.p, sums.p);
    // backward
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
        hipLaunchKernelGGL(k_lf_layer_back1, dim3(gb), dim3(MP_THREADS), 0, s, n, gcur, g.a[l + 1].p, g.h[l].p, g.m[l].p,
                           wl, g.gdir.p, g.gm.p, gp);
        hipLaunchKernelGGL(k_lf_layer_back2, dim3(gb), dim3(MP_THREADS), 0, s, n, g.outer.p, g.cols.p, feat, wl, g.gm.p,
                           g.h[l].p, g.gdir.p, gnext, gp);
        std::swap(gcur, gnext);
    }
    hipLaunchKernelGGL(k_lf_input_back, dim3(gb), dim3(MP_THREADS), 0, s, n, gcur, g.a[0].p, g.x.p, g.gpart.p);
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

template<class V>
i64 fcg(hipStream_t s, nlf_matrix A, const V* val, nlf_precond M, const double* b, double* x, double tol, i64 max_it,
        double* true_res, i64* updates, int window, KernelTimer& timer) {
    const i64 n = A->rows;
    LearnedGraph& g = *learned_of(M);
    const int lpr = pick_lpr(double(A->nnz_full) / double(std::max<i64>(n, 1)));
    const unsigned sb = spmv_blocks(lpr, n), vb = grid_stride_blocks(n), rb = spmv_blocks(64, n);
    const size_t np = std::max<size_t>(std::max<size_t>(sb, vb), rb);
    DevBuf<double> r{size_t(n)}, z{size_t(n)}, p{size_t(n)}, q{size_t(n)}, pa{np}, pb{np}, sums{2};
    DevBuf<FCGState> st{1};
    auto true_residual = [&]() {
        hipLaunchKernelGGL((k_residual<double, double>), dim3(rb), dim3(SPMV_THREADS), 0, s, n, A->outer.p, A->cols.p, A->val.p,
                           x, b, r.p, pa.p, pb.p);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, pa.p, int(rb), sums.p);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, pb.p, int(rb), sums.p + 1);
        double h[2];
        download(h, sums.p, 2, s);
        return std::make_pair(std::sqrt(h[0]), std::sqrt(h[1]));
    };
    auto spmv_pq = [&](bool sample) {             // q = A p, partial p.q into pb; partial p.r into pa
        if (sample) timer.start(0, s);
        launch_spmv<V, double, double, true>(lpr, n, A->outer.p, A->cols.p, val, p.p, q.p, pb.p, nullptr, s);
        if (sample) timer.stop(0, s);
        hipLaunchKernelGGL(k_dot<double>, dim3(sb), dim3(SPMV_THREADS), 0, s, n, p.p, r.p, pa.p, (const CGState*)nullptr);
    };
    auto apply_m = [&](bool sample) {
        if (sample) timer.start(1, s);
        learned_forward<double, double>(g, g.w.p, r.p, z.p, false, s);
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
                if (est < 0.1 * last_update || est <= tol * bn || it >= max_it) {
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

// Deterministic initial weights; the output layer starts at zero, so the initial network is Jacobi.
int nlf_learned_init_weights(uint64_t seed, float* w) {
    if (!w) return NLF_ERR_ARGUMENT;
    uint64_t st = seed * 0x2545f4914f6cdd1dull + 1;
    for (int i = 0; i < NPARAM; ++i) w[i] = 0.0f;
    for (int c = 0; c < MP_H; ++c) { w[P_AIN + c] = uniform(st, 1.0f); w[P_BIN + c] = uniform(st, 0.1f); }
    const float a = std::sqrt(6.0f / float(2 * MP_H));
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
        hipStream_t s = A->ctx->stream;
        const nlf_mesh m = A->mesh;
        auto P = std::make_unique<nlf_precond_s>();
        P->ctx = A->ctx;
        P->kind = NLF_PRECOND_LEARNED;
        P->rows = A->rows;
        auto g = std::make_unique<LearnedGraph>();
        g->n = A->rows;
        g->nnz = A_local->nnz_full;
        std::vector<i64> lo(size_t(g->n + 1));
        download(lo.data(), A_local->outer.p, lo.size(), s);
        upload(g->outer, lo.data(), lo.size(), s);
        std::vector<i32> lc(size_t(g->nnz));
        download(lc.data(), A_local->cols.p, lc.size(), s);
        upload(g->cols, lc.data(), lc.size(), s);
        g->s.alloc(size_t(g->n));
        g->feat.alloc(size_t(4 * g->nnz));
        g->w.alloc(NPARAM);
        g->x.alloc(size_t(g->n));
        g->h[0].alloc(size_t(g->n) * MP_H);
        g->h[1].alloc(size_t(g->n) * MP_H);
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
        NLF_HIP(hipStreamSynchronize(p->ctx->stream));
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
// matrix values in the Krylov SpMV (nlf_matrix_make_fp32 first).
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
            it = fcg<float>(s, A, A->val32.p, M, b, x, tol, max_it, &out.true_residual, &out.outer_iterations, 400, timer);
        } else {
            it = fcg<double>(s, A, A->val.p, M, b, x, tol, max_it, &out.true_residual, &out.outer_iterations, 400, timer);
        }
        out.spmv_ms = timer.avg(0);
        out.precond_ms = timer.avg(1);
        out.spmv_samples = timer.count[0];
        out.spmv_bytes = A->nnz_full * i64((mode == NLF_MODE_MIXED ? 4 : 8) + 4) + A->rows * i64(8 + 16);
        if (it < 0) {
            // stagnation: fall back to PCG with the fallback preconditioner from the current iterate
            out.fallbacks = 1;
            out.preconditioner = fallback ? fallback->kind : NLF_PRECOND_IDENTITY;
            // PCG stops on sqrt(r'z)/||b|| (conjugate_gradient.hpp:48-79); the true residual decides, so
            // the PCG tolerance is tightened tenfold (at most three times) until the true residual meets tol
            int breakdown = 0;
            CG<double, double> cg(s, A->rows, A->nnz_full, A->outer.p, A->cols.p, A->val.p, fallback);
            const unsigned rb = spmv_blocks(64, A->rows);
            DevBuf<double> r{size_t(A->rows)}, prr{rb}, pbb{rb}, sums{2};
            i64 left = max_it + it, total = -it;
            double pcg_tol = tol;
            for (int round = 0; round < 4; ++round) {
                const i64 more = cg.solve(b, x, pcg_tol, left, &out.residual, &breakdown);
                total += more;
                left -= more;
                hipLaunchKernelGGL((k_residual<double, double>), dim3(rb), dim3(SPMV_THREADS), 0, s, A->rows, A->outer.p,
                                   A->cols.p, A->val.p, x, b, r.p, prr.p, pbb.p);
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
