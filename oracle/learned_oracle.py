"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

FP64 numpy restatement of the learned preconditioner (message-passing network over mesh nodes):
edge features, the network N (linear in the residual), its transpose and the symmetric operator
M = (N + N^T)/2 applied in flexible CG, the loss ||N(r) - e||_A^2 / ||e||_A^2 and its gradient by
reverse-mode differentiation written out by hand.  The reference has no learned component; what pins this file
is mathematics, not the reference:
  * the backward pass is checked against central finite differences of the forward pass;
  * the transpose is checked by the adjoint identity <v, N u> = <N^T v, u> and the symmetric
    operator against the explicit matrix of N built column by column;
  * with zero output weights the network is exactly Jacobi, z = r / diag(K);
  * the loss is zero when z = e and equals 1 when z = 0;
  * edge features: f0/f1 are the symmetrically scaled K_L / K entries, f2 uses the influence function
    of the node's group (oracle/nlf_oracle.cpp, influence_functions_2d.hpp:39-148).
The parameter layout is the one of nlf_learned_num_params() floats; the bias slots are unused.
"""
from __future__ import annotations

import numpy as np

H, F, L = 32, 4, 3
P_AIN, P_BIN, P_LAYER0 = 0, H, 2 * H
P_WE, P_WS = 0, H * F
P_WM = P_WS + H * H
P_B = P_WM + H * H
P_LAYER_SIZE = P_B + H
P_WOUT = P_LAYER0 + L * P_LAYER_SIZE
P_BOUT = P_WOUT + H
NPARAM = P_BOUT + 1


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    assert w.size == NPARAM
    layers = []
    for l in range(L):
        o = P_LAYER0 + l * P_LAYER_SIZE
        layers.append(dict(We=w[o + P_WE:o + P_WS].reshape(H, F), Ws=w[o + P_WS:o + P_WM].reshape(H, H),
                           Wm=w[o + P_WM:o + P_B].reshape(H, H), b=w[o + P_B:o + P_B + H]))
    return w[P_AIN:P_AIN + H], w[P_BIN:P_BIN + H], layers, w[P_WOUT:P_WOUT + H], w[P_BOUT]


def _rows(outer):
    return np.repeat(np.arange(len(outer) - 1), np.diff(outer))


def _segsum(vals, outer):
    """Per-row sums of edge values (rows may not be empty: the graph has self loops)."""
    return np.add.reduceat(vals, outer[:-1], axis=0)


def node_groups(elems, group_begin, n_nodes):
    """Group of the lowest-numbered element containing each node."""
    E = elems.shape[0]
    first = np.full(n_nodes, E, dtype=np.int64)
    np.minimum.at(first, elems.reshape(-1), np.repeat(np.arange(E), elems.shape[1]))
    return np.searchsorted(np.asarray(group_begin), first, side="right") - 1


def features(outer, cols, kl_val, k_full, xy, groups_of_nodes, models, influence):
    """Edge features (nnz x 4) and node scaling s = 1/sqrt(K_uu) on the pattern (outer, cols) of K_L.

    k_full: scipy CSR of K (both triangles).  influence(model, x, y) evaluates phi.
    """
    n = len(outer) - 1
    kd = k_full.diagonal()
    s = np.where(kd > 0, 1.0 / np.sqrt(np.where(kd > 0, kd, 1.0)), 1.0)
    rows = _rows(outer)
    kuv = np.asarray(k_full[rows, cols]).ravel()
    ss = s[rows] * s[cols]
    nonlocal_ = [m.local_weight < 0.999 and m.radius[0] > 0 and m.radius[1] > 0 for m in models]

    def w(u, v):
        g = groups_of_nodes[u]
        if not nonlocal_[g]:
            return 0.0
        m = models[g]
        return (1.0 - m.local_weight) * influence(m, xy[u], xy[v]) / influence(m, xy[u], xy[u])

    f2 = np.array([0.5 * (w(u, v) + w(v, u)) for u, v in zip(rows, cols)])
    feat = np.stack([kl_val * ss, kuv * ss, f2, (rows == cols).astype(np.float64)], axis=1)
    return feat, s


def _scaled(s, r):
    s = np.asarray(s, dtype=np.float64)
    sr = s * np.asarray(r, dtype=np.float64)
    sigma = np.sqrt(np.mean(sr * sr))
    if not (sigma > 0 and np.isfinite(sigma)):
        sigma = 1.0
    return s, sr / sigma, sigma


def forward(w, outer, cols, feat, s, r, keep=False):
    """N(r): x = S r / sigma, h0 = a_in x, h' = W_s h + W_m sum_p (W_e f_p) (.) h_{cols p}, y = w_out . h3,
    N(r) = s sigma (x + y).  Returns N(r) (and the activations when keep)."""
    a_in, _, layers, w_out, _ = unpack(w)
    s, x, sigma = _scaled(s, r)
    h = [np.outer(x, a_in)]
    m = []
    for lay in layers:
        G = feat @ lay["We"].T                      # nnz x H edge gates
        mm = _segsum(G * h[-1][cols], outer)
        m.append(mm)
        h.append(h[-1] @ lay["Ws"].T + mm @ lay["Wm"].T)
    y = h[-1] @ w_out
    z = s * sigma * (x + y)
    if keep:
        return z, dict(x=x, sigma=sigma, h=h, m=m, y=y)
    return z


def transpose_y(w, outer, cols, feat, x):
    """Y^T x where y = Y x is the message path of the network (layers run backwards)."""
    a_in, _, layers, w_out, _ = unpack(w)
    g = np.outer(x, w_out)
    for lay in reversed(layers):
        G = feat @ lay["We"].T
        q = g @ lay["Wm"]                            # W_m^T g per node
        gin = g @ lay["Ws"]
        gin += _segsum(G * q[cols], outer)           # symmetric graph and gates
        g = gin
    return g @ a_in


def transpose(w, outer, cols, feat, s, v):
    """N^T v = S^2 v + S Y^T S v (sigma cancels: the operator is linear)."""
    s = np.asarray(s, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    return s * s * v + s * transpose_y(w, outer, cols, feat, s * v)


def apply_sym(w, outer, cols, feat, s, r):
    """M r = (N r + N^T r) / 2, the operator the flexible CG applies."""
    return 0.5 * (forward(w, outer, cols, feat, s, r) + transpose(w, outer, cols, feat, s, r))


def loss(z, e, A, r):
    """||z - e||_A^2 / ||e||_A^2 with A e = r, in the form the GPU evaluates: (z-e).(Az - r) / e.r."""
    return float((z - e) @ (A @ z - r) / (e @ r))


def loss_grad(w, outer, cols, feat, s, A, r, e):
    """Loss of N and dLoss/dw (FP64) by hand-written reverse mode; bias slots get zero gradient."""
    a_in, _, layers, w_out, _ = unpack(w)
    s = np.asarray(s, dtype=np.float64)
    z, act = forward(w, outer, cols, feat, s, r, keep=True)
    Az = A @ z
    er = float(e @ r)
    val = float((z - e) @ (Az - r) / er)
    gz = 2.0 * (Az - r) / er
    g = np.zeros(NPARAM)
    dy = s * act["sigma"] * gz
    h, m, x = act["h"], act["m"], act["x"]
    g[P_WOUT:P_WOUT + H] = dy @ h[L]
    gh = np.outer(dy, w_out)
    rows = _rows(outer)
    for l in range(L - 1, -1, -1):
        lay = layers[l]
        o = P_LAYER0 + l * P_LAYER_SIZE
        g[o + P_WS:o + P_WM] = (gh.T @ h[l]).ravel()
        g[o + P_WM:o + P_B] = (gh.T @ m[l]).ravel()
        gm = gh @ lay["Wm"]                          # dL/dm
        G = feat @ lay["We"].T
        # m_u = sum_{p in row u} G_p * h_{cols p}: dh_v += G_p * gm_u ; dWe[c,f] += gm_u[c] h_v[c] f_p[f]
        ghin = gh @ lay["Ws"]
        np.add.at(ghin, cols, G * gm[rows])
        t = gm[rows] * h[l][cols]                    # nnz x H
        g[o + P_WE:o + P_WS] = (t.T @ feat).ravel()
        gh = ghin
    g[P_AIN:P_AIN + H] = gh.T @ x
    return val, g
