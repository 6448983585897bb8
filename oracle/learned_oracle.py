"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

FP64 numpy restatement of the learned preconditioner (message-passing network over mesh nodes):
edge features, forward pass, loss ||z - e||_A^2 / ||e||_A^2 and its gradient by reverse-mode
differentiation written out by hand.  The reference has no learned component; what pins this file
is mathematics, not the reference:
  * the backward pass is checked against central finite differences of the forward pass;
  * with zero output weights the network is exactly Jacobi, z = r / diag(K);
  * the loss is zero when z = e and equals 1 when z = 0;
  * edge features: f0/f1 are the symmetrically scaled K_L / K entries, f2 uses the influence function
    of the node's group (oracle/nlf_oracle.cpp, influence_functions_2d.hpp:39-148).
The parameter layout is the one of nlf_learned_num_params() floats.
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


def forward(w, outer, cols, feat, s, r, keep=False):
    """z = M(r).  Returns z (and the activations when keep)."""
    a_in, b_in, layers, w_out, b_out = unpack(w)
    s = np.asarray(s, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    sr = s * r
    sigma = np.sqrt(np.mean(sr * sr))
    if not (sigma > 0 and np.isfinite(sigma)):
        sigma = 1.0
    x = sr / sigma
    a = [np.outer(x, a_in) + b_in]
    h = [np.maximum(a[0], 0.0)]
    m = []
    for lay in layers:
        G = feat @ lay["We"].T                      # nnz x H edge gates
        mm = _segsum(G * h[-1][cols], outer)
        al = h[-1] @ lay["Ws"].T + mm @ lay["Wm"].T + lay["b"]
        m.append(mm)
        a.append(al)
        h.append(np.maximum(al, 0.0))
    y = h[-1] @ w_out + b_out
    z = s * sigma * (x + y)
    if keep:
        return z, dict(x=x, sigma=sigma, a=a, h=h, m=m)
    return z


def loss(z, e, A, r):
    """||z - e||_A^2 / ||e||_A^2 with A e = r, in the form the GPU evaluates: (z-e).(Az - r) / e.r."""
    return float((z - e) @ (A @ z - r) / (e @ r))


def loss_grad(w, outer, cols, feat, s, A, r, e):
    """Loss and dLoss/dw (FP64) by hand-written reverse mode."""
    a_in, b_in, layers, w_out, b_out = unpack(w)
    s = np.asarray(s, dtype=np.float64)
    z, act = forward(w, outer, cols, feat, s, r, keep=True)
    Az = A @ z
    er = float(e @ r)
    val = float((z - e) @ (Az - r) / er)
    gz = 2.0 * (Az - r) / er
    g = np.zeros(NPARAM)
    dy = s * act["sigma"] * gz
    h, a, m, x = act["h"], act["a"], act["m"], act["x"]
    g[P_WOUT:P_WOUT + H] = dy @ h[L]
    g[P_BOUT] = dy.sum()
    gh = np.outer(dy, w_out)
    rows = _rows(outer)
    for l in range(L - 1, -1, -1):
        lay = layers[l]
        o = P_LAYER0 + l * P_LAYER_SIZE
        da = gh * (a[l + 1] > 0)
        g[o + P_WS:o + P_WM] = (da.T @ h[l]).ravel()
        g[o + P_WM:o + P_B] = (da.T @ m[l]).ravel()
        g[o + P_B:o + P_B + H] = da.sum(axis=0)
        gm = da @ lay["Wm"]                          # dL/dm
        G = feat @ lay["We"].T
        # m_u = sum_{p in row u} G_p * h_{cols p}: dh_v += G_p * gm_u ; dWe[c,f] += gm_u[c] h_v[c] f_p[f]
        ghin = da @ lay["Ws"]
        np.add.at(ghin, cols, G * gm[rows])
        t = gm[rows] * h[l][cols]                    # nnz x H
        g[o + P_WE:o + P_WS] = (t.T @ feat).ravel()
        gh = ghin
    da0 = gh * (a[0] > 0)
    g[P_AIN:P_AIN + H] = da0.T @ x
    g[P_BIN:P_BIN + H] = da0.sum(axis=0)
    return val, g
