"""CPU pins of the learned-preconditioner oracle (oracle/learned_oracle.py) and of the weights format.

The reference has no learned component, so the oracle is pinned by mathematics: finite differences
of its own forward pass for the hand-written backward pass, the Jacobi identity at zero output
weights, and closed values of the loss.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from nonlocal_amd import configs
from oracle import learned_oracle as LO
from oracle import oracle as O
from tests.helpers import oracle_assembly, upper_to_full


def _small_graph(seed=0):
    """A real composite problem (8 x 8 Q4, 2 inclusions): pattern of K_L, features and K."""
    p = configs.composite(8, seed=1, count=2, r=0.3)
    om, K, f, xD, inner = oracle_assembly(p)
    _, KL, _, _, _ = oracle_assembly(p, part=0, convection=False)
    n = p.mesh.n_nodes
    Kf = upper_to_full(K.export(0), n)
    KLf = upper_to_full(KL.export(0), n)
    outer, cols, kl = KLf.indptr.astype(np.int64), KLf.indices.astype(np.int64), KLf.data
    groups = LO.node_groups(p.mesh.elems, p.mesh.group_begin, n)
    inf = lambda m, a, b: O.influence(m, a, b)                      # noqa: E731
    feat, s = LO.features(outer, cols, kl, Kf, p.mesh.xy, groups, p.models, inf)
    return p, Kf, outer, cols, feat, s, f


def _random_weights(seed, scale_out=0.3):
    rng = np.random.default_rng(seed)
    w = rng.uniform(-0.5, 0.5, LO.NPARAM)
    w[LO.P_WOUT:LO.P_WOUT + LO.H] = rng.uniform(-scale_out, scale_out, LO.H)
    return w


@pytest.fixture(scope="module")
def graph():
    return _small_graph()


def test_layout_matches_library():
    from nonlocal_amd import learned
    assert learned.num_params() == LO.NPARAM == 6721


def test_features_symmetric_and_scaled(graph):
    p, Kf, outer, cols, feat, s, f = graph
    n = len(outer) - 1
    rows = np.repeat(np.arange(n), np.diff(outer))
    # the edge (v, u) carries the same features as (u, v)
    idx = {(int(u), int(v)): k for k, (u, v) in enumerate(zip(rows, cols))}
    for k, (u, v) in enumerate(zip(rows, cols)):
        np.testing.assert_array_equal(feat[k], feat[idx[(int(v), int(u))]])
    diag = rows == cols
    np.testing.assert_allclose(feat[diag, 1], 1.0, rtol=1e-14)          # K_uu s_u^2 = 1
    assert np.all(feat[diag, 3] == 1.0) and np.all(feat[~diag, 3] == 0.0)
    # f2 on the diagonal is (1 - p1) for nonlocal nodes (phi(x, x) / phi(0) = 1)
    np.testing.assert_allclose(feat[diag, 2], 0.5, rtol=1e-14)
    assert np.all(feat[:, 2] >= 0) and np.all(feat[:, 2] <= 0.5 + 1e-15)


def test_node_groups_first_element():
    elems = np.array([[0, 1, 3, 2], [1, 4, 5, 3]])
    g = LO.node_groups(elems, np.array([0, 1, 2]), 6)
    np.testing.assert_array_equal(g, [0, 0, 0, 0, 1, 1])


def test_zero_output_weights_is_jacobi(graph):
    p, Kf, outer, cols, feat, s, f = graph
    w = _random_weights(1)
    w[LO.P_WOUT:] = 0.0
    r = np.random.default_rng(2).standard_normal(len(s))
    z = LO.forward(w, outer, cols, feat, s, r)
    np.testing.assert_allclose(z, r / Kf.diagonal(), rtol=1e-13)


def test_transpose_adjoint_identity(graph):
    p, Kf, outer, cols, feat, s, f = graph
    w = _random_weights(8)
    rng = np.random.default_rng(9)
    u, v = rng.standard_normal((2, len(s)))
    lhs = v @ LO.forward(w, outer, cols, feat, s, u)
    rhs = LO.transpose(w, outer, cols, feat, s, v) @ u
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_symmetric_operator_against_explicit_matrix(graph):
    """Build N column by column; M = (N + N^T)/2 must match apply_sym, be symmetric, and N must
    be linear with a three-hop stencil."""
    p, Kf, outer, cols, feat, s, f = graph
    w = _random_weights(11)
    n = len(s)
    N = np.stack([LO.forward(w, outer, cols, feat, s, np.eye(n)[j]) for j in range(n)], axis=1)
    r = np.random.default_rng(12).standard_normal(n)
    np.testing.assert_allclose(N @ r, LO.forward(w, outer, cols, feat, s, r), rtol=1e-11, atol=1e-13 * np.abs(N).max())
    Msym = 0.5 * (N + N.T)
    np.testing.assert_allclose(LO.apply_sym(w, outer, cols, feat, s, r), Msym @ r, rtol=1e-11,
                               atol=1e-13 * np.abs(N).max())
    KL = sp.csr_matrix((np.ones(len(cols)), cols, outer), shape=(n, n))
    hop3 = (KL @ KL @ KL).toarray() != 0
    assert np.all(N[~hop3] == 0.0)


def test_loss_closed_values(graph):
    p, Kf, outer, cols, feat, s, f = graph
    r = np.random.default_rng(3).standard_normal(len(s))
    e = sp.linalg.spsolve(Kf.tocsc(), r)
    assert abs(LO.loss(e, e, Kf, r)) < 1e-12
    assert abs(LO.loss(np.zeros_like(e), e, Kf, r) - 1.0) < 1e-12


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_finite_differences(graph, seed):
    p, Kf, outer, cols, feat, s, f = graph
    rng = np.random.default_rng(10 + seed)
    w = _random_weights(seed)
    r = rng.standard_normal(len(s))
    e = sp.linalg.spsolve(Kf.tocsc(), r)
    val, g = LO.loss_grad(w, outer, cols, feat, s, Kf, r, e)
    assert np.isfinite(val) and val > 0
    # every parameter block: input, each layer's W_e/W_s/W_m/b, output
    blocks = [(LO.P_AIN, LO.H), (LO.P_BIN, LO.H), (LO.P_WOUT, LO.H + 1)]
    for l in range(LO.L):
        o = LO.P_LAYER0 + l * LO.P_LAYER_SIZE
        blocks += [(o + LO.P_WE, LO.H * LO.F), (o + LO.P_WS, LO.H * LO.H), (o + LO.P_WM, LO.H * LO.H),
                   (o + LO.P_B, LO.H)]
    idx = np.concatenate([b + rng.choice(n, size=min(n, 6), replace=False) for b, n in blocks])
    gs = np.abs(g).max()
    for i in idx:
        hstep = 1e-6 * max(1.0, abs(w[i]))
        wp, wm = w.copy(), w.copy()
        wp[i] += hstep
        wm[i] -= hstep
        lp = LO.loss(LO.forward(wp, outer, cols, feat, s, r), e, Kf, r)
        lm = LO.loss(LO.forward(wm, outer, cols, feat, s, r), e, Kf, r)
        fd = (lp - lm) / (2 * hstep)
        assert abs(fd - g[i]) <= 1e-6 * gs + 1e-5 * abs(fd), (i, fd, g[i])


def test_planted_backward_mistake_is_caught(graph):
    """A dropped transpose in the message gradient must fail the finite-difference comparison."""
    p, Kf, outer, cols, feat, s, f = graph
    rng = np.random.default_rng(5)
    w = _random_weights(7)
    r = rng.standard_normal(len(s))
    e = sp.linalg.spsolve(Kf.tocsc(), r)
    o = LO.P_LAYER0 + LO.P_WM
    val, g = LO.loss_grad(w, outer, cols, feat, s, Kf, r, e)
    Wm = g[o:o + LO.H * LO.H].reshape(LO.H, LO.H)
    bad = Wm.T.ravel()
    i = int(np.argmax(np.abs(Wm - Wm.T)))
    hstep = 1e-6
    wp, wm = w.copy(), w.copy()
    wp[o + i] += hstep
    wm[o + i] -= hstep
    fd = (LO.loss(LO.forward(wp, outer, cols, feat, s, r), e, Kf, r) -
          LO.loss(LO.forward(wm, outer, cols, feat, s, r), e, Kf, r)) / (2 * hstep)
    assert abs(fd - g[o + i]) < 1e-5 * max(1.0, abs(fd))
    assert abs(fd - bad[i]) > 1e-3 * max(1e-9, abs(fd))


def test_init_weights_deterministic_and_jacobi_start():
    from nonlocal_amd import learned
    w1, w2, w3 = learned.init_weights(5), learned.init_weights(5), learned.init_weights(6)
    np.testing.assert_array_equal(w1, w2)
    assert not np.array_equal(w1, w3)
    assert np.all(w1[LO.P_WOUT:] == 0.0)
    assert np.all(np.abs(w1[LO.P_AIN:LO.P_AIN + LO.H]) <= 1.0)


def test_weights_file_roundtrip(tmp_path):
    from nonlocal_amd import learned
    w = learned.init_weights(3)
    path = str(tmp_path / "w.bin")
    learned.save_weights(path, w, 3, {"steps": 1})
    w2, meta = learned.load_weights(path)
    np.testing.assert_array_equal(w, w2)
    assert meta["seed"] == 3
    with open(path, "r+b") as fh:
        fh.write(b"XXXX")
    with pytest.raises(ValueError):
        learned.load_weights(path)
