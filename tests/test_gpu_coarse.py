"""Two-level learned preconditioner: B r = M r + P A_c^-1 P^T r with an aggregation coarse space.

The coarse part is checked against dense linear algebra: the aggregation (a grid of cells over the
node coordinates, restated below with the same floating-point operations), the Galerkin product
A_c = P^T K P built from the exported matrix, and numpy's inverse.  The two-level flexible CG must
reach the direct FP64 solution and need far fewer iterations than the one-level form.
"""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from nonlocal_amd import learned
from tests.helpers import direct_solution, oracle_assembly, rel_l2

pytestmark = pytest.mark.gpu


def _aggregates(xy, target, dof=1):
    """Coarse row of every fine row, as the library forms it."""
    n_nodes = xy.shape[0]
    target = min(target, max(1, n_nodes // 16))
    target = min(target, 4096 // dof)
    xmin, xmax = xy[:, 0].min(), xy[:, 0].max()
    ymin, ymax = xy[:, 1].min(), xy[:, 1].max()
    w = max(xmax - xmin, 1e-300)
    h = max(ymax - ymin, 1e-300)
    cx = min(target, max(1, math.floor(math.sqrt(float(target) * w / h) + 0.5)))
    cy = max(1, target // cx)
    hx, hy = w / cx, h / cy
    ix = np.minimum(cx - 1, ((xy[:, 0] - xmin) / hx).astype(np.int64))
    iy = np.minimum(cy - 1, ((xy[:, 1] - ymin) / hy).astype(np.int64))
    cell = ix * cy + iy
    used = np.unique(cell)
    cid = np.full(cx * cy, -1, dtype=np.int64)
    cid[used] = np.arange(used.size)
    return cid[cell]


@pytest.fixture(scope="module")
def problem(gpu_ctx):
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights committed")
    print("[coarse] 64x64 composite", flush=True)
    w, _ = learned.load_weights()
    tp = learned.TrainingProblem(64, seed=4321, ctx=gpu_ctx)
    tp.learned.set_weights(w)
    return tp


@pytest.mark.parametrize("target", [128, 1024])
def test_coarse_apply_matches_dense_algebra(problem, gpu_ctx, target):
    """128 aggregates: two 64-row tiles of the blocked inverse; 1024 (264 after the 16-nodes cap):
    five tiles, so the trailing updates cover 4 x 4 tiles."""
    tp = problem
    n = tp.b.numel()
    r = torch.from_numpy(np.random.default_rng(0).standard_normal(n)).cuda()
    tp.learned.set_coarse(0)
    z_m = tp.learned.solve(r).cpu().numpy()
    tp.learned.set_coarse(target)
    rows, sec = tp.learned.coarse_info()
    z = tp.learned.solve(r).cpu().numpy()
    agg = _aggregates(tp.problem.mesh.xy, target)
    assert rows == agg.max() + 1
    csr = tp.K.export(2)
    K = sp.csr_matrix((csr.val, csr.inner, csr.outer), shape=(n, n))
    P = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, rows))
    Ac = (P.T @ K @ P).toarray()
    z_ref = z_m + P @ (np.linalg.inv(Ac) @ (P.T @ r.cpu().numpy()))
    err = rel_l2(z, z_ref)
    print(f"[coarse] {rows} coarse rows, set-up {sec * 1e3:.1f} ms, apply vs dense algebra {err:.2e}", flush=True)
    assert err <= 1e-10
    # reproducible: a second build gives the same operator bit for bit
    tp.learned.set_coarse(target)
    assert np.array_equal(tp.learned.solve(r).cpu().numpy(), z)
    tp.learned.set_coarse(0)


def test_coarse_removed_is_one_level(problem):
    tp = problem
    r = torch.from_numpy(np.random.default_rng(1).standard_normal(tp.b.numel())).cuda()
    tp.learned.set_coarse(0)
    z0 = tp.learned.solve(r)
    tp.learned.set_coarse(256)
    tp.learned.set_coarse(0)
    assert tp.learned.coarse_info() == (0, 0.0)
    assert torch.equal(tp.learned.solve(r), z0)


def test_two_level_fcg(problem):
    """Two-level flexible CG reaches the direct FP64 solution with far fewer iterations."""
    tp = problem
    om, K, f, xD, inner = oracle_assembly(tp.problem)
    x_ref = direct_solution(K, f)
    tp.learned.set_coarse(0)
    _, st1 = learned.fcg_solve(tp.K, tp.learned, tp.b, tol=1e-10, max_it=20000)
    tp.learned.set_coarse(256)
    x, st2 = learned.fcg_solve(tp.K, tp.learned, tp.b, tol=1e-10, max_it=20000)
    tp.learned.set_coarse(0)
    err = rel_l2(x.cpu().numpy(), x_ref)
    print(f"[coarse] one level {st1.iterations} it, two levels {st2.iterations} it, error {err:.2e}", flush=True)
    assert st2.fallbacks == 0 and st2.true_residual <= 1e-10
    assert err <= 1e-8
    assert st2.iterations < 0.6 * st1.iterations
