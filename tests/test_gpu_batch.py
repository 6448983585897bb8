"""Batched solves (block-diagonal batch of uneven instances, lockstep Jacobi-PCG under a captured
hipGraph): per-instance solutions against the oracle, graph replays against eager launches, the
autotuner and its cache."""
import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import batch, configs
from tests.helpers import direct_solution, oracle_assembly, rel_l2

pytestmark = pytest.mark.gpu


def _instances():
    """Six uneven thermal instances, one LOCAL (p1 = 1), radii from 0.9 h to 3 h."""
    rng = np.random.default_rng(7)
    out = []
    for k, n in enumerate([6, 11, 8, 14, 9, 12]):
        p1 = 1.0 if k == 2 else float(rng.uniform(0.2, 0.9))
        r = float(rng.uniform(0.9, 3.0)) / n
        lam = float(rng.uniform(0.5, 2.0))
        out.append(configs.thermal_q4(n, p1, r, lam, name=f"b{k}"))
    return out


@pytest.fixture(scope="module")
def solver(gpu_ctx):
    print("[batch] assembling 6 instances", flush=True)
    return batch.BatchSolver(_instances(), gpu_ctx)


@pytest.fixture(scope="module")
def references():
    refs = []
    for p in _instances():
        om, K, f, xD, inner = oracle_assembly(p)
        refs.append(direct_solution(K, f))
    return refs


def test_block_diagonal_assembly_matches_instances(solver):
    n = solver.node_begin
    csr = solver.K.export(0)
    for i, p in enumerate(solver.problems):
        om, K, f, xD, inner = oracle_assembly(p)
        Ko = K.export(0)
        r0, r1 = n[i], n[i + 1]
        outer = csr.outer[r0:r1 + 1] - csr.outer[r0]
        np.testing.assert_array_equal(outer, Ko.outer)
        np.testing.assert_array_equal(csr.inner[csr.outer[r0]:csr.outer[r1]] - r0, Ko.inner)
        np.testing.assert_allclose(csr.val[csr.outer[r0]:csr.outer[r1]], Ko.val, rtol=1e-12, atol=1e-12 * np.abs(Ko.val).max())


@pytest.mark.parametrize("mode,bound,assembly32", [(A.NLF_MODE_FP64, 1e-10, False), (A.NLF_MODE_MIXED, 1e-8, False),
                                                   (A.NLF_MODE_MIXED, 1e-8, True)])
def test_batch_solutions_match_oracle(solver, references, mode, bound, assembly32):
    x, its, res, st = solver.solve(tol=1e-12, mode=mode, assembly32=assembly32)
    print(f"[batch] mode {mode} fp32 assembly {assembly32}: iterations {its.tolist()}, "
          f"max true residual {res.max():.2e}", flush=True)
    assert np.all(res <= 1e-12)
    for xi, ref in zip(solver.split(x), references):
        assert rel_l2(xi.cpu().numpy(), ref) < bound


def test_graph_replays_match_eager_launches(solver):
    xg1, itg1, _, _ = solver.solve(tol=1e-11, use_graph=True)
    torch.cuda.synchronize()
    xe, ite, _, _ = solver.solve(tol=1e-11, use_graph=False)
    xg2, itg2, _, _ = solver.solve(tol=1e-11, use_graph=True)
    assert torch.equal(xg1, xe) and torch.equal(xg1, xg2)
    np.testing.assert_array_equal(itg1, ite)
    np.testing.assert_array_equal(itg1, itg2)


def test_autotune_and_cache(solver, tmp_path, references):
    cache = str(tmp_path / "tune.txt")
    t1 = solver.autotune(cache)
    assert t1 in (1.0, 2.0, 4.0, 8.0, 16.0)
    with open(cache) as f:
        lines = f.read().split()
    assert len(lines) == 2
    t2 = solver.autotune(cache)
    assert t2 == t1
    with open(cache) as f:
        assert len(f.read().split()) == 2          # reused, not measured again
    x, its, res, st = solver.solve(tol=1e-12)
    for xi, ref in zip(solver.split(x), references):
        assert rel_l2(xi.cpu().numpy(), ref) < 1e-10


@pytest.mark.parametrize("mode,bound", [(A.NLF_MODE_FP64, 1e-10), (A.NLF_MODE_MIXED, 1e-8)])
def test_batch_illt_matches_oracle_in_fewer_iterations(gpu_ctx, references, mode, bound):
    """ILLT of the batch's local matrix (block diagonal, one sweep for every instance): each instance's
    solution is the oracle's, in fewer iterations than with Jacobi, and graph replays equal eager launches."""
    jac = batch.BatchSolver(_instances(), gpu_ctx)
    illt = batch.BatchSolver(_instances(), gpu_ctx, precond=A.NLF_PRECOND_ILLT)
    _, its_j, _, _ = jac.solve(tol=1e-12, mode=mode)
    x, its, res, st = illt.solve(tol=1e-12, mode=mode)
    print(f"[batch illt] mode {mode}: iterations {its.tolist()} (Jacobi {its_j.tolist()}), "
          f"max true residual {res.max():.2e}", flush=True)
    assert st.preconditioner == A.NLF_PRECOND_ILLT
    assert res.max() <= 1e-12
    for xi, ref in zip(illt.split(x), references):
        assert rel_l2(xi.cpu().numpy(), ref) <= bound
    assert its.max() < its_j.max() and its.sum() < its_j.sum()
    x_eager, its_e, _, _ = illt.solve(tol=1e-12, mode=mode, use_graph=False)
    assert np.array_equal(its, its_e) and torch.equal(x, x_eager)
