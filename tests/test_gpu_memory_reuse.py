"""Device blocks released inside a context's calls go back to that context's free list and are
taken again by its later allocations (stream-ordered reuse instead of hipMalloc / hipFree).  A reused
block holds stale data, so a kernel that relied on fresh memory would change its results: the same
solve, repeated after other work of other sizes on the same context, must give the same bits and the
oracle's solution; and handles destroyed after their context must be released without touching it."""
import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, configs
from nonlocal_amd._lib import load
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _solve(problem, ctx, precond):
    g = api.mesh_2d(problem.mesh, ctx)
    fn = api.stationary_heat_equation_solver_2d if problem.physics == 0 else api.equilibrium_equation
    sol = fn(g, problem.models, problem.materials, problem.bcs, precond=precond, tol=1e-12, mode=A.NLF_MODE_FP64)
    return sol.x.cpu().numpy(), int(sol.iterations)


def test_repeated_solves_give_the_same_bits(gpu_ctx):
    a = configs.thermal_q4(40, 0.5, 0.08)          # nonlocal: pair chunks, symmetric storage, ILLT
    print("[memory reuse] thermal 40 / elastic / thermal 25 / thermal 40", flush=True)
    x1, it1 = _solve(a, gpu_ctx, A.NLF_PRECOND_ILLT)
    _solve(configs.cfg2(scale=0.08), gpu_ctx, A.NLF_PRECOND_JACOBI)        # other sizes in between
    x2, it2 = _solve(a, gpu_ctx, A.NLF_PRECOND_ILLT)
    _solve(configs.thermal_q4(25, 0.5, 0.1), gpu_ctx, A.NLF_PRECOND_ILLT)   # smaller: larger blocks reused
    x3, it3 = _solve(a, gpu_ctx, A.NLF_PRECOND_ILLT)
    print(f"  iterations {it1} {it2} {it3}", flush=True)
    assert it1 == it2 == it3
    assert np.array_equal(x1, x2) and np.array_equal(x1, x3)
    om, K, f, xD, inner = H.oracle_assembly(a)
    assert H.rel_l2(x1, H.direct_solution(K, f)) <= 1e-10


def test_handles_outliving_their_context():
    default = api.default_context()
    ctx = api.Context(0)
    try:
        problem = configs.thermal_q4(20, 0.5, 0.12)
        g = api.mesh_2d(problem.mesh, ctx)
        api._search(g, problem.models)
        inner = np.ones(problem.mesh.n_nodes, dtype=np.uint8)
        K = api._assemble(g, 0, problem.models, problem.materials, inner)
        x = torch.randn(K.info()["rows"], dtype=torch.float64, device=ctx.device)
        y = K.spmv(x)
        ctx.synchronize()
        assert torch.isfinite(y).all()
        assert load().nlf_ctx_destroy(ctx.h) == 0                  # the context goes first ...
        ctx.h = None
        del K, g                                                    # ... then its handles
        torch.cuda.synchronize()
    finally:
        torch.cuda.set_stream(default.stream)
