"""nlf_cg_solve_noalloc (the CG solve whose work memory is a fixed device array of the library, no
allocation) against nlf_cg_solve on the same matrix, preconditioner and right-hand side: equal
iterations and solutions, on two calls in a row, the second with a smaller system, and the
size limit."""
import ctypes as C

import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, configs, meshgen
from nonlocal_amd._lib import load
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

NOALLOC_MAX_ROWS = 2097152


def _system(problem, ctx, precond):
    g = api.mesh_2d(problem.mesh, ctx)
    api.find_neighbours(g, H.radii(problem.models))
    f, xD, inner = api.boundary_data(g, problem.dof, problem.bcs)
    K = api._assemble(g, problem.physics, problem.models, problem.materials, inner)
    if problem.physics == 0:
        K.add_convection(problem.bcs)
    b = torch.from_numpy(f).to(ctx.device)
    K.apply_first_kind(torch.from_numpy(xD).to(ctx.device), b)
    KL = api._assemble(g, problem.physics, problem.models, problem.materials, inner, part=A.NLF_PART_LOCAL)
    P = api.preconditioner(KL if precond == A.NLF_PRECOND_ILLT else K, precond)
    return g, K, KL, P, b


def _solve(ctx, K, P, b, noalloc, tol=1e-13):
    x = torch.zeros_like(b)
    st = A.SolveStats()
    if noalloc:
        rc = load().nlf_cg_solve_noalloc(K.h, P.h, api._dev(b), api._dev(x), tol, 100000, C.byref(st))
    else:
        # the same products as the no-alloc solve, which reads the full storage (the symmetric slice
        # storage of nlf_cg_solve is memory the no-alloc entry point does not allocate)
        ctx.set_option(A.NLF_OPT_SPMV_SYMMETRIC, 0)
        try:
            rc = load().nlf_cg_solve(K.h, P.h, api._dev(b), api._dev(x), tol, 100000, A.NLF_MODE_FP64, C.byref(st))
        finally:
            ctx.set_option(A.NLF_OPT_SPMV_SYMMETRIC, 1)
    ctx.check(rc)
    ctx.synchronize()
    return x, st


@pytest.mark.parametrize("precond", [A.NLF_PRECOND_ILLT, A.NLF_PRECOND_JACOBI, A.NLF_PRECOND_IDENTITY])
def test_noalloc_cg_matches_cg_two_calls(gpu_ctx, precond):
    """A 48 x 48 composite, then a 20 x 20 plate: the second call finds the work memory holding the
    first call's data and must not depend on it."""
    problems = [configs.composite(48, r=0.08, count=4), configs.cfg0(scale=0.5)]
    assert problems[1].mesh.n_nodes < problems[0].mesh.n_nodes
    for problem in problems:
        print(f"[noalloc] {problem.name} precond={precond}", flush=True)
        g, K, KL, P, b = _system(problem, gpu_ctx, precond)
        x_ref, st_ref = _solve(gpu_ctx, K, P, b, noalloc=False)
        x, st = _solve(gpu_ctx, K, P, b, noalloc=True)
        print(f"  iterations {st.iterations} / {st_ref.iterations}, true residual {st.true_residual:.2e}, "
              f"bitwise equal {torch.equal(x, x_ref)}", flush=True)
        assert st.iterations == st_ref.iterations
        assert H.rel_l2(x.cpu().numpy(), x_ref.cpu().numpy()) <= 1e-14
        assert st.true_residual <= 1e-12 and abs(st.true_residual - st_ref.true_residual) <= 1e-15
        # and the solution is the oracle's
        om, Ko, fo, xD, inner = H.oracle_assembly(problem)
        assert H.rel_l2(x.cpu().numpy(), H.direct_solution(Ko, fo)) <= 1e-10


def test_noalloc_cg_rejects_too_many_rows(gpu_ctx):
    """A local-theory plate of 1500 x 1500 elements has 2,253,001 rows, more than the fixed work memory
    holds: an argument error before any kernel runs."""
    mesh = meshgen.structured_q4(1500)
    assert mesh.n_nodes > NOALLOC_MAX_ROWS
    problem = configs.thermal_q4(1500, p1=1.0, r=0.0)
    problem.models = [A.model(1.0, 0.0)]
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    f, xD, inner = api.boundary_data(g, 1, problem.bcs)
    K = api._assemble(g, 0, problem.models, problem.materials, inner)
    b = torch.from_numpy(f).to(gpu_ctx.device)
    x = torch.zeros_like(b)
    st = A.SolveStats()
    rc = load().nlf_cg_solve_noalloc(K.h, None, api._dev(b), api._dev(x), 1e-10, 10, C.byref(st))
    assert rc == A.NLF_ERR_ARGUMENT
    assert b"NLF_NOALLOC_MAX_ROWS" in load().nlf_last_error(gpu_ctx.h)
    assert float(x.abs().max()) == 0.0
