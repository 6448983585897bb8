"""FP32 assembly and the mixed-precision solve against the FP64 oracle.

The FP32 value pass (nlf_matrix_assemble_fp32) contracts element pairs in FP32 (MFMA f32 blocks for
thermal Q4, FP32 FMA elsewhere) and accumulates in FP32; coordinates, Jacobian determinants and
the cutoff decisions of phi stay FP64.  Bars:
  * matrix entries within 4e-6 of the largest entry of the oracle's FP64 matrix (FP32 unit
    roundoff is 6e-8; an entry sums about 13 element-pair blocks of up to 16 products each),
    relative Frobenius error within 1e-6;
  * mixed-precision solutions (FP32 inner operator from the FP32 assembly, FP64 true residual and
    refinement) within 1e-8 relative L2 of the oracle's direct FP64 solution.
"""
import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api
from oracle import oracle as O
from tests import helpers as H
from tests.test_gpu_parity import IDS, PROBLEMS

pytestmark = pytest.mark.gpu

ENTRY_BOUND = 4e-6
FROB_BOUND = 1e-6


def _gpu_matrix(problem, ctx, convection=True):
    g = api.mesh_2d(problem.mesh, ctx)
    api.find_neighbours(g, H.radii(problem.models))
    inner = O.OracleMesh(problem.mesh).boundary(problem.bcs, problem.dof)[2]
    K = api._assemble(g, problem.physics, problem.models, problem.materials, inner)
    if convection and problem.physics == 0:
        K.add_convection(problem.bcs)
    return g, K


@pytest.mark.parametrize("problem", PROBLEMS, ids=IDS)
def test_fp32_assembly_entries(gpu_ctx, problem):
    print(f"[fp32 assembly] {problem.name}", flush=True)
    om, K, f, xD, inner = H.oracle_assembly(problem)
    g, Kg = _gpu_matrix(problem, gpu_ctx)
    Kg.assemble_fp32()
    o = K.export(0)
    q = Kg.export(3)
    assert np.array_equal(q.outer, o.outer) and np.array_equal(q.inner, o.inner)
    scale = np.abs(o.val).max()
    err = np.abs(q.val - o.val).max() / scale
    frob = np.linalg.norm(q.val - o.val) / np.linalg.norm(o.val)
    print(f"  nnz {o.val.size}: max entry error {err:.3e} of the largest entry, Frobenius {frob:.3e}", flush=True)
    assert err <= ENTRY_BOUND and frob <= FROB_BOUND
    # the FP32 pass is not the rounded FP64 matrix
    rounded = o.val.astype(np.float32).astype(np.float64)
    assert not np.array_equal(q.val, rounded), "FP32 values equal the rounded FP64 ones: no FP32 pass ran"
    # reproducible run to run (deterministic accumulation order)
    Kg.assemble_fp32()
    assert np.array_equal(Kg.export(3).val, q.val)


def test_fp32_assembly_repeats_boundary_additions(gpu_ctx):
    """Robin terms added before or after the FP32 pass end up in the FP32 values alike."""
    problem = PROBLEMS[1]                                   # Q8 with a convection boundary
    assert any(b.kind[0] == A.NLF_BC_CONVECTION for b in problem.bcs)
    _, before = _gpu_matrix(problem, gpu_ctx, convection=True)
    before.assemble_fp32()
    _, after = _gpu_matrix(problem, gpu_ctx, convection=False)
    after.assemble_fp32()
    after.add_convection(problem.bcs)
    a, b = before.export(4).val, after.export(4).val
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-6 * np.abs(a).max())
    _, none = _gpu_matrix(problem, gpu_ctx, convection=False)
    none.assemble_fp32()
    assert np.abs(none.export(4).val - a).max() > 1e-3 * np.abs(a).max(), "the convection terms are missing"


@pytest.mark.parametrize("problem", [PROBLEMS[0], PROBLEMS[3]], ids=[IDS[0], IDS[3]])
def test_fp32_mfma_path_matches_node_path(gpu_ctx, problem):
    """Thermal Q4: FP32 MFMA pair blocks against the node-centric FP32 contraction."""
    print(f"[fp32 mfma] {problem.name}", flush=True)
    g, K = _gpu_matrix(problem, gpu_ctx)
    K.assemble_fp32()
    a = K.export(4).val
    gpu_ctx.set_option(A.NLF_OPT_ASSEMBLY_PAIRS, 0)
    try:
        K.assemble_fp32()
        b = K.export(4).val
    finally:
        gpu_ctx.set_option(A.NLF_OPT_ASSEMBLY_PAIRS, 1)
    err = np.abs(a - b).max() / np.abs(b).max()
    print(f"  max difference {err:.3e} of the largest entry", flush=True)
    assert err <= ENTRY_BOUND
    assert not np.array_equal(a, b), "the two FP32 contractions round differently; equal values mean one ran twice"


MIXED_CASES = [A.NLF_PRECOND_ILLT, A.NLF_PRECOND_JACOBI]


@pytest.mark.parametrize("problem", PROBLEMS, ids=IDS)
@pytest.mark.parametrize("precond", MIXED_CASES)
def test_mixed_solution_with_fp32_assembly(gpu_ctx, problem, precond):
    print(f"[mixed fp32 assembly] {problem.name} precond={precond}", flush=True)
    om, K, f, xD, inner = H.oracle_assembly(problem)
    x_ref = H.direct_solution(K, f)
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    fn = api.stationary_heat_equation_solver_2d if problem.physics == 0 else api.equilibrium_equation
    sol = fn(g, problem.models, problem.materials, problem.bcs, precond=precond, tol=1e-12, mode=A.NLF_MODE_MIXED,
             assembly32=True)
    err = H.rel_l2(sol.x.cpu().numpy(), x_ref)
    print(f"  iterations {sol.iterations} true residual {sol.true_residual:.2e} rel err {err:.2e}", flush=True)
    assert err <= 1e-8


def test_fp32_spmv_of_assembled_values(gpu_ctx):
    """The FP32 SpMV reads the FP32-assembled values: its product is within FP32 accuracy of the
    oracle's FP64 product."""
    problem = PROBLEMS[3]
    om, K, f, xD, inner = H.oracle_assembly(problem)
    g, Kg = _gpu_matrix(problem, gpu_ctx)
    Kg.assemble_fp32()
    x = np.random.default_rng(3).standard_normal(f.size)
    y = Kg.spmv(torch.from_numpy(x).float().cuda()).double().cpu().numpy()
    assert H.rel_l2(y, K.spmv(x)) <= 1e-5


@pytest.mark.parametrize("problem", [PROBLEMS[0], PROBLEMS[3]], ids=[IDS[0], IDS[3]])
@pytest.mark.parametrize("precond", [A.NLF_PRECOND_JACOBI, A.NLF_PRECOND_ILLT])
def test_split_mode_solution(gpu_ctx, problem, precond):
    """NLF_MODE_SPLIT (values as FP32 + 16-bit remainder, FP64 vectors and true residual) reaches the
    FP64 answer.  The FP64 solve stops on the preconditioned residual (the reference's rule), the
    split one on the true residual, so the split solve is asked for the true residual the FP64 solve
    reached and must need about as many iterations."""
    print(f"[split] {problem.name} precond={precond}", flush=True)
    om, K, f, xD, inner = H.oracle_assembly(problem)
    x_ref = H.direct_solution(K, f)
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    fp64 = api.stationary_heat_equation_solver_2d(g, problem.models, problem.materials, problem.bcs, precond=precond,
                                                  tol=1e-12)
    tol = max(fp64.true_residual, 1e-14)
    sol = api.stationary_heat_equation_solver_2d(g, problem.models, problem.materials, problem.bcs, precond=precond,
                                                 tol=tol, mode=A.NLF_MODE_SPLIT)
    err = H.rel_l2(sol.x.cpu().numpy(), x_ref)
    print(f"  split: {sol.iterations} iterations to {tol:.2e} (FP64 {fp64.iterations}), true residual "
          f"{sol.true_residual:.2e}, rel err {err:.2e}", flush=True)
    assert sol.true_residual <= tol
    assert err <= 1e-9
    assert sol.iterations <= 1.15 * fp64.iterations + 16


def test_split_values_accuracy(gpu_ctx):
    """The split product (FP32 value + 16-bit remainder, FP64 x and sums) is four orders of magnitude
    closer to the FP64 product than the FP32-value product is."""
    from nonlocal_amd._lib import load
    problem = PROBLEMS[3]
    g, Kg = _gpu_matrix(problem, gpu_ctx)
    x = torch.from_numpy(np.random.default_rng(4).standard_normal(Kg.info()["rows"])).cuda()
    y64 = Kg.spmv(x).cpu().numpy()
    Kg.make_split()
    ys = torch.empty_like(x)
    yf = torch.empty_like(x)
    gpu_ctx.check(load().nlf_spmv_lpr(Kg.h, 16 | 0x2000, 0, api._dev(x), api._dev(ys)))   # split values
    Kg.make_fp32()
    Kg.make_split()                          # make_fp32 invalidates the split; rebuild it
    xf = x.float()
    yff = torch.empty_like(xf)
    gpu_ctx.check(load().nlf_spmv_lpr(Kg.h, 16, 1, api._dev(xf), api._dev(yff)))           # FP32 values
    yf = yff.double()
    es = H.rel_l2(ys.cpu().numpy(), y64)
    ef = H.rel_l2(yf.cpu().numpy(), y64)
    print(f"  split product error {es:.2e}, FP32 product error {ef:.2e}", flush=True)
    assert es <= 1e-11
    assert es * 1e4 <= ef
