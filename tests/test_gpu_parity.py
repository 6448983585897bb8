"""GPU parity of the hot path against the CPU oracle (through libnlf.so's C ABI).

Bit-exact: element centres, neighbour lists, CSR patterns (integer / index work).  FP64 values:
matrix entries within 1e-12 of the largest entry, solutions within 1e-10 relative L2 of the direct
FP64 solution of the oracle's matrix (mixed precision: 1e-8).
"""
import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, configs, meshgen
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _problems_small():
    out = [configs.cfg0()]
    p = configs.cfg1(scale=0.1)                      # 20 x 20 Q8, jittered, anisotropic, Robin
    p.models = [configs.bell(0.3, 0.15)]
    out.append(p)
    q = configs.cfg2(scale=0.1)                      # 30 x 30 elastic
    q.models = [configs.bell(0.3, 0.1)]
    out.append(q)
    out.append(configs.composite(48, r=0.08, count=4))
    return out


PROBLEMS = _problems_small()
IDS = [p.name for p in PROBLEMS]


def _gpu_mesh(problem, ctx):
    return api.mesh_2d(problem.mesh, ctx)


@pytest.mark.parametrize("jitter,nn", [(0.0, 4), (0.2, 4), (0.2, 8)])
def test_quadrature_and_centres(gpu_ctx, jitter, nn):
    print(f"[quadrature] nn={nn} jitter={jitter}", flush=True)
    mesh = meshgen.structured_q4(12, jitter, seed=5) if nn == 4 else meshgen.structured_q8(10, jitter, seed=5)
    g = api.mesh_2d(mesh, gpu_ctx)
    xq, jac, D, c = g.quadrature()
    om = O.OracleMesh(mesh)
    oxq, ojac, oD = om.quadrature()
    assert np.array_equal(c, om.centres()), "centres must be bit-exact"
    assert np.array_equal(c, meshgen.element_centres_numpy(mesh))
    np.testing.assert_allclose(xq, oxq, rtol=0, atol=1e-15)
    np.testing.assert_allclose(jac, ojac, rtol=1e-14, atol=0)
    np.testing.assert_allclose(D, oD, rtol=0, atol=1e-15 * np.abs(oD).max())


@pytest.mark.parametrize("problem", PROBLEMS, ids=IDS)
def test_neighbours_bit_exact(gpu_ctx, problem):
    print(f"[neighbours] {problem.name}", flush=True)
    g = _gpu_mesh(problem, gpu_ctx)
    exp = api.find_neighbours(g, H.radii(problem.models))
    off, ids = g.neighbours()
    om = O.OracleMesh(problem.mesh)
    ooff, oids, oexp = om.find_neighbours(H.radii(problem.models))
    assert np.array_equal(exp, oexp), "expanded radii must be bit-exact"
    assert np.array_equal(off, ooff)
    assert np.array_equal(ids, oids), "neighbour lists (content and order) must be bit-exact"


def test_neighbours_strategies(gpu_ctx):
    mesh = meshgen.structured_q4(20, 0.2, seed=9)
    for strategy in (A.NLF_EXPAND_NO, A.NLF_EXPAND_MAX, A.NLF_EXPAND_MIN):
        print(f"[neighbours strategy] {strategy}", flush=True)
        g = api.mesh_2d(mesh, gpu_ctx)
        exp = api.find_neighbours(g, [[0.12, 0.12]], strategy)
        off, ids = g.neighbours()
        ooff, oids, oexp = O.OracleMesh(mesh).find_neighbours([[0.12, 0.12]], strategy)
        assert np.array_equal(exp, oexp)
        assert np.array_equal(off, ooff) and np.array_equal(ids, oids)


def _assemble_both(problem, ctx, part=A.NLF_PART_FULL):
    om, K, f, xD, inner = H.oracle_assembly(problem, part=part)
    g = _gpu_mesh(problem, ctx)
    api.find_neighbours(g, H.radii(problem.models))
    Kg = api._assemble(g, problem.physics, problem.models, problem.materials, inner, part=part)
    if problem.physics == 0 and part == A.NLF_PART_FULL:
        Kg.add_convection(problem.bcs)
    return om, K, f, xD, inner, g, Kg


@pytest.mark.parametrize("problem", PROBLEMS, ids=IDS)
@pytest.mark.parametrize("part", [A.NLF_PART_FULL, A.NLF_PART_LOCAL])
def test_pattern_and_values(gpu_ctx, problem, part):
    print(f"[assembly] {problem.name} part={part}", flush=True)
    om, K, f, xD, inner, g, Kg = _assemble_both(problem, gpu_ctx, part)
    for p in (0, 1):
        o = K.export(p)
        q = Kg.export(p)
        assert np.array_equal(q.outer, o.outer), f"part {p} row offsets"
        assert np.array_equal(q.inner, o.inner), f"part {p} column indices"
        scale = max(np.abs(o.val).max(), 1e-300) if o.val.size else 1.0
        err = np.abs(q.val - o.val).max() / scale if o.val.size else 0.0
        print(f"  part {p}: nnz {o.val.size} max rel entry error {err:.3e}", flush=True)
        assert err <= 1e-12


@pytest.mark.parametrize("problem", [PROBLEMS[0], PROBLEMS[3]], ids=[IDS[0], IDS[3]])
def test_mfma_pair_path_matches_node_path(gpu_ctx, problem):
    """Thermal Q4: the MFMA element-pair contraction (default) against the node-centric FP64 one."""
    print(f"[mfma] {problem.name}", flush=True)
    g = _gpu_mesh(problem, gpu_ctx)
    api.find_neighbours(g, H.radii(problem.models))
    inner = O.OracleMesh(problem.mesh).boundary(problem.bcs, problem.dof)[2]
    a = api._assemble(g, 0, problem.models, problem.materials, inner).export(2)
    gpu_ctx.set_option(A.NLF_OPT_ASSEMBLY_PAIRS, 0)
    try:
        b = api._assemble(g, 0, problem.models, problem.materials, inner).export(2)
    finally:
        gpu_ctx.set_option(A.NLF_OPT_ASSEMBLY_PAIRS, 1)
    c = api._assemble(g, 0, problem.models, problem.materials, inner).export(2)
    assert np.array_equal(a.inner, b.inner)
    err = np.abs(a.val - b.val).max() / np.abs(b.val).max()
    print(f"  max rel difference {err:.3e}", flush=True)
    assert err <= 1e-13
    assert not np.array_equal(a.val, b.val), "the two contractions round differently; equal values mean one ran twice"
    assert np.array_equal(a.val, c.val), "the MFMA path must be bitwise reproducible"


@pytest.mark.parametrize("problem", [PROBLEMS[0], PROBLEMS[3]], ids=[IDS[0], IDS[3]])
def test_pair_chunks_match_single_pass(gpu_ctx, problem):
    """Thermal Q4 pair blocks assembled in node chunks (a 1 MiB budget) against the single pass:
    every value bitwise equal, the bound part included."""
    print(f"[pair chunks] {problem.name}", flush=True)
    g = _gpu_mesh(problem, gpu_ctx)
    api.find_neighbours(g, H.radii(problem.models))
    inner = O.OracleMesh(problem.mesh).boundary(problem.bcs, problem.dof)[2]
    whole = api._assemble(g, 0, problem.models, problem.materials, inner)
    pairs = g.neighbours()[1].size
    assert pairs * 16 * 8 > 2 << 20, "the problem must need several 1 MiB chunks"
    gpu_ctx.set_option(A.NLF_OPT_PAIR_BUDGET_MIB, 1)
    try:
        chunked = api._assemble(g, 0, problem.models, problem.materials, inner)
    finally:
        gpu_ctx.set_option(A.NLF_OPT_PAIR_BUDGET_MIB, 0)
    for part in (1, 2):
        a, b = whole.export(part), chunked.export(part)
        assert np.array_equal(a.inner, b.inner) and np.array_equal(a.val, b.val), f"part {part}"
    print(f"  {pairs} neighbour pairs, {pairs * 128 / 2**20:.1f} MiB of blocks: equal", flush=True)


@pytest.mark.parametrize("problem", PROBLEMS[:2], ids=IDS[:2])
def test_assembly_deterministic(gpu_ctx, problem):
    print(f"[determinism] {problem.name}", flush=True)
    g = _gpu_mesh(problem, gpu_ctx)
    api.find_neighbours(g, H.radii(problem.models))
    inner = O.OracleMesh(problem.mesh).boundary(problem.bcs, problem.dof)[2]
    a = api._assemble(g, problem.physics, problem.models, problem.materials, inner).export(2)
    b = api._assemble(g, problem.physics, problem.models, problem.materials, inner).export(2)
    assert np.array_equal(a.val, b.val), "two assemblies must be bitwise identical"


@pytest.mark.parametrize("problem", PROBLEMS, ids=IDS)
def test_spmv(gpu_ctx, problem):
    print(f"[spmv] {problem.name}", flush=True)
    om, K, f, xD, inner, g, Kg = _assemble_both(problem, gpu_ctx)
    x = np.random.default_rng(0).standard_normal(f.size)
    y_ref = K.spmv(x)
    y = Kg.spmv(torch.from_numpy(x).cuda()).cpu().numpy()
    assert H.rel_l2(y, y_ref) <= 1e-13
    Kg.make_fp32()
    y32 = Kg.spmv(torch.from_numpy(x).float().cuda()).double().cpu().numpy()
    assert H.rel_l2(y32, y_ref) <= 1e-5


def test_first_kind_rhs(gpu_ctx):
    problem = configs.cfg0()
    om, K, f, xD, inner, g, Kg = _assemble_both(problem, gpu_ctx)
    f0, xD0, inner0 = om.boundary(problem.bcs, 1)
    xD0 = xD0 + np.where(inner0 == 0, 0.25, 0.0)         # non-zero Dirichlet values
    ref = O.first_kind_rhs(K, inner0, xD0, f0)
    fd = torch.from_numpy(f0).cuda()
    Kg.apply_first_kind(torch.from_numpy(xD0).cuda(), fd)
    np.testing.assert_allclose(fd.cpu().numpy(), ref, rtol=1e-14, atol=1e-15)


def test_ic0_factor_matches_oracle(gpu_ctx):
    problem = configs.cfg0()
    om, KL, f, xD, inner, g, KLg = _assemble_both(problem, gpu_ctx, A.NLF_PART_LOCAL)
    P = O.OraclePrecond(A.NLF_PRECOND_ILLT, KL)
    Pg = api.preconditioner(KLg, A.NLF_PRECOND_ILLT)
    assert Pg.info()["ok"] and P.ok()[0]
    r = np.random.default_rng(1).standard_normal(f.size)
    z_ref = P.apply(r)
    z = Pg.solve(torch.from_numpy(r).cuda()).cpu().numpy()
    assert H.rel_l2(z, z_ref) <= 1e-12


SOLVE_CASES = [(A.NLF_PRECOND_ILLT, A.NLF_MODE_FP64, 1e-10), (A.NLF_PRECOND_JACOBI, A.NLF_MODE_FP64, 1e-10),
               (A.NLF_PRECOND_IDENTITY, A.NLF_MODE_FP64, 1e-10), (A.NLF_PRECOND_ILLT, A.NLF_MODE_MIXED, 1e-8),
               (A.NLF_PRECOND_JACOBI, A.NLF_MODE_MIXED, 1e-8)]


@pytest.mark.parametrize("problem", PROBLEMS, ids=IDS)
@pytest.mark.parametrize("precond,mode,bound", SOLVE_CASES)
def test_solution(gpu_ctx, problem, precond, mode, bound):
    print(f"[solve] {problem.name} precond={precond} mode={mode}", flush=True)
    om, K, f, xD, inner = H.oracle_assembly(problem)
    x_ref = H.direct_solution(K, f)
    g = _gpu_mesh(problem, gpu_ctx)
    tol = 1e-14 if mode == A.NLF_MODE_FP64 else 1e-12
    fn = api.stationary_heat_equation_solver_2d if problem.physics == 0 else api.equilibrium_equation
    sol = fn(g, problem.models, problem.materials, problem.bcs, precond=precond, tol=tol, mode=mode)
    err = H.rel_l2(sol.x.cpu().numpy(), x_ref)
    print(f"  iterations {sol.iterations} true residual {sol.true_residual:.2e} rel err {err:.2e}", flush=True)
    assert err <= bound


def test_cfg0_against_golden_fixture(gpu_ctx):
    """The GPU path on the 40 x 40 workload against the committed oracle fixture."""
    import hashlib
    import os
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz"))
    problem = configs.cfg0()
    g = _gpu_mesh(problem, gpu_ctx)
    sol = api.stationary_heat_equation_solver_2d(g, problem.models, problem.materials, problem.bcs,
                                                 precond=A.NLF_PRECOND_ILLT, tol=1e-14)
    off, ids = g.neighbours()
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(off).tobytes())
    h.update(np.ascontiguousarray(ids).tobytes())
    assert h.hexdigest() == str(G["cfg0_nb_sha"])
    assert H.rel_l2(sol.x.cpu().numpy(), G["cfg0_solution"]) <= 1e-10


@pytest.mark.parametrize("problem", PROBLEMS, ids=IDS)
def test_row_pattern_dictionary(gpu_ctx, problem):
    """The SpMV through the shared row-pattern dictionary equals the per-nonzero-offset SpMV bit for
    bit (same products in the same order); the dictionary is in use on the uniform meshes."""
    from nonlocal_amd._lib import load
    print(f"[row patterns] {problem.name}", flush=True)
    om, K, f, xD, inner, g, Kg = _assemble_both(problem, gpu_ctx)
    groups, entries = Kg.pattern_info()
    nnz = Kg.info()["nnz_full"]
    print(f"  {groups} distinct rows, {entries} entries for {nnz} nonzeros", flush=True)
    if problem.name.startswith("cfg0"):
        assert 0 < entries * 4 <= nnz, "a uniform one-material Q4 mesh must use the dictionary"
    x = torch.from_numpy(np.random.default_rng(2).standard_normal(f.size)).cuda()
    y_dict = torch.empty_like(x)
    y_off = torch.empty_like(x)
    for lpr in (4, 16, 32):
        gpu_ctx.check(load().nlf_spmv_lpr(Kg.h, lpr, 0, api._dev(x), api._dev(y_dict)))
        gpu_ctx.check(load().nlf_spmv_lpr(Kg.h, lpr | 0x1000, 0, api._dev(x), api._dev(y_off)))
        assert torch.equal(y_dict, y_off)
    assert H.rel_l2(y_dict.cpu().numpy(), K.spmv(x.cpu().numpy())) <= 1e-13
    gpu_ctx.set_option(A.NLF_OPT_ROW_PATTERNS, 0)
    try:
        Kn = api._assemble(g, problem.physics, problem.models, problem.materials, inner)
    finally:
        gpu_ctx.set_option(A.NLF_OPT_ROW_PATTERNS, 1)
    assert Kn.pattern_info() == (0, 0)
