"""GPU parity beyond the bell influence and the small radii of test_gpu_parity.py:

* every diameter-expanding strategy (NO, MAX, MIN, MEAN), bit-exact, on one- and multi-group meshes;
* every influence family in the assembly (fast polynomial, bell, general polynomial with p = 4,
  q = 3 and with a non-integer q, exponential, constant), through the MFMA pair kernel (thermal Q4)
  and the node path (Q8, elasticity), against the oracle's entries;
* meshes whose elements start at different corners (a node then holds one local index in two
  elements), whose assembly accumulates in rank rounds;
* convection with a nonzero ambient temperature;
* the configurations' ratio of influence radius to element size (r/h ~ 10, 300-400 neighbours per
  element), entries against the oracle and solutions against the oracle's own solve.
"""
import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, batch, configs, meshgen
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

STRATEGIES = [A.NLF_EXPAND_NO, A.NLF_EXPAND_MAX, A.NLF_EXPAND_MIN, A.NLF_EXPAND_MEAN]


def _meshes():
    return [("q4_jitter", meshgen.structured_q4(22, 0.2, seed=9), [[0.12, 0.1]]),
            ("q8_jitter", meshgen.structured_q8(10, 0.2, seed=3), [[0.2, 0.2]]),
            ("composite", meshgen.composite_q4(40, count=4, seed=7)[0], None),
            ("q4_rotated", H.rotate_local_order(meshgen.structured_q4(18, 0.15, seed=2)), [[0.15, 0.15]])]


@pytest.mark.parametrize("strategy", STRATEGIES)
@pytest.mark.parametrize("case", range(4), ids=[c[0] for c in _meshes()])
def test_neighbours_every_strategy_bit_exact(gpu_ctx, strategy, case):
    name, mesh, radii = _meshes()[case]
    radii = radii or [[0.09, 0.09]] * mesh.n_groups
    print(f"[strategy] {name} strategy={strategy}", flush=True)
    g = api.mesh_2d(mesh, gpu_ctx)
    exp = api.find_neighbours(g, radii, strategy)
    off, ids = g.neighbours()
    ooff, oids, oexp = O.OracleMesh(mesh).find_neighbours(radii, strategy)
    assert np.array_equal(exp, oexp), "expanded radii must be bit-exact"
    assert np.array_equal(off, ooff) and np.array_equal(ids, oids)


# ---- influence families ------------------------------------------------------------------------
FAMILIES = {
    "fast_polynomial": dict(influence=A.NLF_INFLUENCE_POLYNOMIAL, p=2.0, q=1.0),
    "bell": dict(influence=A.NLF_INFLUENCE_POLYNOMIAL, p=2.0, q=2.0),
    "polynomial_p4_q3": dict(influence=A.NLF_INFLUENCE_POLYNOMIAL, p=4.0, q=3.0),
    "polynomial_p2_q1.5": dict(influence=A.NLF_INFLUENCE_POLYNOMIAL, p=2.0, q=1.5),
    "exponential": dict(influence=A.NLF_INFLUENCE_EXPONENTIAL, p=2.0, q=0.5),
    "constant": dict(influence=A.NLF_INFLUENCE_CONSTANT, p=2.0, q=1.0),
}


def _family_problem(kind, family, rotated=False):
    f = FAMILIES[family]
    if kind == "q4":
        p = configs.composite(24, r=0.15, count=2)
        p.models = [A.model(0.4, 0.15, **f) for _ in p.models]
    elif kind == "q8":
        p = configs.cfg1(scale=0.05)                       # 10 x 10 jittered Q8, anisotropic, Robin
        p.models = [A.model(0.3, 0.25, **f)]
    else:
        p = configs.cfg2(scale=0.04)                       # 12 x 12 elastic
        p.models = [A.model(0.3, 0.25, **f)]
    if rotated:
        p.mesh = H.rotate_local_order(p.mesh)
    p.name = f"{kind}_{family}" + ("_rotated" if rotated else "")
    return p


def _compare_entries(problem, ctx, bound=1e-12):
    om, K, f, xD, inner = H.oracle_assembly(problem)
    g = api.mesh_2d(problem.mesh, ctx)
    api.find_neighbours(g, H.radii(problem.models))
    Kg = api._assemble(g, problem.physics, problem.models, problem.materials, inner)
    if problem.physics == 0:
        Kg.add_convection(problem.bcs)
    for part in (0, 1):
        o, q = K.export(part), Kg.export(part)
        assert np.array_equal(q.outer, o.outer) and np.array_equal(q.inner, o.inner), f"part {part} pattern"
        if o.val.size:
            err = np.abs(q.val - o.val).max() / np.abs(o.val).max()
            print(f"  part {part}: nnz {o.val.size} max rel entry error {err:.3e}", flush=True)
            assert err <= bound, f"part {part}"
    return om, K, f, inner, g, Kg


@pytest.mark.parametrize("family", list(FAMILIES))
@pytest.mark.parametrize("kind", ["q4", "q8", "elastic"])
def test_assembly_every_influence_family(gpu_ctx, kind, family):
    problem = _family_problem(kind, family)
    print(f"[influence] {problem.name}", flush=True)
    _compare_entries(problem, gpu_ctx)


@pytest.mark.parametrize("family", ["fast_polynomial", "polynomial_p4_q3", "exponential", "constant"])
def test_q4_pair_kernel_matches_node_path(gpu_ctx, family):
    """Thermal Q4 with each influence class: the MFMA pair kernel (k_pairs_mfma<1> for the fast
    polynomial, <0> for the others) against the node-centric FP64 contraction."""
    problem = _family_problem("q4", family)
    print(f"[pairs vs node] {problem.name}", flush=True)
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    api.find_neighbours(g, H.radii(problem.models))
    inner = O.OracleMesh(problem.mesh).boundary(problem.bcs, 1)[2]
    a = api._assemble(g, 0, problem.models, problem.materials, inner).export(2)
    gpu_ctx.set_option(A.NLF_OPT_ASSEMBLY_PAIRS, 0)
    try:
        b = api._assemble(g, 0, problem.models, problem.materials, inner).export(2)
    finally:
        gpu_ctx.set_option(A.NLF_OPT_ASSEMBLY_PAIRS, 1)
    assert np.array_equal(a.inner, b.inner)
    err = np.abs(a.val - b.val).max() / np.abs(b.val).max()
    print(f"  max rel difference {err:.3e}", flush=True)
    assert err <= 1e-13


# ---- meshes whose elements start at different corners ------------------------------------------
@pytest.mark.parametrize("kind", ["q4", "q8", "elastic"])
def test_rotated_local_order_assembly(gpu_ctx, kind):
    problem = _family_problem(kind, "bell", rotated=True)
    assert not H.is_injective(problem.mesh)
    print(f"[rotated] {problem.name}", flush=True)
    om, K, f, inner, g, Kg = _compare_entries(problem, gpu_ctx)
    a = Kg.export(2)
    b = api._assemble(g, problem.physics, problem.models, problem.materials, inner)
    if problem.physics == 0:
        b.add_convection(problem.bcs)
    b = b.export(2)
    assert np.array_equal(a.val, b.val), "rank-round accumulation must be bitwise reproducible"
    # the solution: same integrals as the unrotated mesh, so the same temperatures
    fn = api.stationary_heat_equation_solver_2d if problem.physics == 0 else api.equilibrium_equation
    sol = fn(g, problem.models, problem.materials, problem.bcs, tol=1e-14)
    x_ref = H.direct_solution(K, f)
    err = H.rel_l2(sol.x.cpu().numpy(), x_ref)
    print(f"  solution rel err {err:.2e}", flush=True)
    assert err <= 1e-10


def test_rotated_mesh_matches_unrotated_solution(gpu_ctx):
    """The element integrals do not depend on the starting corner: the rotated and the plain mesh give
    the same temperatures (to rounding; the centres may move by an ulp, so the radius is chosen away
    from lattice distances)."""
    plain = _family_problem("q4", "bell")
    rot = _family_problem("q4", "bell", rotated=True)
    xs = []
    for p in (plain, rot):
        g = api.mesh_2d(p.mesh, gpu_ctx)
        xs.append(api.stationary_heat_equation_solver_2d(g, p.models, p.materials, p.bcs, tol=1e-14).x.cpu().numpy())
    assert H.rel_l2(xs[1], xs[0]) <= 1e-10


# ---- convection with an ambient temperature ------------------------------------------------------
def test_convection_with_ambient_temperature(gpu_ctx):
    problem = configs.cfg1(scale=0.1)                      # 20 x 20 jittered Q8, anisotropic
    problem.models = [configs.bell(0.3, 0.15)]
    problem.bcs[3] = A.boundary(A.NLF_BC_CONVECTION, heat_transfer=0.5, ambient=1.5)
    problem.bcs[2] = A.boundary(A.NLF_BC_CONVECTION, heat_transfer=2.0, ambient=-0.25)
    print("[convection] ambient 1.5 on y = 1, -0.25 on y = 0", flush=True)
    om, K, f, inner, g, Kg = _compare_entries(problem, gpu_ctx)
    fg, xDg, innerg = api.boundary_data(g, 1, problem.bcs)
    fo, xDo, innero = om.boundary(problem.bcs, 1)
    assert np.array_equal(innerg, innero)
    np.testing.assert_allclose(fg, fo, rtol=0, atol=1e-15 * np.abs(fo).max())
    assert np.abs(fo).max() > 0.01                          # h T_a int N_i on the Robin edges
    sol = api.stationary_heat_equation_solver_2d(g, problem.models, problem.materials, problem.bcs, tol=1e-14)
    err = H.rel_l2(sol.x.cpu().numpy(), H.direct_solution(K, f))
    print(f"  solution rel err {err:.2e}", flush=True)
    assert err <= 1e-10


# ---- the configurations' r/h --------------------------------------------------------------------
def _rh_problems():
    c = configs.composite(100, r=0.1)                       # r/h = 10, as the 1000 x 1000 composite
    c.name = "composite_100_r0.1"
    q8 = configs.cfg1(scale=0.2)                            # 40 x 40 jittered Q8, r/h = 10
    q8.models = [configs.bell(0.3, 0.25)]
    q8.bcs[3] = A.boundary(A.NLF_BC_CONVECTION, heat_transfer=0.5, ambient=0.75)
    q8.name = "q8_40_r0.25"
    el = configs.cfg2(scale=0.1)                            # 30 x 30 elastic, r/h = 9
    el.models = [configs.bell(0.3, 0.3)]
    el.name = "elastic_30_r0.3"
    return [c, q8, el]


@pytest.mark.parametrize("case", range(3), ids=[p.name for p in _rh_problems()])
def test_configuration_radius_entries_and_solution(gpu_ctx, case):
    problem = _rh_problems()[case]
    print(f"[r/h] {problem.name}", flush=True)
    om, K, f, inner, g, Kg = _compare_entries(problem, gpu_ctx)
    off, ids = g.neighbours()
    print(f"  neighbours per element: max {np.diff(off).max()}, mean {np.diff(off).mean():.0f}", flush=True)
    assert np.diff(off).max() >= 300
    x_ref = H.oracle_cg_solution(problem, om, K, f, inner)
    fn = api.stationary_heat_equation_solver_2d if problem.physics == 0 else api.equilibrium_equation
    for precond, mode, bound in [(A.NLF_PRECOND_ILLT, A.NLF_MODE_FP64, 1e-10),
                                 (A.NLF_PRECOND_JACOBI, A.NLF_MODE_MIXED, 1e-8)]:
        sol = fn(g, problem.models, problem.materials, problem.bcs, precond=precond, mode=mode,
                 tol=1e-14 if mode == A.NLF_MODE_FP64 else 1e-12)
        x = sol.x.cpu().numpy()
        err = H.rel_l2(x, x_ref)
        res = np.linalg.norm(f - K.spmv(x)) / np.linalg.norm(f)
        print(f"  precond {precond} mode {mode}: it {sol.iterations}, rel err {err:.2e}, "
              f"residual on the oracle's system {res:.2e}", flush=True)
        assert err <= bound
    if problem.physics == 0:
        sol = fn(g, problem.models, problem.materials, problem.bcs, precond=A.NLF_PRECOND_LEARNED, tol=1e-12)
        err = H.rel_l2(sol.x.cpu().numpy(), x_ref)
        print(f"  learned FCG: it {sol.iterations}, rel err {err:.2e}", flush=True)
        assert err <= 1e-8


def test_batch_at_configuration_radius(gpu_ctx):
    """Eight instances of side 32-64 with r/h from 3 to 10, assembled and solved as one batch."""
    rng = np.random.default_rng(21)
    probs = []
    for k in range(8):
        n = int(rng.integers(32, 65))
        rh = float(rng.uniform(3.0, 10.0)) if k else 10.0
        probs.append(configs.thermal_q4(n, float(rng.uniform(0.2, 0.9)), rh / n, float(rng.uniform(0.5, 2.0)),
                                        name=f"rb{k}"))
    print(f"[batch r/h] sides {[p.mesh.n for p in probs]}", flush=True)
    bs = batch.BatchSolver(probs, gpu_ctx)
    n = bs.node_begin
    csr = bs.K.export(0)
    refs = []
    for i, p in enumerate(probs):
        om, K, f, xD, inner = H.oracle_assembly(p)
        Ko = K.export(0)
        r0, r1 = n[i], n[i + 1]
        assert np.array_equal(csr.outer[r0:r1 + 1] - csr.outer[r0], Ko.outer)
        assert np.array_equal(csr.inner[csr.outer[r0]:csr.outer[r1]] - r0, Ko.inner)
        err = np.abs(csr.val[csr.outer[r0]:csr.outer[r1]] - Ko.val).max() / np.abs(Ko.val).max()
        assert err <= 1e-12
        refs.append(H.oracle_cg_solution(p, om, K, f, inner))
    for mode, bound in [(A.NLF_MODE_FP64, 1e-10), (A.NLF_MODE_MIXED, 1e-8)]:
        x, its, res, st = bs.solve(tol=1e-12, mode=mode)
        errs = [H.rel_l2(xi.cpu().numpy(), ref) for xi, ref in zip(bs.split(x), refs)]
        print(f"  mode {mode}: iterations {its.tolist()}, max rel err {max(errs):.2e}", flush=True)
        assert max(errs) <= bound


def test_export_rows_matches_full_export(gpu_ctx):
    problem = configs.cfg0()
    om, K, f, inner, g, Kg = _compare_entries(problem, gpu_ctx)
    full = Kg.export(0)
    o = K.export_rows(100, 260)
    for r0, r1 in [(0, 1), (100, 260), (1600, 1681), (500, 500)]:
        part = Kg.export_rows(r0, r1)
        b, e = full.outer[r0], full.outer[r1]
        assert np.array_equal(part.outer, full.outer[r0:r1 + 1] - b)
        assert np.array_equal(part.inner, full.inner[b:e]) and np.array_equal(part.val, full.val[b:e])
    part = Kg.export_rows(100, 260)
    assert np.array_equal(part.outer, o.outer) and np.array_equal(part.inner, o.inner)


def test_combine_outside_pattern_leaves_matrix_untouched(gpu_ctx):
    """K := alpha K + beta C with C's pattern outside K's: an argument error, and K keeps its values."""
    from nonlocal_amd import _abi
    problem = configs.cfg0(scale=0.5)
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    api.find_neighbours(g, H.radii(problem.models))
    inner = O.OracleMesh(problem.mesh).boundary(problem.bcs, 1)[2]
    KL = api._assemble(g, 0, problem.models, problem.materials, inner, part=A.NLF_PART_LOCAL)
    K = api._assemble(g, 0, problem.models, problem.materials, inner)
    before = KL.export(2)
    with pytest.raises(api.NLFError) as err:
        api.combine(KL, 2.0, K, 1.0)
    assert err.value.code == _abi.NLF_ERR_ARGUMENT
    after = KL.export(2)
    assert np.array_equal(before.val, after.val)
