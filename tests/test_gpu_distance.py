"""GPU parity of the distances other than the Euclidean one: powered_distance with n = 1, integer n,
real n and the maximum norm, and powered_distance_with_rotation (search_function.hpp:28-90), against
the oracle, through the C ABI.

* Neighbour lists (content and order) bit-exact, on a jittered mesh and on one centred on the
  origin, where the rotated distance of an element centred at (0, 0) is 0 / 0 and its list is empty.
* Matrices: the reference's upper CSR and bound part with the same pattern, entries within 1e-12 of
  the largest, for thermal Q4 (the MFMA pair path with the general influence), thermal Q8 and
  elastic Q4 (the node path).  With the rotated ellipse e' in N(e) does not imply e in N(e'); the
  device's full storage must then be the transpose-completed upper part, checked entry by entry.
* SpMV within 1e-13 and the solution within 1e-10 of the direct solution of the oracle's matrix;
  heat flux within 1e-12 of the largest flux.
"""
import dataclasses

import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, configs, meshgen
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

ROT = A.NLF_DISTANCE_ELLIPSE_WITH_ROTATION
SEARCH = [("n1", A.distance(A.NLF_DISTANCE_LP, 1)), ("n3", A.distance(A.NLF_DISTANCE_LP, 3)),
          ("n1.5", A.distance(A.NLF_DISTANCE_LP, 1.5)), ("inf", A.distance(A.NLF_DISTANCE_LP, "inf")),
          ("rotation", A.distance(ROT))]


def _centred(mesh):
    return dataclasses.replace(mesh, xy=mesh.xy - 0.5)


@pytest.mark.parametrize("name,dist", SEARCH, ids=[s[0] for s in SEARCH])
@pytest.mark.parametrize("where", ["jitter", "origin"])
def test_neighbours_bit_exact(gpu_ctx, name, dist, where):
    mesh = meshgen.structured_q4(31, jitter=0.2, seed=3) if where == "jitter" else _centred(meshgen.structured_q4(25))
    print(f"[distance neighbours] {name} {where}", flush=True)
    for strategy in (A.NLF_EXPAND_NO, A.NLF_EXPAND_MAX):
        g = api.mesh_2d(mesh, gpu_ctx)
        exp = api.find_neighbours(g, [[0.11, 0.07]], strategy, distances=[dist])
        off, ids = g.neighbours()
        ooff, oids, oexp = O.OracleMesh(mesh).find_neighbours([[0.11, 0.07]], strategy, distances=[dist])
        assert np.array_equal(exp, oexp)
        assert np.array_equal(off, ooff), "neighbour counts"
        assert np.array_equal(ids, oids), "neighbour lists (content and order)"
    if name == "rotation" and where == "origin":
        assert (np.diff(off) == 0).sum() == 1, "the element centred at the origin has an empty list"


def _thermal(mesh, model, name):
    mat = A.thermal_material(A.NLF_CONDUCTIVITY_ANISOTROPIC, (1.0, 0.25, 0.1))
    return configs.Problem(name, mesh, 0, [model], [mat], configs.thermal_bcs(robin_top=0.5), {})


def _problems():
    q4 = meshgen.structured_q4(24, jitter=0.2, seed=2)
    q4c = _centred(meshgen.structured_q4(24, jitter=0.2, seed=7))
    q8 = meshgen.structured_q8(12, jitter=0.2, seed=1)
    out = [
        _thermal(q4, A.model(0.4, (0.12, 0.09), n=1, p=2.0, q=1.0), "q4_n1_poly"),
        _thermal(q4, A.model(0.4, (0.12, 0.09), n=3, p=2.0, q=2.0), "q4_n3_bell"),
        _thermal(q4, A.model(0.4, 0.1, n="inf", influence=A.NLF_INFLUENCE_CONSTANT), "q4_inf_constant"),
        _thermal(q4, A.model(0.6, 0.08, n=1.5, influence=A.NLF_INFLUENCE_EXPONENTIAL, p=2.0, q=2.0), "q4_n1.5_exp"),
        _thermal(q4c, A.model(0.5, (0.15, 0.08), distance=ROT, p=2.0, q=1.0), "q4_rotation_fast"),
        _thermal(q4c, A.model(0.5, (0.12, 0.07), distance=ROT, p=2.0, q=2.0), "q4_rotation_bell"),
        _thermal(q8, A.model(0.3, 0.15, n=3, p=3.0, q=1.0), "q8_n3_p3"),
        _thermal(_centred(q8), A.model(0.3, (0.2, 0.1), distance=ROT, p=2.0, q=2.0), "q8_rotation_bell"),
    ]
    el = configs.cfg2(scale=0.08)
    out.append(dataclasses.replace(el, name="elastic_q4_inf", models=[A.model(0.3, 0.12, n="inf", p=2.0, q=2.0)]))
    out.append(dataclasses.replace(el, name="elastic_q4_rotation", mesh=dataclasses.replace(el.mesh, xy=el.mesh.xy + 0.1),
                                   models=[A.model(0.3, (0.12, 0.08), distance=ROT, p=2.0, q=2.0)]))
    return out


PROBLEMS = _problems()
IDS = [p.name for p in PROBLEMS]


def _assemble_both(problem, ctx):
    om, K, f, xD, inner = H.oracle_assembly(problem)
    g = api.mesh_2d(problem.mesh, ctx)
    api._search(g, problem.models)
    Kg = api._assemble(g, problem.physics, problem.models, problem.materials, inner)
    if problem.physics == 0:
        Kg.add_convection(problem.bcs)
    return om, K, f, inner, g, Kg


@pytest.mark.parametrize("problem", PROBLEMS, ids=IDS)
def test_assembly_and_spmv(gpu_ctx, problem):
    print(f"[distance assembly] {problem.name}", flush=True)
    om, K, f, inner, g, Kg = _assemble_both(problem, gpu_ctx)
    for p in (0, 1):
        o, q = K.export(p), Kg.export(p)
        assert np.array_equal(q.outer, o.outer), f"part {p} row offsets"
        assert np.array_equal(q.inner, o.inner), f"part {p} column indices"
        scale = max(np.abs(o.val).max(), 1e-300) if o.val.size else 1.0
        err = np.abs(q.val - o.val).max() / scale if o.val.size else 0.0
        print(f"  part {p}: nnz {o.val.size} max rel entry error {err:.3e}", flush=True)
        assert err <= 1e-12
    # the full storage is the upper part completed by its transpose (explicit zeros included)
    n = f.size
    up = Kg.export(0)
    ru = np.repeat(np.arange(n), np.diff(up.outer))
    off = up.inner != ru
    r_all = np.concatenate([ru, up.inner[off]])
    c_all = np.concatenate([up.inner, ru[off]])
    v_all = np.concatenate([up.val, up.val[off]])
    order = np.lexsort((c_all, r_all))
    full = Kg.export(2)
    print(f"  full storage nnz {full.inner.size} (upper part and its transpose {order.size})", flush=True)
    assert np.array_equal(np.repeat(np.arange(n), np.diff(full.outer)), r_all[order])
    assert np.array_equal(full.inner, c_all[order])
    assert np.array_equal(full.val, v_all[order]), "the lower part must mirror the upper part exactly"
    x = np.random.default_rng(0).standard_normal(n)
    y = Kg.spmv(torch.from_numpy(x).cuda()).cpu().numpy()
    assert H.rel_l2(y, K.spmv(x)) <= 1e-13


def test_rotation_neighbours_are_asymmetric(gpu_ctx):
    """The rotated-ellipse lists of the thermal cases are not a symmetric relation, so the transpose
    completion above is exercised."""
    problem = PROBLEMS[IDS.index("q4_rotation_bell")]
    om = O.OracleMesh(problem.mesh)
    off, ids, _ = om.find_neighbours(H.radii(problem.models), distances=H.distances(problem.models))
    pairs = {(e, int(j)) for e in range(problem.mesh.n_elems) for j in ids[off[e]:off[e + 1]]}
    assert any((j, e) not in pairs for e, j in pairs)


# The upper triangle of an asymmetric kernel read as a symmetric matrix (the reference's form for the
# rotated ellipse) need not be positive definite: the plane-stress case above has a negative
# eigenvalue (-1.9e-3 of 1.44), where CG breaks down, so the elastic solve takes the maximum norm.
SOLVE = [p for p in PROBLEMS if p.physics == 0][::2] + [PROBLEMS[IDS.index("elastic_q4_inf")]]


@pytest.mark.parametrize("problem", SOLVE, ids=[p.name for p in SOLVE])
def test_solution(gpu_ctx, problem):
    print(f"[distance solve] {problem.name}", flush=True)
    om, K, f, xD, inner = H.oracle_assembly(problem)
    x_ref = H.direct_solution(K, f)
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    fn = api.stationary_heat_equation_solver_2d if problem.physics == 0 else api.equilibrium_equation
    sol = fn(g, problem.models, problem.materials, problem.bcs, precond=A.NLF_PRECOND_ILLT, tol=1e-14,
             mode=A.NLF_MODE_FP64)
    err = H.rel_l2(sol.x.cpu().numpy(), x_ref)
    print(f"  iterations {sol.iterations} rel err {err:.2e}", flush=True)
    assert err <= 1e-10


@pytest.mark.parametrize("problem", [p for p in PROBLEMS if p.physics == 0], ids=[p.name for p in PROBLEMS if p.physics == 0])
def test_flux(gpu_ctx, problem):
    print(f"[distance flux] {problem.name}", flush=True)
    T = np.random.default_rng(21).standard_normal(problem.mesh.n_nodes)
    om = O.OracleMesh(problem.mesh)
    om.find_neighbours(H.radii(problem.models), distances=H.distances(problem.models))
    n_ref, q_ref = om.heat_flux(problem.models, problem.materials, T)
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    sol = api.heat_equation_solution_2d(g, problem.models, problem.materials, torch.from_numpy(T))
    fx, fy = sol.calc_flux()
    qx, qy = sol.flux_in_qnodes()
    scale = np.abs(q_ref).max()
    eq = max(np.abs(qx.cpu().numpy() - q_ref[0]).max(), np.abs(qy.cpu().numpy() - q_ref[1]).max()) / scale
    en = max(np.abs(fx.cpu().numpy() - n_ref[0]).max(), np.abs(fy.cpu().numpy() - n_ref[1]).max()) / scale
    print(f"  quadrature points {eq:.2e}, nodes {en:.2e}", flush=True)
    assert eq <= 1e-12 and en <= 1e-12
