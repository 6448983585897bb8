"""GPU parity of the Neumann problem (nlf_assemble_ex with the Lagrange row and column, CG on the
indefinite bordered system) and of coefficients per quadrature point, against the oracle; and the
reference's flux_stability known answers on the GPU."""
import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, meshgen
from oracle import oracle as O
from tests import helpers as H
from tests.test_neumann_oracle import THERMAL_KINDS, _plate

pytestmark = pytest.mark.gpu


def _compare(om, Ko, g, physics, models, mats, inner, neumann=False, coefficients=None, part=A.NLF_PART_FULL):
    Kg = api._assemble(g, physics, models, mats, inner, part=part, neumann=neumann, coefficients=coefficients)
    for p in (0, 1):
        o, q = Ko.export(p), Kg.export(p)
        assert np.array_equal(o.outer, q.outer) and np.array_equal(o.inner, q.inner), f"part {p} pattern"
        if o.val.size:
            err = np.abs(q.val - o.val).max() / np.abs(o.val).max()
            print(f"  part {p}: nnz {o.val.size} max rel entry error {err:.2e}", flush=True)
            assert err <= 1e-12
    return Kg


@pytest.mark.parametrize("nonlocal_", [False, True], ids=["local", "nonlocal"])
@pytest.mark.parametrize("nn", [4, 8])
def test_neumann_matrix_and_solution(gpu_ctx, nonlocal_, nn):
    mesh = meshgen.structured_q4(12, 0.15, seed=8) if nn == 4 else meshgen.structured_q8(7, 0.15, seed=8)
    bcs = [A.boundary(A.NLF_BC_FLUX, -1.0), A.boundary(A.NLF_BC_FLUX, 1.0), A.boundary(), A.boundary()]
    model = A.model(0.4, 0.2, p=2, q=2) if nonlocal_ else A.model(1.0, 0.0)
    print(f"[neumann] nn={nn} nonlocal={nonlocal_}", flush=True)
    res = O.thermal_problem(mesh, [model], [A.thermal_material()], bcs, tol=1e-15, energy=0.25)
    x_ref = H.direct_solution(res["K"], res["f"])
    g = api.mesh_2d(mesh, gpu_ctx)
    api.find_neighbours(g, H.radii([model]))
    inner = np.ones(mesh.n_nodes, np.uint8)
    _compare(res["mesh"], res["K"], g, 0, [model], [A.thermal_material()], inner, neumann=True)
    for precond in (A.NLF_PRECOND_ILLT, A.NLF_PRECOND_JACOBI):
        sol = api.stationary_heat_equation_solver_2d(g, [model], [A.thermal_material()], bcs, precond=precond,
                                                     tol=1e-15, energy=0.25)
        x = sol.x.cpu().numpy()
        assert x.size == mesh.n_nodes + 1
        err = H.rel_l2(x, x_ref)
        print(f"  precond {precond}: iterations {sol.iterations}, rel err {err:.2e}", flush=True)
        assert err <= 1e-10
        assert abs(H.integrate_nodes(res["mesh"], mesh, x[:mesh.n_nodes]) - 0.25) < 1e-12


def test_flux_stability_known_answers_on_the_gpu(gpu_ctx):
    """flux_stability.cpp:35-61 through the GPU driver.

    Parity first: the GPU flux of the oracle's direct solution equals the oracle's flux of it.  The
    flux is a gradient of a field spanning +-5.12 on elements of h = 0.125, so one evaluation carries
    about 8 |T| eps / h = 3.5e-14 of rounding, and the GPU's fused multiply-adds round differently
    from the oracle's unfused build: the parity bar is 1e-13.  Then the known answers at the bars that
    rounding floor allows: centre-line temperature 0 (3e-14) and flux -1 (1e-13), temperature
    integral 0 (the reference's 3.3e-14) and x-flux integral -10 (1e-12, 1e-13 of its size: the
    1e-15 stop criterion of CG leaves a few 1e-14 on the temperature at the plate's ends)."""
    mesh, models, mats, bcs, values = _plate()
    res = O.thermal_problem(mesh, models, mats, bcs, precond=A.NLF_PRECOND_ILLT, tol=1e-15, values=values)
    om = res["mesh"]
    direct = H.direct_solution(res["K"], res["f"])[:mesh.n_nodes]
    g = api.mesh_2d(mesh, gpu_ctx)
    q_o, _ = om.heat_flux(models, mats, direct)
    post = api.heat_equation_solution_2d(g, models, mats, torch.from_numpy(direct).to(gpu_ctx.device))
    q_g = [c.cpu().numpy() for c in post.calc_flux()]
    parity = max(np.abs(q_g[d] - q_o[d]).max() for d in (0, 1))
    sol = api.stationary_heat_equation_solver_2d(g, models, mats, bcs, precond=A.NLF_PRECOND_ILLT, tol=1e-15,
                                                 values=values)
    T = sol.x[:mesh.n_nodes].cpu().numpy()
    fx = sol.solution.calc_flux()[0].cpu().numpy()
    centre = np.abs(mesh.xy[:, 0]) < np.finfo(float).eps
    t_err = np.abs(T[centre]).max()
    f_err = np.abs(fx[centre] + 1.0).max()
    ti = H.integrate_nodes(om, mesh, T)
    fi = H.integrate_nodes(om, mesh, fx) + 10.0
    print(f"[flux_stability] flux parity {parity:.1e}; iterations {sol.iterations}: |T| {t_err:.1e}, "
          f"|q_x + 1| {f_err:.1e}, int T {ti:.1e}, int q_x + 10 {fi:.1e}", flush=True)
    assert parity <= 1e-13
    assert t_err <= 3e-14 and f_err <= 1e-13
    assert abs(ti) <= 3.3e-14 and abs(fi) <= 1e-12


def test_unbalanced_fluxes_are_refused(gpu_ctx):
    mesh = meshgen.structured_q4(6)
    bcs = [A.boundary(A.NLF_BC_FLUX, -0.5), A.boundary(A.NLF_BC_FLUX, 0.6), A.boundary(), A.boundary()]
    g = api.mesh_2d(mesh, gpu_ctx)
    with pytest.raises(api.NLFError, match="unsolvable Neumann"):
        api.stationary_heat_equation_solver_2d(g, [A.model(1.0, 0.0)], [A.thermal_material()], bcs)


@pytest.mark.parametrize("kind", list(THERMAL_KINDS))
@pytest.mark.parametrize("nn", [4, 8])
def test_conductivity_per_point(gpu_ctx, kind, nn):
    mesh = meshgen.structured_q4(10, 0.2, seed=3) if nn == 4 else meshgen.structured_q8(6, 0.2, seed=3)
    nq = 4 if nn == 4 else 9
    nc = THERMAL_KINDS[kind]
    rng = np.random.default_rng(13)
    coeff = rng.uniform(0.5, 2.0, size=(mesh.n_elems, nq, nc))
    if nc == 3:
        coeff[..., 2] *= 0.2
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0), A.boundary(A.NLF_BC_FLUX, 1.0), A.boundary(),
           A.boundary(A.NLF_BC_CONVECTION, heat_transfer=0.5, ambient=1.0)]
    mat = A.thermal_material(kind, (1.0, 1.0, 0.0))
    print(f"[coefficients] conductivity kind {kind} nn={nn}", flush=True)
    res = O.thermal_problem(mesh, [A.model(1.0, 0.0)], [mat], bcs, tol=1e-15, coefficients=[coeff])
    g = api.mesh_2d(mesh, gpu_ctx)
    om = O.OracleMesh(mesh)
    om.find_neighbours([[0.0, 0.0]])
    f, xD, inner = om.boundary(bcs, 1)
    Ko = om.assemble(0, [A.model(1.0, 0.0)], [mat], inner, coefficients=[coeff])
    _compare(om, Ko, g, 0, [A.model(1.0, 0.0)], [mat], inner, coefficients=[coeff])
    sol = api.stationary_heat_equation_solver_2d(g, [A.model(1.0, 0.0)], [mat], bcs, tol=1e-15, coefficients=[coeff])
    err = H.rel_l2(sol.x.cpu().numpy(), H.direct_solution(res["K"], res["f"]))
    print(f"  solution rel err {err:.2e}", flush=True)
    assert err <= 1e-10


@pytest.mark.parametrize("name", ["isotropic", "orthotropic", "anisotropic_0.4"])
def test_hooke_per_point(gpu_ctx, name):
    from tests.test_elastic_oracle import MATERIALS
    mesh = meshgen.structured_q8(5, 0.2, seed=6)
    nq = 9
    base = MATERIALS[name]
    kind, h = O.hooke(base)
    rng = np.random.default_rng(14)
    coeff = h[None, None, :] * rng.uniform(0.5, 2.0, size=(mesh.n_elems, nq, 1))
    inner = np.ones(2 * mesh.n_nodes, np.uint8)
    inner[np.ravel([[2 * v, 2 * v + 1] for v in mesh.boundary_nodes("left")])] = 0
    print(f"[coefficients] Hooke {name}", flush=True)
    om = O.OracleMesh(mesh)
    om.find_neighbours([[0.0, 0.0]])
    Ko = om.assemble(1, [A.model(1.0, 0.0)], [base], inner, coefficients=[coeff])
    g = api.mesh_2d(mesh, gpu_ctx)
    _compare(om, Ko, g, 1, [A.model(1.0, 0.0)], [base], inner, coefficients=[coeff])


def test_coefficients_per_point_of_a_nonlocal_problem_are_refused(gpu_ctx):
    mesh = meshgen.structured_q4(6)
    g = api.mesh_2d(mesh, gpu_ctx)
    api.find_neighbours(g, [[0.3, 0.3]])
    with pytest.raises(api.NLFError) as err:
        api._assemble(g, 0, [A.model(0.5, 0.3)], [A.thermal_material()], np.ones(mesh.n_nodes, np.uint8),
                      coefficients=[np.ones((mesh.n_elems, 4, 1))])
    assert err.value.code == A.NLF_ERR_UNSUPPORTED
