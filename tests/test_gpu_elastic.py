"""GPU parity of orthotropic and anisotropic plane elasticity and of the thermal load
(nlf_assemble physics 1, nlf_strain_stress with a temperature change, nlf_temperature_condition)
against the oracle: matrix entries, thermal loads and stresses within 1e-12 of the largest value,
solutions within 1e-10, and the stress-free thermal expansion patch on the GPU."""
import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, meshgen
from oracle import oracle as O
from tests import helpers as H
from tests.test_elastic_oracle import EXPANSIONS, MATERIALS, _all_dirichlet, _with_expansion, free_expansion

pytestmark = pytest.mark.gpu


def _problem(nn, nonlocal_):
    mesh = meshgen.structured_q4(14, 0.2, seed=41) if nn == 4 else meshgen.structured_q8(8, 0.2, seed=42)
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0, A.NLF_BC_FIRST_KIND, 0.0), A.boundary(A.NLF_BC_FLUX, 0.7, A.NLF_BC_FLUX, -0.2),
           A.boundary(), A.boundary(A.NLF_BC_NONE, 0.0, A.NLF_BC_FLUX, 0.3)]
    model = A.model(0.35, 0.22, p=2, q=2) if nonlocal_ else A.model(1.0, 0.0)
    return mesh, model, bcs


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("name", list(MATERIALS))
@pytest.mark.parametrize("nn", [4, 8])
@pytest.mark.parametrize("nonlocal_", [False, True], ids=["local", "nonlocal"])
def test_stiffness_every_hooke_kind(gpu_ctx, name, nn, nonlocal_):
    mesh, model, bcs = _problem(nn, nonlocal_)
    print(f"[elastic] {name} nn={nn} nonlocal={nonlocal_}", flush=True)
    om = O.OracleMesh(mesh)
    om.find_neighbours(H.radii([model]))
    f, xD, inner = om.boundary(bcs, 2)
    K = om.assemble(1, [model], [MATERIALS[name]], inner)
    g = api.mesh_2d(mesh, gpu_ctx)
    api.find_neighbours(g, H.radii([model]))
    Kg = api._assemble(g, 1, [model], [MATERIALS[name]], inner)
    for part in (0, 1):
        o, q = K.export(part), Kg.export(part)
        assert np.array_equal(o.outer, q.outer) and np.array_equal(o.inner, q.inner)
        if o.val.size:
            err = _rel(q.val, o.val)
            print(f"  part {part}: max rel entry error {err:.2e}", flush=True)
            assert err <= 1e-12
    if name.startswith("anisotropic") and nonlocal_:
        f = O.first_kind_rhs(K, inner, xD, f)
        sol = api.equilibrium_equation(g, [model], [MATERIALS[name]], bcs, tol=1e-14)
        err = H.rel_l2(sol.x.cpu().numpy(), H.direct_solution(K, f))
        print(f"  solution rel err {err:.2e}", flush=True)
        assert err <= 1e-10


@pytest.mark.parametrize("ename", list(EXPANSIONS))
@pytest.mark.parametrize("name", ["isotropic", "orthotropic", "anisotropic_0.4"])
@pytest.mark.parametrize("nn", [4, 8])
@pytest.mark.parametrize("nonlocal_", [False, True], ids=["local", "nonlocal"])
def test_thermal_load_and_stresses(gpu_ctx, ename, name, nn, nonlocal_):
    mesh, model, bcs = _problem(nn, nonlocal_)
    mat = _with_expansion(MATERIALS[name], EXPANSIONS[ename])
    print(f"[thermal] {name} expansion {ename} nn={nn} nonlocal={nonlocal_}", flush=True)
    rng = np.random.default_rng(7)
    dT = rng.uniform(-1.0, 2.0, mesh.n_nodes)
    om = O.OracleMesh(mesh)
    om.find_neighbours(H.radii([model]))
    f_ref = om.temperature_condition([model], [mat], dT)
    g = api.mesh_2d(mesh, gpu_ctx)
    api.find_neighbours(g, H.radii([model]))
    f = torch.zeros(2 * mesh.n_nodes, dtype=torch.float64, device=gpu_ctx.device)
    api.temperature_condition(g, [model], [mat], dT, f)
    err = _rel(f.cpu().numpy(), f_ref)
    print(f"  thermal load max rel error {err:.2e}", flush=True)
    assert err <= 1e-12
    u = rng.standard_normal(2 * mesh.n_nodes) * 1e-2
    ns, nst, q = om.strain_stress([model], [mat], u, dT)
    post = api.mechanical_solution_2d(g, [model], [mat], torch.from_numpy(u), delta_temperature=dT)
    post.calc_strain_and_stress()
    gs = torch.stack(post.strain()).cpu().numpy()
    gt = torch.stack(post.stress()).cpu().numpy()
    gq = torch.stack(post.stress_in_qnodes()).cpu().numpy()
    assert _rel(gs, ns) <= 1e-12, "node strains (before the thermal strain is taken out)"
    assert _rel(gt, nst) <= 1e-12 and _rel(gq, q) <= 1e-12


@pytest.mark.parametrize("name", list(MATERIALS))
@pytest.mark.parametrize("nn", [4, 8])
def test_free_thermal_expansion_on_the_gpu(gpu_ctx, name, nn):
    """The stress-free thermal expansion patch through the whole GPU driver: every side held at the
    free-expansion field, a uniform temperature change; the solution is that field and its stresses
    vanish."""
    mesh = meshgen.structured_q4(9, 0.2, seed=43) if nn == 4 else meshgen.structured_q8(6, 0.2, seed=44)
    mat = _with_expansion(MATERIALS[name], EXPANSIONS["anisotropic"])
    field = free_expansion(EXPANSIONS["anisotropic"])
    bcs, values = _all_dirichlet(field)
    dT = np.full(mesh.n_nodes, 1.7)
    g = api.mesh_2d(mesh, gpu_ctx)
    sol = api.equilibrium_equation(g, [A.model(1.0, 0.0)], [mat], bcs, tol=1e-15, values=values, delta_temperature=dT)
    x = sol.x.cpu().numpy()
    ux, uy = field(mesh.xy[:, 0], mesh.xy[:, 1])
    assert np.abs(x[0::2] - ux).max() < 1e-11 and np.abs(x[1::2] - uy).max() < 1e-11
    sol.solution.calc_strain_and_stress()
    q = torch.stack(sol.solution.stress_in_qnodes()).cpu().numpy()
    assert np.abs(q).max() < 1e-9
