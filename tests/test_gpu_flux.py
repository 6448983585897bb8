"""GPU heat flux (nlf_heat_flux) against the oracle's calc_flux and the reference's known answers.

* Parity: quadrature-point and node fluxes within 1e-12 (relative to the largest flux) of the
  oracle on the same temperature, for Q4 and Q8, every conductivity kind, single- and multi-group
  meshes with local and nonlocal groups.  The sums run in another order than the oracle's loops, so
  the comparison is to rounding, not bit for bit.
* Size-independent property at a larger size: the weak divergence of the quadrature-point flux is
  minus the assembled operator times T (sum_e sum_q w_q D_i(q) . q(x_q) = -(K T)_i).
* The reference's solver tests on its SU2 meshes (isotropic_solid_ring.cpp,
  isotropic_composite_ring.cpp): temperature and flux within its tolerances, through the GPU solve.
"""
import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, configs, meshgen
from nonlocal_amd._lib import load
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

ANISO = A.thermal_material(A.NLF_CONDUCTIVITY_ANISOTROPIC, (1.0, 0.25, 0.1))
ORTHO = A.thermal_material(A.NLF_CONDUCTIVITY_ORTHOTROPIC, (1.3, 0.4, 0.0))
ISO = A.thermal_material(A.NLF_CONDUCTIVITY_ISOTROPIC, (1.7, 0.0, 0.0))


def _cases():
    comp, _ = meshgen.composite_q4(40, count=4, rmin=0.08, rmax=0.15, seed=9)
    return {
        "cfg0_q4_bell": (configs.cfg0().mesh, [configs.bell(0.5, 0.1)], [A.thermal_material()]),
        "q8_jitter_aniso_bell": (meshgen.structured_q8(16, jitter=0.2, seed=1), [configs.bell(0.3, 0.15)], [ANISO]),
        "q4_jitter_ortho_constant": (meshgen.structured_q4(24, jitter=0.2, seed=2),
                                     [A.model(0.4, 0.12, influence=A.NLF_INFLUENCE_CONSTANT)], [ORTHO]),
        "q8_exponential": (meshgen.structured_q8(12), [A.model(0.6, 0.2, influence=A.NLF_INFLUENCE_EXPONENTIAL,
                                                                p=2.0, q=1.0)], [ISO]),
        "q4_fast_polynomial": (meshgen.structured_q4(24, jitter=0.2, seed=6), [A.model(0.5, 0.12, q=1.0)], [ANISO]),
        "q4_local": (meshgen.structured_q4(20, jitter=0.2, seed=4), [A.model(1.0, 0.0)], [ANISO]),
        "composite_local_and_nonlocal": (comp, [A.model(1.0, 0.0)] + [configs.bell(0.5, 0.1)] * (comp.n_groups - 1),
                                         [ISO] + [A.thermal_material(A.NLF_CONDUCTIVITY_ISOTROPIC, (10.0, 0, 0))] *
                                         (comp.n_groups - 1)),
    }


CASES = _cases()


@pytest.mark.parametrize("case", list(CASES))
def test_flux_matches_oracle(gpu_ctx, case):
    print(f"[flux parity] {case}", flush=True)
    mesh, models, materials = CASES[case]
    T = np.random.default_rng(21).standard_normal(mesh.n_nodes)
    om = O.OracleMesh(mesh)
    om.find_neighbours(H.radii(models))
    n_ref, q_ref = om.heat_flux(models, materials, T)
    g = api.mesh_2d(mesh, gpu_ctx)
    sol = api.heat_equation_solution_2d(g, models, materials, torch.from_numpy(T))
    with pytest.raises(RuntimeError):
        sol.flux()
    fx, fy = sol.calc_flux()
    qx, qy = sol.flux_in_qnodes()
    scale = np.abs(q_ref).max()
    eq = max(np.abs(qx.cpu().numpy() - q_ref[0]).max(), np.abs(qy.cpu().numpy() - q_ref[1]).max()) / scale
    en = max(np.abs(fx.cpu().numpy() - n_ref[0]).max(), np.abs(fy.cpu().numpy() - n_ref[1]).max()) / scale
    print(f"  quadrature points {eq:.2e}, nodes {en:.2e} of the largest flux", flush=True)
    assert eq <= 1e-12 and en <= 1e-12
    assert sol.calc_flux()[0] is fx                    # computed once, like the reference's calc_flux


def test_flux_needs_neighbours(gpu_ctx):
    mesh = meshgen.structured_q4(6)
    g = api.mesh_2d(mesh, gpu_ctx)
    T = torch.zeros(mesh.n_nodes, dtype=torch.float64, device="cuda")
    out = torch.empty(2 * mesh.n_nodes, dtype=torch.float64, device="cuda")
    marr = A.array_of(A.Model, [configs.bell(0.5, 0.3)])
    mats = A.array_of(A.ThermalMaterial, [ISO])
    assert load().nlf_heat_flux(g.h, marr, mats, api._dev(T), api._dev(out), None) == A.NLF_ERR_STATE


def test_flux_weak_divergence_composite(gpu_ctx):
    """200 x 200 composite (matrix local, 8 nonlocal inclusions of lambda = 10) with the whole
    operator (no boundary conditions): B^T W q = -K T to rounding."""
    print("[flux weak divergence] composite 200x200", flush=True)
    mesh, _ = meshgen.composite_q4(200, count=8, rmin=0.05, rmax=0.12, seed=3)
    models = [A.model(1.0, 0.0)] + [configs.bell(0.5, 0.02)] * (mesh.n_groups - 1)
    materials = [ISO] + [A.thermal_material(A.NLF_CONDUCTIVITY_ISOTROPIC, (10.0, 0, 0))] * (mesh.n_groups - 1)
    g = api.mesh_2d(mesh, gpu_ctx)
    api.find_neighbours(g, H.radii(models))
    K = api._assemble(g, 0, models, materials, np.ones(mesh.n_nodes, dtype=np.uint8))
    T = torch.from_numpy(np.random.default_rng(5).standard_normal(mesh.n_nodes)).cuda()
    sol = api.heat_equation_solution_2d(g, models, materials, T)
    sol.calc_flux()
    qx, qy = (v.cpu().numpy() for v in sol.flux_in_qnodes())
    xq, jac, D, _ = g.quadrature()
    w = O.tables(4)[0]
    E = mesh.n_elems
    q = np.stack([qx.reshape(E, 4), qy.reshape(E, 4)])
    contrib = (D[..., 0] * (w * q[0])[:, None, :] + D[..., 1] * (w * q[1])[:, None, :]).sum(axis=2)
    lhs = np.zeros(mesh.n_nodes)
    np.add.at(lhs, mesh.elems.ravel(), contrib.ravel())
    rhs = -K.spmv(T).cpu().numpy()
    err = np.linalg.norm(lhs - rhs) / np.linalg.norm(rhs)
    print(f"  ||B'Wq + K T|| / ||K T|| = {err:.2e}", flush=True)
    assert err <= 1e-11


@pytest.mark.parametrize("composite", [False, True], ids=["solid_ring", "composite_ring"])
def test_ring_known_answers(gpu_ctx, composite):
    print(f"[ring] composite={composite}", flush=True)
    mesh, models, materials, bcs, temperature, flux = H.ring_problem(composite)
    g = api.mesh_2d(mesh, gpu_ctx)
    sol = api.stationary_heat_equation_solver_2d(g, models, materials, bcs, tol=1e-14)
    fx, fy = sol.solution.calc_flux()
    x, y = mesh.xy[:, 0], mesh.xy[:, 1]
    ex, ey = flux(x, y)
    tol_T, tol_q = H.RING_TOLERANCES[composite]
    eT = H.norm_error(sol.x.cpu().numpy(), temperature(x, y))
    eqx = H.norm_error(fx.cpu().numpy(), ex)
    eqy = H.norm_error(fy.cpu().numpy(), ey)
    print(f"  temperature {eT:.3e} (<= {tol_T}), flux {eqx:.3e} {eqy:.3e} (<= {tol_q})", flush=True)
    assert eT <= tol_T and eqx <= tol_q and eqy <= tol_q
