"""CPU pins of the oracle's strain/stress post-processing and position-dependent boundary values.

calc_strain_and_stress (mechanical_solution_2d.hpp:192-232) and boundary conditions given as
functions of the point (boundary_condition_second_kind_2d.hpp:12-33, boundary_condition_first_kind_2d
.hpp:17-26) are pinned by checks that do not restate them:
  * a linear displacement has the exact strain sym(grad u) and stress C eps everywhere;
  * the weak divergence of the quadrature-point stresses equals the assembled operator times u,
    sum_e sum_q w_q B_i(q)^T sigma(x_q) = (K u)_i, for local and nonlocal groups;
  * the total load of a linear traction is its exact integral along the boundary edges;
  * the reference's own mechanical tests on its SU2 meshes (isotropic_lame_solid_ring.cpp,
    isotropic_lame_composite_ring.cpp, saint_venant.cpp), with its tolerances.
"""
import numpy as np
import pytest

from nonlocal_amd import _abi as A
from nonlocal_amd import configs, meshgen
from oracle import oracle as O
from tests import helpers as H

STEEL = A.elastic_material(2.1, 0.3)


def _hooke(mat):
    v = mat.young / (1 - mat.poisson ** 2)
    return np.array([[v, mat.poisson * v, 0.0], [mat.poisson * v, v, 0.0], [0.0, 0.0, 0.5 * v * (1 - mat.poisson)]])


@pytest.mark.parametrize("nn", [4, 8])
def test_linear_displacement_exact_strain_and_stress(nn):
    mesh = meshgen.structured_q4(8, jitter=0.2, seed=1) if nn == 4 else meshgen.structured_q8(6, jitter=0.2, seed=1)
    G = np.array([[0.3, -0.7], [1.1, 0.4]])                  # grad u: rows u_x, u_y; columns d/dx, d/dy
    u = mesh.xy @ G.T + np.array([0.2, -0.5])
    om = O.OracleMesh(mesh)
    ns, nst, q = om.strain_stress([A.model(1.0, 0.0)], [STEEL], u.ravel())
    eps = np.array([G[0, 0], G[1, 1], 0.5 * (G[0, 1] + G[1, 0])])
    sig = _hooke(STEEL) @ np.array([eps[0], eps[1], 2 * eps[2]])   # engineering shear in the Voigt product
    assert np.abs(ns - eps[:, None]).max() < 1e-13
    assert np.abs(q - sig[:, None]).max() < 1e-12
    assert np.abs(nst - sig[:, None]).max() < 1e-12


def _weak_divergence(om, mesh, qs):
    """sum_e sum_q w_q B_i(q)^T sigma(x_q), interleaved (x, y) per node."""
    xq, jac, D = om.quadrature()
    w = O.tables(mesh.nn)[0]
    E = D.shape[0]
    s = qs.reshape(3, E, -1)
    fx = (D[..., 0] * (w * s[0])[:, None, :] + D[..., 1] * (w * s[2])[:, None, :]).sum(axis=2)
    fy = (D[..., 0] * (w * s[2])[:, None, :] + D[..., 1] * (w * s[1])[:, None, :]).sum(axis=2)
    out = np.zeros(2 * mesh.n_nodes)
    np.add.at(out, 2 * mesh.elems.ravel(), fx.ravel())
    np.add.at(out, 2 * mesh.elems.ravel() + 1, fy.ravel())
    return out


CASES = {
    "q4_bell": (lambda: meshgen.structured_q4(12), [configs.bell(0.4, 0.25)], [STEEL]),
    "q8_jitter_fast": (lambda: meshgen.structured_q8(6, jitter=0.2, seed=2), [A.model(0.5, 0.35, q=1.0)], [STEEL]),
    "q4_composite_groups": (lambda: meshgen.composite_q4(16, count=2, rmin=0.15, rmax=0.25, seed=5)[0], None, None),
}


@pytest.mark.parametrize("case", list(CASES))
def test_stress_weak_divergence(case):
    build, models, materials = CASES[case]
    mesh = build()
    if models is None:
        models = [A.model(1.0, 0.0)] + [configs.bell(0.5, 0.2)] * (mesh.n_groups - 1)
        materials = [STEEL] + [A.elastic_material(0.7, 0.2)] * (mesh.n_groups - 1)
    om = O.OracleMesh(mesh)
    om.find_neighbours(H.radii(models))
    inner = np.ones(2 * mesh.n_nodes, dtype=np.uint8)
    K = om.assemble(1, models, materials, inner)
    u = np.random.default_rng(3).standard_normal(2 * mesh.n_nodes)
    ns, nst, q = om.strain_stress(models, materials, u)
    lhs = _weak_divergence(om, mesh, q)
    rhs = H.upper_to_full(K.export(0), 2 * mesh.n_nodes) @ u
    assert np.linalg.norm(lhs - rhs) <= 1e-12 * np.linalg.norm(rhs)
    local = [A.model(1.0, 0.0)] * len(models)
    _, _, q_local = om.strain_stress(local, materials, u)
    assert np.linalg.norm(_weak_divergence(om, mesh, q_local) - rhs) > 1e-3 * np.linalg.norm(rhs)


@pytest.mark.parametrize("nn", [4, 8])
def test_linear_traction_total_load(nn):
    """sum_i f_i over a flux group = the integral of g along its edges (exact for linear g)."""
    mesh = meshgen.structured_q4(7, jitter=0.2, seed=3) if nn == 4 else meshgen.structured_q8(5, jitter=0.2, seed=3)
    om = O.OracleMesh(mesh)
    bcs = [A.boundary(A.NLF_BC_FLUX, 0.0, A.NLF_BC_FLUX, 0.0)] * 4
    f, xD, inner = om.boundary(bcs, dof=2, values=lambda x, y, g, d: (1.0 + 2.0 * x - 3.0 * y) * (d + 1))
    # the boundary of the unit square (jitter leaves boundary nodes in place): int (1 + 2x - 3y) ds = 4 + 4 - 6
    assert abs(f[0::2].sum() - 2.0) < 1e-13 and abs(f[1::2].sum() - 4.0) < 1e-13
    # first kind: the value at the node
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0, A.NLF_BC_NONE, 0.0)] * 4
    f, xD, inner = om.boundary(bcs, dof=2, values=lambda x, y, g, d: 5.0 + x + 10 * y)
    nodes = np.unique(mesh.bedges.ravel())
    assert np.allclose(xD[2 * nodes], 5.0 + mesh.xy[nodes, 0] + 10 * mesh.xy[nodes, 1], rtol=0, atol=1e-15)
    assert inner[2 * nodes].max() == 0 and inner[2 * nodes + 1].min() == 1


@pytest.mark.parametrize("composite", [False, True], ids=["solid_ring", "composite_ring"])
def test_lame_ring_known_answers(composite):
    mesh, models, materials, bcs, values, expected = H.lame_ring_problem(composite)
    r = O.elastic_problem(mesh, models, materials, bcs, tol=1e-14, values=values)
    u = r["u"]
    ns, nst, _ = r["mesh"].strain_stress(models, materials, u)
    got = {"ux": u[0::2], "uy": u[1::2], "exx": ns[0], "eyy": ns[1], "exy": ns[2],
           "sxx": nst[0], "syy": nst[1], "sxy": nst[2]}
    x, y = mesh.xy[:, 0], mesh.xy[:, 1]
    for k, tol in H.LAME_TOLERANCES[composite].items():
        assert H.norm_error(got[k], expected[k](x, y)) <= tol, k


def test_saint_venant_known_answers():
    mesh, models, materials, bcs, values = H.saint_venant_problem()
    r = O.elastic_problem(mesh, models, materials, bcs, tol=1e-15, values=values)
    u = r["u"]
    _, nst, _ = r["mesh"].strain_stress(models, materials, u)
    centre = np.abs(mesh.xy[:, 0]) < np.finfo(float).eps
    assert centre.sum() == 9
    assert np.abs(u[0::2][centre]).max() <= 2e-14                    # saint_venant.cpp:51-56
    assert np.abs(nst[0][centre] - 1.0).max() <= 1.3e-9               # :58-63
    assert abs(H.integrate_nodes(r["mesh"], mesh, u[0::2])) <= 7.5e-13          # :65-70
    assert abs(H.integrate_nodes(r["mesh"], mesh, nst[0]) - 10.0) <= 1e-9       # :72-77
