"""CPU pins of the oracle's heat capacity matrix, source term and implicit Euler time stepping.

heat_capacity_matrix_2d (heat_capacity_matrix_2d.hpp:31-62), integrate_right_part (right_part_2d.hpp:
10-35) and nonstationary_heat_equation_solver_2d (nonstationary_heat_equation_solver_2d.hpp:66-106)
are pinned by mathematics that does not restate them:
  * the textbook mass matrix of a square bilinear element, and sum_ij C_ij = rho c |Omega|;
  * the source integral of a constant and of a linear source is exact on mapped elements;
  * a generalised eigenvector K v = lambda C v decays by exactly 1 / (1 + tau lambda) per step;
  * the stationary solution is a fixed point of the step;
  * the reference's own source-term test (boundary_conditions_2d.cpp:84-113).
"""
import math

import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp

from nonlocal_amd import _abi as A
from nonlocal_amd import configs, meshgen, su2
from oracle import oracle as O
from tests import helpers as H


def _full(csr, n):
    U = sp.csr_matrix((csr.val, csr.inner, csr.outer), shape=(n, n))
    return (U + sp.triu(U, 1).T).toarray()


def test_capacity_matrix_of_one_square_element():
    h = 0.3
    mesh = meshgen.Mesh2D(xy=np.array([[0, 0], [h, 0], [h, h], [0, h]], dtype=float),
                          elems=np.array([[0, 1, 2, 3]], dtype=np.int32), nn=4,
                          group_begin=np.array([0, 1]), group_names=["g"],
                          bedges=np.zeros((0, 2), dtype=np.int32), bgroup_begin=np.array([0]), bgroup_names=[])
    om = O.OracleMesh(mesh)
    Cm = O.capacity_matrix(om, [2.5], np.ones(4, dtype=np.uint8))
    ref = 2.5 * h * h / 36 * np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]])
    np.testing.assert_allclose(_full(Cm.export(0), 4), ref, rtol=1e-14, atol=0)


@pytest.mark.parametrize("nn", [4, 8])
def test_capacity_matrix_total(nn):
    mesh = meshgen.structured_q4(7, jitter=0.2, seed=4) if nn == 4 else meshgen.structured_q8(5, jitter=0.2, seed=4)
    om = O.OracleMesh(mesh)
    Cm = O.capacity_matrix(om, [1.7], np.ones(mesh.n_nodes, dtype=np.uint8))
    M = _full(Cm.export(0), mesh.n_nodes)
    assert abs(M.sum() - 1.7) < 1e-13                    # unit square
    assert np.all(np.linalg.eigvalsh(M) > 0)


@pytest.mark.parametrize("nn", [4, 8])
def test_source_integrals(nn):
    mesh = meshgen.structured_q4(7, jitter=0.2, seed=5) if nn == 4 else meshgen.structured_q8(5, jitter=0.2, seed=5)
    om = O.OracleMesh(mesh)
    assert abs(om.right_part(1, lambda x, y: 3.0).sum() - 3.0) < 1e-13
    assert abs(om.right_part(1, lambda x, y: 1 + 2 * x - 4 * y).sum() - 0.0) < 1e-13      # int over the unit square
    f = om.right_part(2, lambda x, y: (x, 2.0))
    assert abs(f[0::2].sum() - 0.5) < 1e-13 and abs(f[1::2].sum() - 2.0) < 1e-13


def _strip():
    """8 x 8 Q4, T = 0 on x = 0 and x = 1, insulated elsewhere, lambda = 1.3, rho c = 2."""
    mesh = meshgen.structured_q4(8, jitter=0.15, seed=6)
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0), A.boundary(A.NLF_BC_FIRST_KIND, 0.0), A.boundary(), A.boundary()]
    return mesh, [A.thermal_material(A.NLF_CONDUCTIVITY_ISOTROPIC, (1.3, 0, 0))], bcs


@pytest.mark.parametrize("model", [A.model(1.0, 0.0), configs.bell(0.5, 0.3)], ids=["local", "nonlocal"])
def test_eigenmode_decay(model):
    mesh, materials, bcs = _strip()
    om = O.OracleMesh(mesh)
    om.find_neighbours(H.radii([model]))
    f, xD, inner = om.boundary(bcs)
    K = _full(om.assemble(0, [model], materials, inner).export(0), mesh.n_nodes)
    Cm = _full(O.capacity_matrix(om, [2.0], inner).export(0), mesh.n_nodes)
    idx = np.flatnonzero(inner)
    lam, vec = la.eigh(K[np.ix_(idx, idx)], Cm[np.ix_(idx, idx)])
    for k in (0, 5):
        v = np.zeros(mesh.n_nodes)
        v[idx] = vec[:, k]
        tau, steps = 0.01, 6
        out = O.transient(mesh, [model], materials, [2.0], bcs, tau, v, steps)
        np.testing.assert_allclose(out[-1], v / (1 + tau * lam[k]) ** steps, rtol=0, atol=1e-12 * np.abs(v).max())


def test_stationary_solution_is_a_fixed_point():
    p = configs.cfg0(scale=0.3)                           # 12 x 12, T = 0 on x = 0, unit flux on x = 1
    p.models = [configs.bell(0.5, 0.25)]
    om, K, f, xD, inner = H.oracle_assembly(p)
    T_star = H.direct_solution(K, f)
    out = O.transient(p.mesh, p.models, p.materials, [3.0], p.bcs, 0.05, T_star, 3)
    assert np.abs(out[-1] - T_star).max() <= 1e-12 * np.abs(T_star).max()
    # and from zero it approaches it
    out = O.transient(p.mesh, p.models, p.materials, [3.0], p.bcs, 5.0, np.zeros_like(T_star), 40)
    assert np.abs(out[-1] - T_star).max() <= 1e-6 * np.abs(T_star).max()


def source_test_problem():
    """boundary_conditions_2d.cpp:28-113 ("all_boundaries_temperature"): the 16-element unit square,
    anisotropic lambda = (10, 10, 0), T = f(x) g(y) with f(x) = (-3 a (x - 1.01))^(-1/3),
    a = 0.7 sigma / 10, g(y) = exp(-(y - 0.5)^2 / 0.1), its values on all four sides and the source that
    makes it exact."""
    sigma, lam, d, m = 5.67036713e-8, 10.0, 0.1, 0.5
    alpha = 0.7 * sigma / lam

    def f(x):
        return np.power(-3.0 * alpha * (x - 1.01), -1.0 / 3.0)

    def g(y):
        return np.exp(-(y - m) ** 2 / d)

    def ref(x, y):
        return f(x) * g(y)

    def source(x, y):
        a1 = 4 * alpha * alpha * lam
        a2 = 2 * lam / d
        return -ref(x, y) * (a1 * f(x) ** 6 + a2 * (2.0 / d * (y - m) ** 2 - 1.0))

    mesh = su2.read(f"{H.MESH_DIR}/square_1x1_16el_uniform_mesh.su2")
    names = mesh.bgroup_names
    sides = {"Left": lambda x, y: ref(0.0, y), "Right": lambda x, y: ref(1.0, y),
             "Top": lambda x, y: ref(x, 1.0), "Bottom": lambda x, y: ref(x, 0.0)}
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0) for _ in names]

    def values(x, y, group, dof):
        return float(sides[names[group]](x, y))

    materials = [A.thermal_material(A.NLF_CONDUCTIVITY_ANISOTROPIC, (lam, lam, 0.0))]
    return mesh, [A.model(1.0, 0.0)], materials, bcs, values, source, ref


def test_source_term_known_answer():
    mesh, models, materials, bcs, values, source, ref = source_test_problem()
    r = O.thermal_problem(mesh, models, materials, bcs, tol=1e-14, values=values, source=source)
    T = r["T"]
    exact = ref(mesh.xy[:, 0], mesh.xy[:, 1])
    err = np.sum((T - exact) ** 2) / np.sum(exact ** 2)
    assert err < 1e-3                                     # check_solution(..., 1e-3), boundary_conditions_2d.cpp:112
    # the source matters: without it the error is far larger
    r0 = O.thermal_problem(mesh, models, materials, bcs, tol=1e-14, values=values)
    assert np.sum((r0["T"] - exact) ** 2) / np.sum(exact ** 2) > 10 * err
