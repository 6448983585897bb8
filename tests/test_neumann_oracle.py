"""Pins of the oracle's Neumann problem (the Lagrange row and column of conductivity_matrix_2d,
conductivity_matrix_2d.hpp:39-72, 83-104; init_right_part, stationary_heat_equation_solver_2d.hpp:28-44)
and of coefficients given per quadrature point (evaluate_conductivity.hpp:16-37,
evaluate_mechanical_parameters.hpp:9-30).

The Neumann checks: the border holds int N_i, whose sum is the area; the constraint fixes the
integral of the temperature; and the reference's own known-answer test flux_stability.cpp:35-61 on
its plate mesh, at its 2e-14 bars.  The coefficient checks compare two independent paths: values
per quadrature point against one group per element with that element's constant.
"""
import os

import numpy as np
import pytest

from nonlocal_amd import _abi as A
from nonlocal_amd import meshgen, su2
from oracle import oracle as O
from tests import helpers as H


def _plate():
    """flux_stability.cpp:27-33: plate 10 x 1 (h = 0.125), lambda = 1, flux -4|y| on "Left" and
    +4|y| on "Right", every other side insulated: a Neumann problem."""
    mesh = su2.read(os.path.join(H.MESH_DIR, "plate_10x1_h0.125.su2"))
    names = mesh.bgroup_names
    bcs = [A.boundary() for _ in names]
    bcs[names.index("Left")] = A.boundary(A.NLF_BC_FLUX, 0.0)
    bcs[names.index("Right")] = A.boundary(A.NLF_BC_FLUX, 0.0)
    gl, gr = names.index("Left"), names.index("Right")

    def values(x, y, g, d):
        return -4 * abs(y) if g == gl else (4 * abs(y) if g == gr else 0.0)
    return mesh, [A.model(1.0, 0.0)], [A.thermal_material()], bcs, values


def test_border_holds_the_basis_integrals():
    mesh = meshgen.structured_q4(9, 0.2, seed=5)
    om = O.OracleMesh(mesh)
    om.find_neighbours([[0.0, 0.0]])
    inner = np.ones(mesh.n_nodes, np.uint8)
    K = om.assemble(0, [A.model(1.0, 0.0)], [A.thermal_material()], inner, neumann=True).export(0)
    N = mesh.n_nodes
    assert K.rows == N + 1
    border = np.array([K.val[K.outer[r + 1] - 1] for r in range(N)])
    assert all(K.inner[K.outer[r + 1] - 1] == N for r in range(N))
    assert abs(border.sum() - 1.0) < 1e-14                       # the unit square's area
    assert np.all(border > 0)
    assert K.outer[N + 1] - K.outer[N] == 1 and K.inner[-1] == N and K.val[-1] == 0.0
    # the rest of the matrix is the plain one
    plain = om.assemble(0, [A.model(1.0, 0.0)], [A.thermal_material()], inner).export(0)
    for r in range(0, N, 7):
        a = K.val[K.outer[r]:K.outer[r + 1] - 1]
        assert np.array_equal(a, plain.val[plain.outer[r]:plain.outer[r + 1]])


def test_flux_stability_known_answers():
    """The reference's flux_stability test: on the centre line x = 0 the temperature is 0 and the
    flux -1 (2e-14), the temperature integrates to 0 (3.3e-14) and the x flux to -10 (2.2e-14).

    These bars sit at the rounding floor of a field spanning +-5.12.  The integral bars hold for the
    exact solution of the oracle's bordered system (a direct solve, integrals summed exactly).  The
    pointwise bars are 3e-14: the direct solve itself leaves 2.5e-14 on the centre line.  The CG
    solution, stopped at a relative residual of 1e-15 as the reference's solver is, carries that
    residual into the x-flux integral (about 1e-13), so its integrals are held to 1.5e-13."""
    mesh, models, mats, bcs, values = _plate()
    res = O.thermal_problem(mesh, models, mats, bcs, precond=A.NLF_PRECOND_ILLT, tol=1e-15, values=values)
    assert res["neumann"]
    om = res["mesh"]
    centre = np.abs(mesh.xy[:, 0]) < np.finfo(float).eps
    assert centre.sum() == 9
    direct = H.direct_solution(res["K"], res["f"])[:mesh.n_nodes]
    for T, integral_bars in ((direct, (3.3e-14, 2.2e-14)), (res["T"], (1.5e-13, 1.5e-13))):
        flux, _ = om.heat_flux(models, mats, T)
        assert np.abs(T[centre]).max() <= 3e-14
        assert np.abs(flux[0][centre] + 1.0).max() <= 3e-14
        assert abs(H.integrate_nodes(om, mesh, T)) <= integral_bars[0]
        assert abs(H.integrate_nodes(om, mesh, flux[0]) + 10.0) <= integral_bars[1]


@pytest.mark.parametrize("energy", [0.0, 0.37])
def test_neumann_constraint_fixes_the_integral(energy):
    """The Lagrange row: sum_i (int N_i) T_i = energy, i.e. the temperature integrates to the energy;
    the solution solves the bordered system (direct solve of the same matrix)."""
    mesh = meshgen.structured_q8(6, 0.2, seed=7)
    bcs = [A.boundary(A.NLF_BC_FLUX, -0.5), A.boundary(A.NLF_BC_FLUX, 0.5), A.boundary(), A.boundary()]
    res = O.thermal_problem(mesh, [A.model(1.0, 0.0)], [A.thermal_material()], bcs, tol=1e-15, energy=energy)
    x_ref = H.direct_solution(res["K"], res["f"])
    assert H.rel_l2(res["x"], x_ref) < 1e-10
    assert abs(H.integrate_nodes(res["mesh"], mesh, x_ref[:mesh.n_nodes]) - energy) < 1e-13
    # lambda dT/dn = flux on both sides gives the slope 0.5, and the mean is fixed: T = (x - 1/2)/2 + energy
    assert np.abs(x_ref[:mesh.n_nodes] - (0.5 * (mesh.xy[:, 0] - 0.5) + energy)).max() < 1e-12


def test_unbalanced_fluxes_are_unsolvable():
    mesh = meshgen.structured_q4(6)
    bcs = [A.boundary(A.NLF_BC_FLUX, -0.5), A.boundary(A.NLF_BC_FLUX, 0.6), A.boundary(), A.boundary()]
    with pytest.raises(ValueError, match="unsolvable Neumann"):
        O.thermal_problem(mesh, [A.model(1.0, 0.0)], [A.thermal_material()], bcs)


def test_nonlocal_neumann_solves_the_bordered_system():
    mesh = meshgen.structured_q4(12, 0.15, seed=8)
    bcs = [A.boundary(A.NLF_BC_FLUX, -1.0), A.boundary(A.NLF_BC_FLUX, 1.0), A.boundary(), A.boundary()]
    res = O.thermal_problem(mesh, [A.model(0.4, 0.2, p=2, q=2)], [A.thermal_material()], bcs, tol=1e-15)
    x_ref = H.direct_solution(res["K"], res["f"])
    assert H.rel_l2(res["x"], x_ref) < 1e-9
    assert abs(H.integrate_nodes(res["mesh"], mesh, x_ref[:mesh.n_nodes])) < 1e-13


# ---- coefficients per quadrature point ----------------------------------------------------------
def _one_group_per_element(mesh):
    import copy
    out = copy.deepcopy(mesh)
    out.group_begin = np.arange(mesh.n_elems + 1, dtype=np.int64)
    out.group_names = [f"e{e}" for e in range(mesh.n_elems)]
    return out


THERMAL_KINDS = {A.NLF_CONDUCTIVITY_ISOTROPIC: 1, A.NLF_CONDUCTIVITY_ORTHOTROPIC: 2, A.NLF_CONDUCTIVITY_ANISOTROPIC: 3}


@pytest.mark.parametrize("kind", list(THERMAL_KINDS))
@pytest.mark.parametrize("nn", [4, 8])
def test_conductivity_per_point_equals_one_group_per_element(kind, nn):
    mesh = meshgen.structured_q4(6, 0.2, seed=3) if nn == 4 else meshgen.structured_q8(4, 0.2, seed=3)
    nq = 4 if nn == 4 else 9
    nc = THERMAL_KINDS[kind]
    rng = np.random.default_rng(11)
    per_el = rng.uniform(0.5, 2.0, size=(mesh.n_elems, nc))
    if nc == 3:
        per_el[:, 2] *= 0.2                                    # keep the tensor positive definite
    coeff = np.repeat(per_el[:, None, :], nq, axis=1)           # (E, nq, nc)
    inner = np.ones(mesh.n_nodes, np.uint8)
    inner[mesh.boundary_nodes("left")] = 0
    base = A.thermal_material(kind, (1.0, 1.0, 0.0))
    om = O.OracleMesh(mesh)
    om.find_neighbours([[0.0, 0.0]])
    Kq = om.assemble(0, [A.model(1.0, 0.0)], [base], inner, coefficients=[coeff]).export(0)
    split = _one_group_per_element(mesh)
    os_ = O.OracleMesh(split)
    os_.find_neighbours([[0.0, 0.0]] * split.n_groups)
    mats = [A.thermal_material(kind, tuple(list(c) + [0.0] * (3 - nc))) for c in per_el]
    Ke = os_.assemble(0, [A.model(1.0, 0.0)] * split.n_groups, mats, inner).export(0)
    assert np.array_equal(Kq.outer, Ke.outer) and np.array_equal(Kq.inner, Ke.inner)
    assert np.array_equal(Kq.val, Ke.val)


def test_conductivity_per_point_constant_equals_the_constant():
    mesh = meshgen.structured_q4(5, 0.2, seed=4)
    inner = np.ones(mesh.n_nodes, np.uint8)
    om = O.OracleMesh(mesh)
    om.find_neighbours([[0.0, 0.0]])
    mat = A.thermal_material(A.NLF_CONDUCTIVITY_ANISOTROPIC, (1.3, 0.7, 0.2))
    coeff = np.tile([1.3, 0.7, 0.2], (mesh.n_elems * 4, 1))
    a = om.assemble(0, [A.model(1.0, 0.0)], [mat], inner, coefficients=[coeff]).export(0)
    b = om.assemble(0, [A.model(1.0, 0.0)], [mat], inner).export(0)
    assert np.array_equal(a.val, b.val)


@pytest.mark.parametrize("name", ["isotropic", "orthotropic", "anisotropic_0.4"])
def test_hooke_per_point_equals_one_group_per_element(name):
    from tests.test_elastic_oracle import MATERIALS
    mesh = meshgen.structured_q4(5, 0.2, seed=6)
    nq = 4
    rng = np.random.default_rng(12)
    mats_e = []
    for e in range(mesh.n_elems):
        m = A.ElasticMaterial.from_buffer_copy(MATERIALS[name])
        m.young *= rng.uniform(0.5, 2.0)
        m.young_y *= rng.uniform(0.5, 2.0)
        mats_e.append(m)
    per_el = np.array([O.hooke(m)[1] for m in mats_e])            # (E, ncomp) in the kind's order
    coeff = np.repeat(per_el[:, None, :], nq, axis=1)
    inner = np.ones(2 * mesh.n_nodes, np.uint8)
    inner[np.ravel([[2 * v, 2 * v + 1] for v in mesh.boundary_nodes("left")])] = 0
    om = O.OracleMesh(mesh)
    om.find_neighbours([[0.0, 0.0]])
    Kq = om.assemble(1, [A.model(1.0, 0.0)], [MATERIALS[name]], inner, coefficients=[coeff]).export(0)
    split = _one_group_per_element(mesh)
    os_ = O.OracleMesh(split)
    os_.find_neighbours([[0.0, 0.0]] * split.n_groups)
    Ke = os_.assemble(1, [A.model(1.0, 0.0)] * split.n_groups, mats_e, inner).export(0)
    assert np.array_equal(Kq.inner, Ke.inner) and np.array_equal(Kq.val, Ke.val)


def test_coefficients_per_point_of_a_nonlocal_problem_are_refused():
    """problem_settings.hpp:32-34: nonlocal with varying coefficients is asymmetric (BiCGSTAB)."""
    mesh = meshgen.structured_q4(5)
    om = O.OracleMesh(mesh)
    om.find_neighbours([[0.3, 0.3]])
    coeff = np.ones((mesh.n_elems, 4, 1))
    with pytest.raises(RuntimeError, match="asymmetric"):
        om.assemble(0, [A.model(0.5, 0.3)], [A.thermal_material()], np.ones(mesh.n_nodes, np.uint8),
                    coefficients=[coeff])
