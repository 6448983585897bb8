"""Pins of the oracle's orthotropic and anisotropic plane elasticity (elastic_parameters.hpp:62-133,
stiffness_matrix_2d.hpp:44-123, mechanical_parameters_2d.hpp:27-48) and of the thermal load and
thermal strains (temperature_condition_2d.hpp:20-170, mechanical_solution_2d.hpp:135-166).

The checks are closed forms (the inverse compliance, the 4th-order tensor rotation), patch tests
that bilinear and serendipity elements pass exactly for any constant Hooke matrix (a linear
displacement, and the stress-free thermal expansion), and dense double integrals on tiny meshes.
"""
import math

import numpy as np
import pytest

from nonlocal_amd import _abi as A
from nonlocal_amd import meshgen
from oracle import oracle as O
from tests import helpers as H

ORTHO = dict(young=(3.0, 1.2), poisson=(0.3, 0.12), shear=0.55)      # E_x nu_1 = E_y nu_0 (consistent)


def voigt_of(material):
    """3x3 Voigt matrix [[11, 12, 16], [12, 22, 26], [16, 26, 66]] computed independently of the
    oracle: the inverse plane-stress compliance, rotated as a 4th-order tensor."""
    if material.kind == A.NLF_ELASTIC_ISOTROPIC:
        E, nu = material.young, material.poisson
        Ex = Ey = E
        nxy = nyx = nu
        G = E / (2.0 * (1.0 + nu))
    else:
        Ex, Ey, nxy, nyx, G = material.young, material.young_y, material.poisson, material.poisson_y, material.shear
    S = np.array([[1 / Ex, -nyx / Ey, 0.0], [-nxy / Ex, 1 / Ey, 0.0], [0.0, 0.0, 1 / G]])
    # the reference writes C12 = E_x nu_1 / (1 - nu_0 nu_1): the inverse of S when E_x nu_1 = E_y nu_0
    C0 = np.linalg.inv(S)
    return rotate_voigt(C0, material.angle) if material.kind == A.NLF_ELASTIC_ANISOTROPIC else C0


def rotate_voigt(C0, angle):
    """C'_ijkl = R_ia R_jb R_kc R_ld C_abcd with R the counter-clockwise rotation by `angle`."""
    C = np.zeros((2, 2, 2, 2))
    C[0, 0, 0, 0], C[1, 1, 1, 1] = C0[0, 0], C0[1, 1]
    C[0, 0, 1, 1] = C[1, 1, 0, 0] = C0[0, 1]
    for idx in [(0, 1, 0, 1), (0, 1, 1, 0), (1, 0, 0, 1), (1, 0, 1, 0)]:
        C[idx] = C0[2, 2]
    for idx in [(0, 0, 0, 1), (0, 0, 1, 0), (0, 1, 0, 0), (1, 0, 0, 0)]:
        C[idx] = C0[0, 2]
    for idx in [(1, 1, 0, 1), (1, 1, 1, 0), (0, 1, 1, 1), (1, 0, 1, 1)]:
        C[idx] = C0[1, 2]
    R = np.array([[math.cos(angle), -math.sin(angle)], [math.sin(angle), math.cos(angle)]])
    Cp = np.einsum("ia,jb,kc,ld,abcd->ijkl", R, R, R, R, C)
    return np.array([[Cp[0, 0, 0, 0], Cp[0, 0, 1, 1], Cp[0, 0, 0, 1]],
                     [Cp[0, 0, 1, 1], Cp[1, 1, 1, 1], Cp[1, 1, 0, 1]],
                     [Cp[0, 0, 0, 1], Cp[1, 1, 0, 1], Cp[0, 1, 0, 1]]])


def voigt_from_oracle(material):
    kind, h = O.hooke(material)
    if kind == A.NLF_ELASTIC_ISOTROPIC:
        h11, h12, h66 = h
        return np.array([[h11, h12, 0], [h12, h11, 0], [0, 0, h66]])
    if kind == A.NLF_ELASTIC_ORTHOTROPIC:
        h11, h12, h66, h22 = h
        return np.array([[h11, h12, 0], [h12, h22, 0], [0, 0, h66]])
    h11, h12, h66, h22, h16, h26 = h
    return np.array([[h11, h12, h16], [h12, h22, h26], [h16, h26, h66]])


MATERIALS = {
    "isotropic": A.elastic_material(2.0, 0.25),
    "orthotropic": A.orthotropic_material(**ORTHO),
    "anisotropic_0.4": A.anisotropic_material(**ORTHO, angle=0.4),
    "anisotropic_2.2": A.anisotropic_material(**ORTHO, angle=2.2),
}


@pytest.mark.parametrize("name", list(MATERIALS))
def test_hooke_matrix_is_the_rotated_inverse_compliance(name):
    C = voigt_from_oracle(MATERIALS[name])
    assert np.allclose(C, voigt_of(MATERIALS[name]), rtol=0, atol=1e-14 * np.abs(C).max())


def test_anisotropic_at_angle_zero_is_orthotropic():
    a = voigt_from_oracle(A.anisotropic_material(**ORTHO, angle=0.0))
    b = voigt_from_oracle(A.orthotropic_material(**ORTHO))
    assert np.allclose(a, b, rtol=0, atol=1e-15 * np.abs(b).max())
    # a quarter turn swaps the axes
    c = voigt_from_oracle(A.anisotropic_material(**ORTHO, angle=math.pi / 2))
    assert np.allclose(c[[1, 0, 2]][:, [1, 0, 2]], b, rtol=0, atol=1e-14 * np.abs(b).max())


def _mesh(nn):
    return meshgen.structured_q4(7, 0.2, seed=31) if nn == 4 else meshgen.structured_q8(5, 0.2, seed=32)


def _all_dirichlet(field):
    """first-kind on all four sides with the values of field(x, y) -> (u_x, u_y)"""
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0, A.NLF_BC_FIRST_KIND, 0.0) for _ in range(4)]
    return bcs, (lambda x, y, g, d: field(x, y)[d])


LINEAR = np.array([[0.3, -0.2], [0.15, 0.4]])          # u = LINEAR @ (x, y) + (0.1, -0.05)


def linear_u(x, y):
    return LINEAR[0, 0] * x + LINEAR[0, 1] * y + 0.1, LINEAR[1, 0] * x + LINEAR[1, 1] * y - 0.05


@pytest.mark.parametrize("name", list(MATERIALS))
@pytest.mark.parametrize("nn", [4, 8])
def test_dirichlet_patch_every_hooke_kind(name, nn):
    mesh = _mesh(nn)
    bcs, values = _all_dirichlet(linear_u)
    res = O.elastic_problem(mesh, [A.model(1.0, 0.0)], [MATERIALS[name]], bcs, tol=1e-15, values=values)
    ux, uy = linear_u(mesh.xy[:, 0], mesh.xy[:, 1])
    x = H.direct_solution(res["K"], res["f"])
    assert np.abs(x[0::2] - ux).max() < 1e-12 and np.abs(x[1::2] - uy).max() < 1e-12


@pytest.mark.parametrize("name", list(MATERIALS))
@pytest.mark.parametrize("nn", [4, 8])
def test_traction_patch_every_hooke_kind(name, nn):
    """Left edge held at the linear field, the other sides loaded with the tractions of its constant
    stress sigma = C eps (C from voigt_of): the linear field is the solution, so the assembled C
    must be that C."""
    mesh = _mesh(nn)
    exx, eyy, exy = LINEAR[0, 0], LINEAR[1, 1], 0.5 * (LINEAR[0, 1] + LINEAR[1, 0])
    sxx, syy, sxy = voigt_of(MATERIALS[name]) @ np.array([exx, eyy, 2 * exy])
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0, A.NLF_BC_FIRST_KIND, 0.0),
           A.boundary(A.NLF_BC_FLUX, sxx, A.NLF_BC_FLUX, sxy),            # x = 1, normal +x
           A.boundary(A.NLF_BC_FLUX, -sxy, A.NLF_BC_FLUX, -syy),          # y = 0, normal -y
           A.boundary(A.NLF_BC_FLUX, sxy, A.NLF_BC_FLUX, syy)]            # y = 1, normal +y

    def values(x, y, g, d):
        if g == 0:
            return linear_u(x, y)[d]
        return bcs[g].value[d]
    res = O.elastic_problem(mesh, [A.model(1.0, 0.0)], [MATERIALS[name]], bcs, tol=1e-15, values=values)
    ux, uy = linear_u(mesh.xy[:, 0], mesh.xy[:, 1])
    x = H.direct_solution(res["K"], res["f"])
    assert np.abs(x[0::2] - ux).max() < 1e-11 and np.abs(x[1::2] - uy).max() < 1e-11
    # and the stresses of that field are sigma
    _, nst, q = res["mesh"].strain_stress([A.model(1.0, 0.0)], [MATERIALS[name]], x)
    assert np.allclose(q[0], sxx, atol=1e-10) and np.allclose(q[1], syy, atol=1e-10) and np.allclose(q[2], sxy, atol=1e-10)


@pytest.mark.parametrize("name", ["orthotropic", "anisotropic_0.4"])
def test_nonlocal_stiffness_dense_brute_force(name):
    """K_NL by an all-pairs double loop over elements and Gauss points with B^T C B (numpy, C from
    voigt_of) equals the oracle's sparse assembly when the search covers the support of phi."""
    mesh = meshgen.structured_q4(3, 0.15, seed=8)
    om = O.OracleMesh(mesh)
    r = 0.6
    model = A.model(0.0, r, search_radius=3.0, p=2, q=2)
    om.find_neighbours([[3.0, 3.0]])
    N = mesh.n_nodes
    K = H.upper_to_full(om.assemble(1, [model], [MATERIALS[name]], np.ones(2 * N, np.uint8)).export(0), 2 * N).toarray()
    xq, jac, D = om.quadrature()
    w = O.tables(4)[0]
    C = voigt_of(MATERIALS[name])
    norm = 3.0 / (math.pi * r * r)

    def B(e, q):
        b = np.zeros((3, 8))
        for i in range(4):
            dx, dy = D[e, i, q]
            b[:, 2 * i] = (dx, 0.0, dy)
            b[:, 2 * i + 1] = (0.0, dy, dx)
        return b
    ref = np.zeros((2 * N, 2 * N))
    for eL in range(mesh.n_elems):
        dofL = np.ravel([[2 * v, 2 * v + 1] for v in mesh.elems[eL]])
        for eN in range(mesh.n_elems):
            dofN = np.ravel([[2 * v, 2 * v + 1] for v in mesh.elems[eN]])
            blk = np.zeros((8, 8))
            for qL in range(4):
                for qN in range(4):
                    d2 = np.sum((xq[eL, qL] - xq[eN, qN]) ** 2) / (r * r)
                    if d2 >= 1:
                        continue
                    blk += w[qL] * w[qN] * norm * (1 - d2) ** 2 * B(eL, qL).T @ C @ B(eN, qN)
            ref[np.ix_(dofL, dofN)] += blk
    assert np.allclose(K, ref, atol=1e-13 * np.abs(ref).max())


# ---- thermal load and thermal strains --------------------------------------------------------------
EXPANSIONS = {"isotropic": 0.02, "orthotropic": (0.03, -0.01), "anisotropic": (0.03, -0.01, 0.015)}
DT = 1.7


def free_expansion(expansion, dt=DT):
    """u = (alpha dT) . x: the displacement of free thermal expansion (tensor shear alpha_xy)."""
    a = [expansion] * 2 + [0.0] if not hasattr(expansion, "__len__") else list(expansion) + [0.0] * (3 - len(expansion))
    return lambda x, y: (dt * (a[0] * x + a[2] * y), dt * (a[2] * x + a[1] * y))


def _with_expansion(material, expansion):
    m = A.ElasticMaterial.from_buffer_copy(material)
    A._set_expansion(m, expansion)
    return m


@pytest.mark.parametrize("ename", list(EXPANSIONS))
@pytest.mark.parametrize("name", list(MATERIALS))
@pytest.mark.parametrize("nn", [4, 8])
def test_free_thermal_expansion_is_stress_free(ename, name, nn):
    """Uniform temperature change, every side held at the free-expansion field: that field solves the
    problem (the thermal load of each interior node equals the elastic force of the field), and its
    stresses vanish once the thermal strain is subtracted."""
    mesh = _mesh(nn)
    mat = _with_expansion(MATERIALS[name], EXPANSIONS[ename])
    field = free_expansion(EXPANSIONS[ename])
    bcs, values = _all_dirichlet(field)
    dT = np.full(mesh.n_nodes, DT)
    res = O.elastic_problem(mesh, [A.model(1.0, 0.0)], [mat], bcs, tol=1e-15, values=values, delta_temperature=dT)
    ux, uy = field(mesh.xy[:, 0], mesh.xy[:, 1])
    x = H.direct_solution(res["K"], res["f"])
    assert np.abs(x[0::2] - ux).max() < 1e-13 and np.abs(x[1::2] - uy).max() < 1e-13
    strain, stress, q = res["mesh"].strain_stress([A.model(1.0, 0.0)], [mat], x, dT)
    assert np.abs(q).max() < 1e-12 * np.abs(voigt_of(mat)).max()
    # without the temperature change the same field is stressed
    _, _, q0 = res["mesh"].strain_stress([A.model(1.0, 0.0)], [mat], x)
    assert np.abs(q0).max() > 1e-3


@pytest.mark.parametrize("ename", ["isotropic", "orthotropic"])
@pytest.mark.parametrize("name", ["isotropic", "orthotropic"])
def test_free_thermal_expansion_nonlocal(ename, name):
    """For a nonlocal group the thermal load and the stiffness share their influence integrals, so
    K u = f for the free-expansion field row by row when the thermal stress has no shear part."""
    mesh = meshgen.structured_q4(10, 0.15, seed=33)
    mat = _with_expansion(MATERIALS[name], EXPANSIONS[ename])
    field = free_expansion(EXPANSIONS[ename])
    bcs, values = _all_dirichlet(field)
    model = A.model(0.4, 0.25, p=2, q=2)
    dT = np.full(mesh.n_nodes, DT)
    res = O.elastic_problem(mesh, [model], [mat], bcs, tol=1e-15, values=values, delta_temperature=dT)
    ux, uy = field(mesh.xy[:, 0], mesh.xy[:, 1])
    x = H.direct_solution(res["K"], res["f"])
    assert np.abs(x[0::2] - ux).max() < 1e-12 and np.abs(x[1::2] - uy).max() < 1e-12


def test_nonlocal_thermal_load_dense():
    """The nonlocal thermal load by a dense numpy double loop in the reference's form
    (temperature_condition_2d.hpp:106-124: w_qL D_i,x (S_xx + S_xy), w_qL D_i,y (S_yy + S_xy) with
    S = sum_qNL w_qNL phi |J| C alpha dT) plus the local part, against the oracle, with a varying
    temperature and anisotropic Hooke and expansion."""
    mesh = meshgen.structured_q4(3, 0.15, seed=9)
    om = O.OracleMesh(mesh)
    r, p1 = 0.6, 0.3
    model = A.model(p1, r, search_radius=3.0, p=2, q=2)
    om.find_neighbours([[3.0, 3.0]])
    mat = _with_expansion(MATERIALS["anisotropic_0.4"], EXPANSIONS["anisotropic"])
    rng = np.random.default_rng(4)
    dT = rng.uniform(-1, 2, mesh.n_nodes)
    f = om.temperature_condition([model], [mat], dT)
    xq, jac, D = om.quadrature()
    w, qN = O.tables(4)[:2]
    C = voigt_of(mat)
    a = np.array(EXPANSIONS["anisotropic"])
    tq = dT[mesh.elems] @ qN                                             # (E, nq)
    sig = np.einsum("ij,eqj->eqi", C, tq[..., None] * np.array([a[0], a[1], 2 * a[2]]))   # (E, nq, 3)
    norm = 3.0 / (math.pi * r * r)
    ref = np.zeros(2 * mesh.n_nodes)
    for eL in range(mesh.n_elems):
        S = np.zeros((4, 3))
        for qL in range(4):
            for eN in range(mesh.n_elems):
                for qN_ in range(4):
                    d2 = np.sum((xq[eL, qL] - xq[eN, qN_]) ** 2) / (r * r)
                    if d2 < 1:
                        S[qL] += w[qN_] * norm * (1 - d2) ** 2 * jac[eN, qN_] * sig[eN, qN_]
        for i, v in enumerate(mesh.elems[eL]):
            for q in range(4):
                dx, dy = w[q] * D[eL, i, q]
                ref[2 * v] += (1 - p1) * dx * (S[q, 0] + S[q, 2]) + p1 * (dx * sig[eL, q, 0] + dy * sig[eL, q, 2])
                ref[2 * v + 1] += (1 - p1) * dy * (S[q, 1] + S[q, 2]) + p1 * (dy * sig[eL, q, 1] + dx * sig[eL, q, 2])
    assert np.allclose(f, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_group_without_expansion_takes_no_load():
    mesh, _ = meshgen.composite_q4(8, count=1, seed=3)
    om = O.OracleMesh(mesh)
    mats = [_with_expansion(MATERIALS["isotropic"], 0.02), MATERIALS["orthotropic"]]
    f = om.temperature_condition([A.model(1.0, 0.0)] * 2, mats, np.ones(mesh.n_nodes))
    only = [_with_expansion(MATERIALS["isotropic"], 0.02), _with_expansion(MATERIALS["orthotropic"], 0.0)]
    g = om.temperature_condition([A.model(1.0, 0.0)] * 2, only, np.ones(mesh.n_nodes))
    assert np.array_equal(f, g) and np.abs(f).max() > 0
