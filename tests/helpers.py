"""Shared helpers for the parity tests: run a problem through the oracle and through libnlf.so."""
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from nonlocal_amd import _abi as A
from oracle import oracle as O


def radii(models):
    return [[m.search_radius[0] or m.radius[0], m.search_radius[1] or m.radius[1]] if m.local_weight < 0.999
            else [0.0, 0.0] for m in models]


def oracle_assembly(problem, part=A.NLF_PART_FULL, convection=True):
    om = O.OracleMesh(problem.mesh)
    om.find_neighbours(radii(problem.models))
    f, xD, inner = om.boundary(problem.bcs, dof=problem.dof)
    K = om.assemble(problem.physics, problem.models, problem.materials, inner, part=part)
    if convection and problem.physics == 0 and part == A.NLF_PART_FULL:
        om.convection(K, problem.bcs, inner)
    f = O.first_kind_rhs(K, inner, xD, f)
    return om, K, f, xD, inner


def upper_to_full(csr, n):
    U = sp.csr_matrix((csr.val, csr.inner, csr.outer), shape=(n, n))
    return (U + sp.triu(U, 1).T).tocsr()


def direct_solution(K, f):
    """FP64 direct solve of the oracle's system (the exact answer CG converges to)."""
    n = f.size
    M = upper_to_full(K.export(0), n).tocsc()
    return spla.spsolve(M, f)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---- the reference's solver-level known-answer tests on its own SU2 meshes ------------------------
MESH_DIR = os.path.join(os.path.dirname(__file__), "golden", "meshes")


def norm_error(actual, expected):
    """unit_tests::norm_error (tests/utils/error.hpp:22-26): max |a - e| / max |e| over the nodes."""
    actual = np.asarray(actual, dtype=np.float64)
    expected = np.asarray(expected, dtype=np.float64)
    return float(np.abs(actual - expected).max() / np.abs(expected).max())


def ring_problem(composite: bool):
    """isotropic_solid_ring.cpp:30-36 / isotropic_composite_ring.cpp:36-45: T = 0 on "Inner", T = 1 on
    "Outer", other boundaries insulated, local theory; lambda = 1, or 1 (Inner_Material) and 2
    (Outer_Material).  Returns (mesh, models, materials, bcs, expected T(x, y), expected flux(x, y))."""
    from nonlocal_amd import su2
    mesh = su2.read(os.path.join(MESH_DIR, "composite_ring.su2" if composite else "solid_ring.su2"))
    bcs = [A.boundary() for _ in mesh.bgroup_names]
    bcs[mesh.bgroup_names.index("Inner")] = A.boundary(A.NLF_BC_FIRST_KIND, 0.0)
    bcs[mesh.bgroup_names.index("Outer")] = A.boundary(A.NLF_BC_FIRST_KIND, 1.0)
    models = [A.model(1.0, 0.0) for _ in mesh.group_names]
    lam = {"DEFAULT": 1.0, "Inner_Material": 1.0, "Outer_Material": 2.0}
    materials = [A.thermal_material(A.NLF_CONDUCTIVITY_ISOTROPIC, (lam[g], 0.0, 0.0)) for g in mesh.group_names]
    # analytic solutions of the reference tests (T_in = 0, T_out = 1, radii 0.5 | 0.75 | 1)
    if composite:
        k = 2.0 / 1.0
        coeff = 1.0 / ((k - 1.0) * np.log(0.75) - k * np.log(0.5))

        def temperature(x, y):
            r = np.hypot(x, y)
            return np.where(r < 0.75, 1.0 - 1.0 + k * coeff * np.log(r / 0.5), 1.0 + coeff * np.log(r))

        def flux(x, y):
            r2 = x * x + y * y
            return -coeff * k * x / r2, -coeff * k * y / r2
    else:
        coeff = 1.0 / np.log(1.0 / 0.5)

        def temperature(x, y):
            return 1.0 + coeff * np.log(np.hypot(x, y))

        def flux(x, y):
            r2 = x * x + y * y
            return -x * coeff / r2, -y * coeff / r2
    return mesh, models, materials, bcs, temperature, flux


# tolerances of the reference tests: (temperature, flux)
RING_TOLERANCES = {False: (1.4e-3, 1.7e-2), True: (1.2e-3, 2.1e-2)}
