"""Shared helpers for the parity tests: run a problem through the oracle and through libnlf.so."""
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
