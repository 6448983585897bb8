"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

ctypes front end of oracle/liboracle.so, the FP64 OpenMP C++ restatement of NonLocFEM's 2D
nonlocal hot path.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from nonlocal_amd import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i64 = C.c_int64
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_mesh_create.restype = C.c_void_p
        L.orc_mesh_create.argtypes = [_i64, _f64p, _i64, C.c_int, _i32p, C.c_int, _i64p, _i64, C.c_int, _i32p,
                                      C.c_int, _i64p]
        L.orc_mesh_destroy.argtypes = [C.c_void_p]
        L.orc_centres.argtypes = [C.c_void_p, _f64p]
        L.orc_quadrature.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
        L.orc_tables.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.orc_basis.argtypes = [C.c_int, C.c_double, C.c_double, _f64p, _f64p, _f64p]
        L.orc_find_neighbours.restype = _i64
        L.orc_find_neighbours.argtypes = [C.c_void_p, _f64p, C.c_int, _f64p]
        L.orc_find_neighbours_ex.restype = _i64
        L.orc_find_neighbours_ex.argtypes = [C.c_void_p, _f64p, C.c_void_p, C.c_int, _f64p]
        L.orc_neighbours_export.argtypes = [C.c_void_p, _i64p, _i32p]
        L.orc_influence.restype = C.c_double
        L.orc_influence.argtypes = [C.POINTER(_abi.Model), _f64p, _f64p]
        L.orc_influence_norm.restype = C.c_double
        L.orc_influence_norm.argtypes = [C.POINTER(_abi.Model)]
        L.orc_influence_points.argtypes = [C.POINTER(_abi.Model), _f64p, _i64, _f64p, _f64p]
        L.orc_distance_points.argtypes = [C.POINTER(_abi.Distance), _f64p, _f64p, _i64, _f64p, _f64p]
        L.orc_assemble.restype = C.c_void_p
        L.orc_assemble.argtypes = [C.c_void_p, C.c_int, C.POINTER(_abi.Model), C.c_void_p, _u8p, C.c_int, C.c_int,
                                   C.c_int]
        L.orc_assemble_ex.restype = C.c_void_p
        L.orc_assemble_ex.argtypes = [C.c_void_p, C.c_int, C.POINTER(_abi.Model), C.c_void_p, _u8p, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.POINTER(C.POINTER(C.c_double))]
        L.orc_matrix_destroy.argtypes = [C.c_void_p]
        L.orc_matrix_info.argtypes = [C.c_void_p, C.c_int, _i64p, _i64p]
        L.orc_matrix_export.argtypes = [C.c_void_p, C.c_int, _i64p, _i64p, _f64p]
        L.orc_matrix_set_values.argtypes = [C.c_void_p, C.c_int, _f64p]
        L.orc_matrix_export_rows.argtypes = [C.c_void_p, C.c_int, _i64, _i64, _i64p, _i64p, _f64p]
        L.orc_boundary.argtypes = [C.c_void_p, C.c_int, C.POINTER(_abi.Boundary), _f64p, _f64p, _u8p]
        L.orc_convection.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(_abi.Boundary), _u8p]
        L.orc_first_kind_rhs.argtypes = [C.c_void_p, _u8p, _f64p, _f64p]
        L.orc_spmv.argtypes = [C.c_void_p, _f64p, _f64p]
        L.orc_precond_create.restype = C.c_void_p
        L.orc_precond_create.argtypes = [C.c_int, C.c_void_p]
        L.orc_precond_destroy.argtypes = [C.c_void_p]
        L.orc_precond_ok.argtypes = [C.c_void_p, _f64p]
        L.orc_precond_apply.argtypes = [C.c_void_p, _f64p, _f64p, _i64]
        L.orc_precond_factor.argtypes = [C.c_void_p, _i64p, _i64p, _f64p, _f64p]
        L.orc_cg.restype = _i64
        L.orc_cg.argtypes = [C.c_void_p, C.c_void_p, _f64p, _f64p, C.c_double, _i64, _f64p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_nearest_qnodes.argtypes = [C.c_int, _i32p]
        L.orc_element_areas.argtypes = [C.c_void_p, _f64p]
        L.orc_qnodes_to_nodes.argtypes = [C.c_void_p, _f64p, _f64p]
        L.orc_strain_stress.restype = C.c_int
        L.orc_strain_stress.argtypes = [C.c_void_p, C.POINTER(_abi.Model), C.POINTER(_abi.ElasticMaterial), _f64p,
                                        _f64p, _f64p, _f64p, _f64p]
        L.orc_hooke.restype = C.c_int
        L.orc_hooke.argtypes = [C.POINTER(_abi.ElasticMaterial), C.POINTER(C.c_int), _f64p]
        L.orc_temperature_condition.restype = C.c_int
        L.orc_temperature_condition.argtypes = [C.c_void_p, C.POINTER(_abi.Model), C.POINTER(_abi.ElasticMaterial),
                                                _f64p, _f64p]
        L.orc_boundary_fn.argtypes = [C.c_void_p, C.c_int, C.POINTER(_abi.Boundary), C.c_void_p, C.c_void_p, _f64p,
                                      _f64p, _u8p]
        L.orc_right_part.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, _f64p]
        L.orc_heat_flux.restype = C.c_int
        L.orc_heat_flux.argtypes = [C.c_void_p, C.POINTER(_abi.Model), C.POINTER(_abi.ThermalMaterial), _f64p, _f64p,
                                    _f64p]
        L.orc_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def set_threads(n: int) -> int:
    return lib().orc_set_threads(int(n))


def tables(nn: int):
    nq = 4 if nn == 4 else 9
    w = np.zeros(nq)
    qN = np.zeros(nn * nq)
    qx = np.zeros(nn * nq)
    qy = np.zeros(nn * nq)
    N0 = np.zeros(nn)
    lib().orc_tables(nn, _ptr(w, _f64p), _ptr(qN, _f64p), _ptr(qx, _f64p), _ptr(qy, _f64p), _ptr(N0, _f64p))
    return w, qN.reshape(nn, nq), qx.reshape(nn, nq), qy.reshape(nn, nq), N0


def nearest_qnodes(nn: int) -> np.ndarray:
    """Index of the quadrature point nearest to each node of the reference element."""
    out = np.zeros(nn, dtype=np.int32)
    lib().orc_nearest_qnodes(nn, _ptr(out, _i32p))
    return out


def basis(nn: int, x: float, y: float):
    """Shape functions and their xi/eta derivatives at (x, y) of the reference element."""
    N = np.zeros(nn)
    Nx = np.zeros(nn)
    Ny = np.zeros(nn)
    lib().orc_basis(nn, x, y, _ptr(N, _f64p), _ptr(Nx, _f64p), _ptr(Ny, _f64p))
    return N, Nx, Ny


@dataclass
class CSR:
    outer: np.ndarray
    inner: np.ndarray
    val: np.ndarray

    @property
    def rows(self):
        return self.outer.size - 1

    def dense(self, symmetric_upper=False):
        n = self.rows
        A = np.zeros((n, n))
        for r in range(n):
            for k in range(self.outer[r], self.outer[r + 1]):
                A[r, self.inner[k]] += self.val[k]
        if symmetric_upper:
            A = A + np.triu(A, 1).T
        return A

    def to_scipy(self, n_cols=None):
        import scipy.sparse as sp
        n = self.rows
        return sp.csr_matrix((self.val, self.inner, self.outer), shape=(n, n_cols or n))


class OracleMatrix:
    def __init__(self, handle, rows):
        self.h = handle
        self.rows = rows

    def __del__(self, lib=lib):             # bound now: module globals go first at exit
        if getattr(self, "h", None) and _LIB is not None:
            lib().orc_matrix_destroy(self.h)
            self.h = None

    def export(self, part=0) -> CSR:
        rows = _i64()
        nnz = _i64()
        lib().orc_matrix_info(self.h, part, C.byref(rows), C.byref(nnz))
        outer = np.zeros(rows.value + 1, dtype=np.int64)
        inner = np.zeros(nnz.value, dtype=np.int64)
        val = np.zeros(nnz.value)
        lib().orc_matrix_export(self.h, part, _ptr(outer, _i64p), _ptr(inner, _i64p), _ptr(val, _f64p))
        return CSR(outer, inner, val)

    def export_rows(self, row0, row1, part=0) -> CSR:
        """Rows [row0, row1) of one part, offsets from 0."""
        outer = np.zeros(row1 - row0 + 1, dtype=np.int64)
        lib().orc_matrix_export_rows(self.h, part, row0, row1, _ptr(outer, _i64p), None, None)
        inner = np.zeros(max(int(outer[-1]), 1), dtype=np.int64)
        val = np.zeros(max(int(outer[-1]), 1))
        lib().orc_matrix_export_rows(self.h, part, row0, row1, _ptr(outer, _i64p), _ptr(inner, _i64p),
                                     _ptr(val, _f64p))
        return CSR(outer, inner[:outer[-1]], val[:outer[-1]])

    def set_values(self, val, part=0):
        val = np.ascontiguousarray(val, dtype=np.float64)
        lib().orc_matrix_set_values(self.h, part, _ptr(val, _f64p))

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().orc_spmv(self.h, _ptr(x, _f64p), _ptr(y, _f64p))
        return y


class OracleMesh:
    """Oracle mesh: precompute, centres and neighbour lists (mesh_2d + find_neighbours)."""

    def __init__(self, mesh):
        self.mesh = mesh
        xy = np.ascontiguousarray(mesh.xy, dtype=np.float64)
        el = np.ascontiguousarray(mesh.elems, dtype=np.int32)
        gb = np.ascontiguousarray(mesh.group_begin, dtype=np.int64)
        be = np.ascontiguousarray(mesh.bedges, dtype=np.int32)
        bb = np.ascontiguousarray(mesh.bgroup_begin, dtype=np.int64)
        self._keep = (xy, el, gb, be, bb)
        self.h = lib().orc_mesh_create(mesh.n_nodes, _ptr(xy, _f64p), mesh.n_elems, mesh.nn, _ptr(el, _i32p),
                                       len(gb) - 1, _ptr(gb, _i64p), be.shape[0], be.shape[1], _ptr(be, _i32p),
                                       len(bb) - 1, _ptr(bb, _i64p))
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self, lib=lib):             # bound now: module globals go first at exit
        if getattr(self, "h", None) and _LIB is not None:
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def centres(self):
        c = np.zeros((self.mesh.n_elems, 2))
        lib().orc_centres(self.h, _ptr(c, _f64p))
        return c

    def quadrature(self):
        nq = 4 if self.mesh.nn == 4 else 9
        E, nn = self.mesh.n_elems, self.mesh.nn
        xq = np.zeros((E, nq, 2))
        jac = np.zeros((E, nq))
        D = np.zeros((E, nn, nq, 2))
        lib().orc_quadrature(self.h, _ptr(xq, _f64p), _ptr(jac, _f64p), _ptr(D, _f64p))
        return xq, jac, D

    def find_neighbours(self, radii, strategy=_abi.NLF_EXPAND_MAX, distances=None):
        """radii: (G, 2) search radii per group; distances: one nlf_distance per group (None: Euclidean).
        Returns (offsets, ids, expanded_radii)."""
        radii = np.ascontiguousarray(np.asarray(radii, dtype=np.float64).reshape(-1, 2))
        expanded = np.zeros_like(radii)
        if distances is None:
            total = lib().orc_find_neighbours(self.h, _ptr(radii, _f64p), strategy, _ptr(expanded, _f64p))
        else:
            arr = (_abi.Distance * len(distances))(*distances)
            total = lib().orc_find_neighbours_ex(self.h, _ptr(radii, _f64p), C.cast(arr, C.c_void_p), strategy,
                                                 _ptr(expanded, _f64p))
        if total < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        off = np.zeros(self.mesh.n_elems + 1, dtype=np.int64)
        ids = np.zeros(total, dtype=np.int32)
        lib().orc_neighbours_export(self.h, _ptr(off, _i64p), _ptr(ids, _i32p))
        return off, ids, expanded

    def boundary(self, bcs, dof=1, values=None):
        """bcs: list of nlf_boundary per boundary group; values: optional callable (x, y, group, dof) ->
        value of the first-kind and second-kind (flux / pressure) groups at a point.  Returns
        (f, xD, is_inner)."""
        rows = dof * self.mesh.n_nodes
        arr = _abi.array_of(_abi.Boundary, bcs)
        f = np.zeros(rows)
        xD = np.zeros(rows)
        inner = np.ones(rows, dtype=np.uint8)
        fn = _abi.boundary_callback(values)
        lib().orc_boundary_fn(self.h, dof, arr, C.cast(fn, C.c_void_p) if fn else None, None, _ptr(f, _f64p),
                              _ptr(xD, _f64p), _ptr(inner, _u8p))
        return f, xD, inner

    def assemble(self, physics, models, materials, is_inner, symmetric=True, part=_abi.NLF_PART_FULL,
                 faithful_reset=False, neumann=False, coefficients=None) -> OracleMatrix:
        """neumann: the Lagrange row and column of a Neumann problem (conductivity_matrix_2d::compute
        with is_neumann); coefficients: per group None or the group's values at its quadrature points,
        ((e - group_begin) * nq + q) * ncomp + k."""
        marr = _abi.array_of(_abi.Model, models)
        mat_t = _abi.ThermalMaterial if physics == 0 else _abi.ElasticMaterial
        matarr = _abi.array_of(mat_t, materials)
        inner = np.ascontiguousarray(is_inner, dtype=np.uint8)
        if neumann or coefficients is not None:
            keep = [None if c is None else np.ascontiguousarray(c, dtype=np.float64).ravel()
                    for c in (coefficients or [None] * len(models))]
            arr = (C.POINTER(C.c_double) * len(keep))(*[C.POINTER(C.c_double)() if c is None else _ptr(c, _f64p)
                                                       for c in keep])
            h = lib().orc_assemble_ex(self.h, physics, marr, C.cast(matarr, C.c_void_p), _ptr(inner, _u8p),
                                      int(symmetric), part, int(faithful_reset), int(neumann), arr)
        else:
            h = lib().orc_assemble(self.h, physics, marr, C.cast(matarr, C.c_void_p), _ptr(inner, _u8p),
                                   int(symmetric), part, int(faithful_reset))
        if not h:
            raise RuntimeError(lib().orc_last_error().decode())
        return OracleMatrix(h, inner.size + int(bool(neumann)))

    def right_part(self, dof, source):
        """int s N_i (integrate_right_part, base/right_part_2d.hpp:10-35); source(x, y) -> value (dof 2:
        a pair), evaluated point by point."""
        f = np.zeros(dof * self.mesh.n_nodes)
        fn = _abi.boundary_callback(lambda x, y, g, d: source(x, y) if dof == 1 else source(x, y)[d])
        lib().orc_right_part(self.h, dof, C.cast(fn, C.c_void_p), None, _ptr(f, _f64p))
        return f

    def element_areas(self):
        a = np.zeros(self.mesh.n_elems)
        lib().orc_element_areas(self.h, _ptr(a, _f64p))
        return a

    def qnodes_to_nodes(self, x):
        """Node values from quadrature-point values [E*nq] (mesh_2d_utils.hpp:82-101)."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        out = np.zeros(self.mesh.n_nodes)
        lib().orc_qnodes_to_nodes(self.h, _ptr(x, _f64p), _ptr(out, _f64p))
        return out

    def heat_flux(self, models, materials, T):
        """heat_equation_solution_2d::calc_flux (heat_equation_solution_2d.hpp:99-137).  Needs the
        neighbour lists when a group is nonlocal.  Returns (node flux (2, N), quad-point flux (2, E*nq))."""
        nq = 4 if self.mesh.nn == 4 else 9
        T = np.ascontiguousarray(T, dtype=np.float64)
        q = np.zeros((2, self.mesh.n_elems * nq))
        n = np.zeros((2, self.mesh.n_nodes))
        marr = _abi.array_of(_abi.Model, models)
        matarr = _abi.array_of(_abi.ThermalMaterial, materials)
        if lib().orc_heat_flux(self.h, marr, matarr, _ptr(T, _f64p), _ptr(q, _f64p), _ptr(n, _f64p)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return n, q

    def strain_stress(self, models, materials, u, delta_temperature=None):
        """mechanical_solution_2d::calc_strain_and_stress (mechanical_solution_2d.hpp:192-232).  u: the
        displacement (x, y per node); delta_temperature: node temperature changes whose thermal strains
        are subtracted before the stresses.  Returns (node strains (3, N), node stresses (3, N),
        quad-point stresses (3, E*nq)), components xx, yy, xy."""
        nq = 4 if self.mesh.nn == 4 else 9
        u = np.ascontiguousarray(u, dtype=np.float64)
        dT = None if delta_temperature is None else np.ascontiguousarray(delta_temperature, dtype=np.float64)
        q = np.zeros((3, self.mesh.n_elems * nq))
        ns = np.zeros((3, self.mesh.n_nodes))
        nst = np.zeros((3, self.mesh.n_nodes))
        marr = _abi.array_of(_abi.Model, models)
        matarr = _abi.array_of(_abi.ElasticMaterial, materials)
        if lib().orc_strain_stress(self.h, marr, matarr, _ptr(u, _f64p), None if dT is None else _ptr(dT, _f64p),
                                   _ptr(q, _f64p), _ptr(ns, _f64p), _ptr(nst, _f64p)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return ns, nst, q

    def temperature_condition(self, models, materials, delta_temperature, f=None):
        """temperature_condition (temperature_condition_2d.hpp:128-170): f (2N, x and y per node) plus the
        thermal load of the node temperature changes."""
        f = np.zeros(2 * self.mesh.n_nodes) if f is None else np.array(f, dtype=np.float64)
        dT = np.ascontiguousarray(delta_temperature, dtype=np.float64)
        marr = _abi.array_of(_abi.Model, models)
        matarr = _abi.array_of(_abi.ElasticMaterial, materials)
        if lib().orc_temperature_condition(self.h, marr, matarr, _ptr(dT, _f64p), _ptr(f, _f64p)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return f

    def convection(self, matrix: OracleMatrix, bcs, is_inner):
        arr = _abi.array_of(_abi.Boundary, bcs)
        inner = np.ascontiguousarray(is_inner, dtype=np.uint8)
        lib().orc_convection(self.h, matrix.h, arr, _ptr(inner, _u8p))


def first_kind_rhs(matrix: OracleMatrix, is_inner, xD, f):
    inner = np.ascontiguousarray(is_inner, dtype=np.uint8)
    xD = np.ascontiguousarray(xD, dtype=np.float64)
    f = np.array(f, dtype=np.float64)
    lib().orc_first_kind_rhs(matrix.h, _ptr(inner, _u8p), _ptr(xD, _f64p), _ptr(f, _f64p))
    return f


class OraclePrecond:
    def __init__(self, kind, matrix: OracleMatrix | None):
        self.h = lib().orc_precond_create(kind, matrix.h if matrix is not None else None)
        self.kind = kind

    def __del__(self, lib=lib):
        if getattr(self, "h", None) and _LIB is not None:
            lib().orc_precond_destroy(self.h)
            self.h = None

    def ok(self):
        s = C.c_double()
        return bool(lib().orc_precond_ok(self.h, C.byref(s))), s.value

    def apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        lib().orc_precond_apply(self.h, _ptr(r, _f64p), _ptr(z, _f64p), r.size)
        return z


def cg(matrix: OracleMatrix, precond: OraclePrecond, b, x0=None, tol=1e-15, max_it=10000):
    """conjugate_gradient::solve restated; returns (x, iterations, residual)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    res = C.c_double()
    it = lib().orc_cg(matrix.h, precond.h, _ptr(b, _f64p), _ptr(x, _f64p), tol, max_it, C.byref(res))
    return x, it, res.value


# ---------------------------------------------------------------------------------------------
# Problem drivers (stationary_heat_equation_solver_2d.hpp:46-120, equilibrium_equation_2d.hpp:20-62)
# ---------------------------------------------------------------------------------------------
def is_neumann(bcs):
    """init_problem_settings (thermal/init_problem_settings.hpp:43): a stationary problem whose conditions
    are all fluxes (insulated groups hold none) is a Neumann problem."""
    return all(b.kind[0] in (_abi.NLF_BC_NONE, _abi.NLF_BC_FLUX) for b in bcs)


NEUMANN_PROBLEM_ERROR_THRESHOLD = 1e-10          # nonlocal_constants.hpp:45


def thermal_problem(mesh, models, materials, bcs, precond=_abi.NLF_PRECOND_ILLT, tol=1e-12, max_it=100000,
                    radii=None, faithful_reset=False, values=None, source=None, energy=0.0, coefficients=None):
    """Solve the stationary thermal problem; returns a dict with neighbours, matrices, f and T.  A
    Neumann problem (every condition a flux) is solved with the Lagrange row (f[N] = energy,
    stationary_heat_equation_solver_2d.hpp:28-44); T is then the first N entries."""
    om = OracleMesh(mesh)
    if radii is None:
        radii = [[m.search_radius[0] or m.radius[0], m.search_radius[1] or m.radius[1]]
                 if m.local_weight < 0.999 else [0.0, 0.0] for m in models]
    off, ids, expanded = om.find_neighbours(radii)
    f, xD, inner = om.boundary(bcs, dof=1, values=values)
    neumann = is_neumann(bcs)
    K = om.assemble(0, models, materials, inner, symmetric=True, faithful_reset=faithful_reset, neumann=neumann,
                    coefficients=coefficients)
    om.convection(K, bcs, inner)
    if source is not None:
        f = f + om.right_part(1, source)
    if neumann:
        if abs(np.sum(f)) > NEUMANN_PROBLEM_ERROR_THRESHOLD:
            raise ValueError("It's unsolvable Neumann problem.")
        f = np.append(f, energy)
        inner = np.append(inner, 1).astype(np.uint8)
        xD = np.append(xD, 0.0)
    f = first_kind_rhs(K, inner, xD, f)
    if precond == _abi.NLF_PRECOND_ILLT:
        KL = om.assemble(0, models, materials, inner[:mesh.n_nodes], symmetric=True, part=_abi.NLF_PART_LOCAL,
                         neumann=neumann, coefficients=coefficients)
        P = OraclePrecond(precond, KL)
        if not P.ok()[0]:
            P = OraclePrecond(_abi.NLF_PRECOND_IDENTITY, None)
    else:
        KL = None
        P = OraclePrecond(precond, K if precond == _abi.NLF_PRECOND_JACOBI else None)
    x, it, res = cg(K, P, f, tol=tol, max_it=max_it)
    return dict(mesh=om, neighbours=(off, ids, expanded), K=K, KL=KL, f=f, xD=xD, inner=inner, T=x[:mesh.n_nodes],
                x=x, iterations=it, residual=res, neumann=neumann)


def elastic_problem(mesh, models, materials, bcs, tol=1e-12, max_it=100000, radii=None, values=None,
                    delta_temperature=None):
    """equilibrium_equation (equilibrium_equation_2d.hpp:20-62): second-kind conditions, the thermal load
    of delta_temperature (node values) when given, first-kind conditions, CG with ILLT when nonlocal."""
    om = OracleMesh(mesh)
    if radii is None:
        radii = [[m.radius[0], m.radius[1]] if m.local_weight < 0.999 else [0.0, 0.0] for m in models]
    off, ids, expanded = om.find_neighbours(radii)
    f, xD, inner = om.boundary(bcs, dof=2, values=values)
    K = om.assemble(1, models, materials, inner, symmetric=True)
    if delta_temperature is not None:
        f = om.temperature_condition(models, materials, delta_temperature, f)
    f = first_kind_rhs(K, inner, xD, f)
    nonlocal_ = any(m.local_weight < 0.999 for m in models)
    if nonlocal_:
        KL = om.assemble(1, models, materials, inner, symmetric=True, part=_abi.NLF_PART_LOCAL)
        P = OraclePrecond(_abi.NLF_PRECOND_ILLT, KL)
        if not P.ok()[0]:
            P = OraclePrecond(_abi.NLF_PRECOND_IDENTITY, None)
    else:
        KL = None
        P = OraclePrecond(_abi.NLF_PRECOND_IDENTITY, None)
    u, it, res = cg(K, P, f, tol=tol, max_it=max_it)
    return dict(mesh=om, neighbours=(off, ids, expanded), K=K, KL=KL, f=f, xD=xD, inner=inner, u=u,
                iterations=it, residual=res)


def hooke(material):
    """(kind, values) of the Hooke matrix of a material, in the reference's index order of the kind:
    isotropic {11, 12, 66}, orthotropic {11, 12, 66, 22}, anisotropic {11, 12, 66, 22, 16, 26}."""
    kind = C.c_int()
    h = np.zeros(6)
    if lib().orc_hooke(C.byref(material), C.byref(kind), _ptr(h, _f64p)) != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return kind.value, h[:{0: 3, 1: 4, 2: 6}[kind.value]]


def influence(model, x, y) -> float:
    """phi(x, y) of one model (influence_functions_2d.hpp:39-148), as the assembly evaluates it."""
    xa = np.ascontiguousarray(x, dtype=np.float64)
    ya = np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_influence(C.byref(model), _ptr(xa, _f64p), _ptr(ya, _f64p))


def influence_points(model, x, ys) -> np.ndarray:
    """phi(x, y) for every row y of ys (k, 2)."""
    xa = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    ya = np.ascontiguousarray(np.asarray(ys, dtype=np.float64).reshape(-1, 2))
    out = np.empty(ya.shape[0])
    if lib().orc_influence_points(C.byref(model), _ptr(xa, _f64p), ya.shape[0], _ptr(ya, _f64p), _ptr(out, _f64p)):
        raise RuntimeError(lib().orc_last_error().decode())
    return out


def distance_points(dist, x, r, ys) -> np.ndarray:
    """The search distance(x, y, r) of an nlf_distance for every row y of ys (k, 2)."""
    xa = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    ra = np.ascontiguousarray(np.asarray(r, dtype=np.float64))
    ya = np.ascontiguousarray(np.asarray(ys, dtype=np.float64).reshape(-1, 2))
    out = np.empty(ya.shape[0])
    if lib().orc_distance_points(C.byref(dist), _ptr(xa, _f64p), _ptr(ra, _f64p), ya.shape[0], _ptr(ya, _f64p),
                                 _ptr(out, _f64p)):
        raise RuntimeError(lib().orc_last_error().decode())
    return out


# ---------------------------------------------------------------------------------------------
# Nonstationary heat conduction (nonstationary_heat_equation_solver_2d.hpp:17-106), restated with
# the oracle's matrices and scipy's sparse direct solver.
# ---------------------------------------------------------------------------------------------
def capacity_matrix(om: OracleMesh, capacities, is_inner) -> OracleMatrix:
    """heat_capacity_matrix_2d::calc_matrix (heat_capacity_matrix_2d.hpp:50-62): rho c int N_i N_j."""
    models = [_abi.model(1.0, 0.0) for _ in capacities]
    marr = _abi.array_of(_abi.Model, models)
    caps = (C.c_double * len(capacities))(*[float(v) for v in capacities])
    inner = np.ascontiguousarray(is_inner, dtype=np.uint8)
    h = lib().orc_assemble(om.h, 2, marr, C.cast(caps, C.c_void_p), _ptr(inner, _u8p), 1, _abi.NLF_PART_LOCAL, 0)
    if not h:
        raise RuntimeError(lib().orc_last_error().decode())
    return OracleMatrix(h, inner.size)


def transient(mesh, models, materials, capacities, bcs, tau, T0, steps, values=None, source=None):
    """compute() then `steps` calc_step()s: inner() := tau K + C with the Dirichlet diagonal set to 1
    (:76-83), bound() := tau K_bound; each step rhs = tau (f_2nd + f_source) + C_inner T_n (:96-101),
    f -= K_bound x_D and f[D] = x_D (:102), then the linear system (solved directly here; the
    reference runs CG to 1e-15).  Returns the list of temperatures after each step."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    om = OracleMesh(mesh)
    radii = [[m.search_radius[0] or m.radius[0], m.search_radius[1] or m.radius[1]]
             if m.local_weight < 0.999 else [0.0, 0.0] for m in models]
    om.find_neighbours(radii)
    n = mesh.n_nodes
    f2, xD, inner = om.boundary(bcs, dof=1, values=values)
    K = om.assemble(0, models, materials, inner)
    om.convection(K, bcs, inner)
    Cm = capacity_matrix(om, capacities, inner)

    def full(csr):
        U = sp.csr_matrix((csr.val, csr.inner, csr.outer), shape=(n, n))
        return (U + sp.triu(U, 1).T).tocsr()

    A = (tau * full(K.export(0)) + full(Cm.export(0))).tolil()
    dirichlet = np.flatnonzero(inner == 0)
    for r in dirichlet:
        A[r, r] = 1.0
    A = A.tocsc()
    kb = K.export(1)
    Kb = tau * sp.csr_matrix((kb.val, kb.inner, kb.outer), shape=(n, n))
    Cf = full(Cm.export(0))
    fs = om.right_part(1, source) if source is not None else np.zeros(n)
    lu = spla.splu(A)
    T = np.array(T0, dtype=np.float64)
    out = []
    for _ in range(steps):
        rhs = tau * (f2 + fs) + Cf @ T
        rhs = rhs - Kb @ xD
        rhs[dirichlet] = xD[dirichlet]
        T = lu.solve(rhs)
        out.append(T.copy())
    return out
