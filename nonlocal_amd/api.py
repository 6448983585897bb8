"""Python mirror of NonLocFEM's 2D solver/assembler API over the C ABI (include/nlf.h).

Names follow the reference: mesh_2d (libraries/mesh/mesh_2d/mesh_2d.hpp:18-81), find_neighbours
(find_neighbours.hpp:119-160), conductivity_matrix_2d (thermal/conductivity_matrix_2d.hpp:11-33),
stiffness_matrix (mechanical/stiffness_matrix_2d.hpp:10-38), conjugate_gradient
(slae/conjugate_gradient.hpp:15-42), stationary_heat_equation_solver_2d
(thermal/stationary_heat_equation_solver_2d.hpp:46-120) and equilibrium_equation
(mechanical/equilibrium_equation_2d.hpp:20-62).  Device memory and the stream come from torch; all
compute runs in libnlf.so.  Errors raise NLFError (the reference throws std::domain_error).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi
from ._abi import (NLF_ERR_STATE, NLF_EXPAND_MAX, NLF_MODE_FP64, NLF_PART_FULL, NLF_PART_LOCAL, NLF_PRECOND_IDENTITY,
                   NLF_PRECOND_ILLT, NLF_PRECOND_JACOBI)
from ._lib import check, load, NLFError  # noqa: F401

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _destroy(name, h):
    """Release a library handle; a no-op while the interpreter is shutting down."""
    try:
        getattr(load(), name)(h)
    except Exception:       # module globals are already torn down at interpreter exit
        pass


def _dev(t: torch.Tensor):
    assert t.is_cuda and t.is_contiguous()
    return C.c_void_p(t.data_ptr())


class Context:
    """One library context bound to torch's current stream on `device`."""

    def __init__(self, device: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU available: the nonlocal hot path runs on the GPU only")
        L = load()
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        # One ordered stream for torch's allocations/copies and the library's kernels: torch's
        # default stream is the legacy null stream (handle 0), which the C ABI reads as "create a
        # private stream", so a dedicated default-priority stream is made current for torch too.
        self.stream = torch.cuda.Stream(self.device)
        torch.cuda.set_stream(self.stream)
        h = C.c_void_p()
        check(L.nlf_ctx_create(device, C.c_void_p(self.stream.cuda_stream), C.byref(h)))
        self.h = h

    def check(self, code):
        check(code, self.h)

    def synchronize(self):
        self.check(load().nlf_ctx_synchronize(self.h))

    def set_option(self, option, value):
        self.check(load().nlf_ctx_set_option(self.h, option, int(value)))

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _destroy("nlf_ctx_destroy", self.h)
            except TypeError:      # module globals already cleared at interpreter exit
                pass
            self.h = None


_DEFAULT_CTX = None


def default_context() -> Context:
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


class mesh_2d:
    """Device-resident mesh with the mesh_2d::init precompute (quad coords, |J|, detJ * grad N)."""

    def __init__(self, mesh, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.mesh = mesh
        xy = np.ascontiguousarray(mesh.xy, dtype=np.float64)
        el = np.ascontiguousarray(mesh.elems, dtype=np.int32)
        gb = np.ascontiguousarray(mesh.group_begin, dtype=np.int64)
        be = np.ascontiguousarray(mesh.bedges, dtype=np.int32)
        bb = np.ascontiguousarray(mesh.bgroup_begin, dtype=np.int64)
        h = C.c_void_p()
        self.ctx.check(load().nlf_mesh_create(self.ctx.h, mesh.n_nodes, _ptr(xy, _f64p), mesh.n_elems, mesh.nn,
                                              _ptr(el, _i32p), len(gb) - 1, _ptr(gb, _i64p), be.shape[0], be.shape[1],
                                              _ptr(be, _i32p), len(bb) - 1, _ptr(bb, _i64p), C.byref(h)))
        self.h = h
        self.influences = None

    def __del__(self, _destroy=_destroy):      # bound now: module globals go first at exit
        if getattr(self, "h", None):
            _destroy("nlf_mesh_destroy", self.h)
            self.h = None

    @property
    def nodes_count(self):
        return self.mesh.n_nodes

    @property
    def elements_count(self):
        return self.mesh.n_elems

    def quadrature(self):
        nq = 4 if self.mesh.nn == 4 else 9
        E, nn = self.mesh.n_elems, self.mesh.nn
        xq = np.zeros((E, nq, 2))
        jac = np.zeros((E, nq))
        D = np.zeros((E, nn, nq, 2))
        c = np.zeros((E, 2))
        self.ctx.check(load().nlf_mesh_export_quadrature(self.h, _ptr(xq, _f64p), _ptr(jac, _f64p), _ptr(D, _f64p),
                                                         _ptr(c, _f64p)))
        return xq, jac, D, c

    def neighbours(self):
        """Host copy of the neighbour lists as CSR (offsets int64[E+1], ids int32)."""
        total = C.c_int64()
        self.ctx.check(load().nlf_neighbours_count(self.h, C.byref(total)))
        off = np.zeros(self.mesh.n_elems + 1, dtype=np.int64)
        ids = np.zeros(max(total.value, 1), dtype=np.int32)
        self.ctx.check(load().nlf_neighbours_export(self.h, _ptr(off, _i64p), _ptr(ids, _i32p)))
        return off, ids[:total.value]


def find_neighbours(mesh: mesh_2d, search_radii, strategy=NLF_EXPAND_MAX, distances=None):
    """find_neighbours + mesh_2d::neighbours: search_radii (G, 2); distances: one nlf_distance per group
    (_abi.distance / _abi.distance_of), None = Euclidean; returns the expanded radii."""
    r = np.ascontiguousarray(np.asarray(search_radii, dtype=np.float64).reshape(-1, 2))
    expanded = np.zeros_like(r)
    if distances is None:
        mesh.ctx.check(load().nlf_find_neighbours(mesh.h, _ptr(r, _f64p), strategy, _ptr(expanded, _f64p)))
    else:
        if len(distances) != r.shape[0]:
            raise ValueError("one distance per group")
        arr = _abi.array_of(_abi.Distance, distances)
        mesh.ctx.check(load().nlf_find_neighbours_ex(mesh.h, _ptr(r, _f64p), arr, strategy, _ptr(expanded, _f64p)))
    mesh.influences = expanded
    return expanded


@dataclass
class CSRHost:
    outer: np.ndarray
    inner: np.ndarray
    val: np.ndarray


class finite_element_matrix:
    """Device matrix handle: inner (upper view of the full storage) and bound parts."""

    def __init__(self, h, ctx: Context, dof: int):
        self.h = h
        self.ctx = ctx
        self.dof = dof

    def __del__(self, _destroy=_destroy):      # bound now: module globals go first at exit
        if getattr(self, "h", None):
            _destroy("nlf_matrix_destroy", self.h)
            self.h = None

    def info(self):
        rows, up, full, bound = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        self.ctx.check(load().nlf_matrix_info(self.h, C.byref(rows), C.byref(up), C.byref(full), C.byref(bound)))
        return dict(rows=rows.value, nnz_upper=up.value, nnz_full=full.value, nnz_bound=bound.value)

    def symv_info(self):
        """(slots, strip doubles) of the symmetric SpMV storage, built here if it applies; None when it
        does not (the products then read the full storage)."""
        a, b = C.c_int64(), C.c_int64()
        code = load().nlf_matrix_symv_info(self.h, C.byref(a), C.byref(b))
        if code == NLF_ERR_STATE:
            return None
        self.ctx.check(code)
        return a.value, b.value

    def pattern_info(self):
        """(distinct offset rows, their total length) of the SpMV's row-pattern dictionary; (0, 0)
        when the SpMV reads per-nonzero offsets."""
        g, e = C.c_int64(), C.c_int64()
        self.ctx.check(load().nlf_matrix_pattern_info(self.h, C.byref(g), C.byref(e)))
        return g.value, e.value

    def timings(self):
        a, b = C.c_double(), C.c_double()
        self.ctx.check(load().nlf_matrix_timings(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def export(self, part=0) -> CSRHost:
        """part 0 = inner (upper CSR, the reference's layout), 1 = bound, 2 = inner full storage;
        3 and 4 = parts 0 and 2 of the FP32 values (make_fp32 / assemble_fp32)."""
        i = self.info()
        nnz = {0: i["nnz_upper"], 1: i["nnz_bound"], 2: i["nnz_full"], 3: i["nnz_upper"], 4: i["nnz_full"]}[part]
        outer = np.zeros(i["rows"] + 1, dtype=np.int64)
        inner = np.zeros(max(nnz, 1), dtype=np.int64)
        val = np.zeros(max(nnz, 1))
        self.ctx.check(load().nlf_matrix_export(self.h, part, _ptr(outer, _i64p), _ptr(inner, _i64p), _ptr(val, _f64p)))
        return CSRHost(outer, inner[:nnz], val[:nnz])

    def export_rows(self, row0, row1) -> CSRHost:
        """Rows [row0, row1) of the inner upper CSR, offsets from 0 (only those rows are copied)."""
        outer = np.zeros(row1 - row0 + 1, dtype=np.int64)
        self.ctx.check(load().nlf_matrix_export_rows(self.h, row0, row1, _ptr(outer, _i64p), None, None))
        nnz = int(outer[-1])
        inner = np.zeros(max(nnz, 1), dtype=np.int64)
        val = np.zeros(max(nnz, 1))
        self.ctx.check(load().nlf_matrix_export_rows(self.h, row0, row1, _ptr(outer, _i64p), _ptr(inner, _i64p),
                                                     _ptr(val, _f64p)))
        return CSRHost(outer, inner[:nnz], val[:nnz])

    def inner(self):
        return self.export(0)

    def bound(self):
        return self.export(1)

    def spmv(self, x: torch.Tensor, y: torch.Tensor | None = None):
        y = torch.empty_like(x) if y is None else y
        if x.dtype == torch.float64:
            self.ctx.check(load().nlf_spmv(self.h, _dev(x), _dev(y)))
        else:
            self.ctx.check(load().nlf_spmv_fp32(self.h, _dev(x), _dev(y)))
        return y

    def make_fp32(self):
        """FP32 values for mixed-precision inner solves: the FP64 values rounded."""
        self.ctx.check(load().nlf_matrix_make_fp32(self.h))

    def assemble_fp32(self):
        """FP32 values for mixed-precision inner solves: the value pass repeated in FP32 (FP32
        element-pair contractions and accumulation), plus the boundary additions made since."""
        self.ctx.check(load().nlf_matrix_assemble_fp32(self.h))

    def make_split(self):
        """Values for NLF_MODE_SPLIT: FP32 rounding plus a 16-bit remainder (40 significand bits)."""
        self.ctx.check(load().nlf_matrix_make_split(self.h))

    def compare_fp32(self):
        """Distance of the FP32 values from the FP64 ones (device reduction over the full storage):
        {"max_abs", "max_entry", "l2", "norm"}."""
        out = np.zeros(4)
        self.ctx.check(load().nlf_matrix_compare_fp32(self.h, _ptr(out, _f64p)))
        return dict(max_abs=out[0], max_entry=out[1], l2=out[2], norm=out[3])

    def add_convection(self, bcs):
        arr = _abi.array_of(_abi.Boundary, bcs)
        self.ctx.check(load().nlf_matrix_add_convection(self.h, arr))

    def apply_first_kind(self, xD: torch.Tensor, f: torch.Tensor):
        self.ctx.check(load().nlf_apply_first_kind(self.h, _dev(xD), _dev(f)))


PHYSICS_CAPACITY = 2      # nlf_assemble physics of the heat capacity matrix (materials: rho c per group)


def _assemble(mesh: mesh_2d, physics, models, materials, is_inner, symmetric=True, part=NLF_PART_FULL,
              neumann=False, coefficients=None):
    """nlf_assemble, or nlf_assemble_ex with neumann (the Lagrange row and column of a Neumann
    problem) and coefficients (per group None or its values at its quadrature points)."""
    marr = _abi.array_of(_abi.Model, models)
    if physics == PHYSICS_CAPACITY:
        matarr = (C.c_double * len(materials))(*[float(v) for v in materials])
    else:
        mat_t = _abi.ThermalMaterial if physics == 0 else _abi.ElasticMaterial
        matarr = _abi.array_of(mat_t, materials)
    inner = np.ascontiguousarray(is_inner, dtype=np.uint8)
    h = C.c_void_p()
    if neumann or coefficients is not None:
        keep = [None if c is None else np.ascontiguousarray(c, dtype=np.float64).ravel()
                for c in (coefficients or [None] * len(models))]
        arr = (C.POINTER(C.c_double) * len(keep))(*[C.POINTER(C.c_double)() if c is None else _ptr(c, _f64p)
                                                   for c in keep])
        opt = _abi.AssembleOptions()
        opt.neumann = int(bool(neumann))
        opt.coefficients = C.cast(arr, C.POINTER(C.POINTER(C.c_double)))
        mesh.ctx.check(load().nlf_assemble_ex(mesh.h, physics, marr, C.cast(matarr, C.c_void_p), _ptr(inner, _u8p),
                                              int(symmetric), part, C.byref(opt), C.byref(h)))
    else:
        mesh.ctx.check(load().nlf_assemble(mesh.h, physics, marr, C.cast(matarr, C.c_void_p), _ptr(inner, _u8p),
                                           int(symmetric), part, C.byref(h)))
    return finite_element_matrix(h, mesh.ctx, 2 if physics == 1 else 1)


def is_neumann(bcs):
    """init_problem_settings (thermal/init_problem_settings.hpp:43): a stationary thermal problem whose
    conditions are all fluxes (insulated groups hold none) is a Neumann problem."""
    return all(b.kind[0] in (_abi.NLF_BC_NONE, _abi.NLF_BC_FLUX) for b in bcs)


NEUMANN_PROBLEM_ERROR_THRESHOLD = 1e-10          # nonlocal_constants.hpp:45


def combine(K: finite_element_matrix, alpha, Cm: finite_element_matrix, beta):
    """K := alpha K + beta C on K's pattern (nlf_matrix_combine); C's pattern must lie within K's."""
    K.ctx.check(load().nlf_matrix_combine(K.h, float(alpha), Cm.h, float(beta)))
    return K


def right_part(mesh: mesh_2d, dof, source, f: torch.Tensor):
    """f += int s N_i (integrate_right_part, base/right_part_2d.hpp:10-35).  source(x, y) takes arrays of
    quadrature-point coordinates and returns the values (dof 2: a pair of arrays)."""
    xq = mesh.quadrature()[0].reshape(-1, 2)
    v = source(xq[:, 0], xq[:, 1])
    v = np.stack([np.broadcast_to(np.asarray(c, dtype=np.float64), xq.shape[:1]) for c in v], axis=1) if dof == 2 \
        else np.broadcast_to(np.asarray(v, dtype=np.float64), xq.shape[:1])
    vd = torch.from_numpy(np.array(v, dtype=np.float64).ravel()).to(mesh.ctx.device)
    mesh.ctx.check(load().nlf_right_part(mesh.h, dof, _dev(vd), _dev(f)))
    return f


class heat_capacity_matrix_2d:
    """Heat capacity matrix assembler (heat_capacity_matrix_2d<T, I, J>, thermal/heat_capacity_matrix_2d.hpp:
    8-62): C_ij = rho c int N_i N_j, local theory, same separation into inner and bound parts."""

    def __init__(self, mesh: mesh_2d):
        self.mesh = mesh
        self._matrix = None

    def calc_matrix(self, capacities, is_inner):
        """capacities: density * capacity per element group."""
        models = [_abi.model(1.0, 0.0) for _ in capacities]
        self._matrix = _assemble(self.mesh, PHYSICS_CAPACITY, models, capacities, is_inner, True, NLF_PART_LOCAL)
        return self._matrix

    def matrix(self) -> finite_element_matrix:
        return self._matrix


class conductivity_matrix_2d:
    """Thermal conductivity matrix assembler (conductivity_matrix_2d<T, I, J>)."""

    def __init__(self, mesh: mesh_2d):
        self.mesh = mesh
        self._matrix = None

    def compute(self, models, materials, is_inner, is_symmetric=True, is_neumann=False, part=NLF_PART_FULL):
        if is_neumann:
            raise NLFError(-4, "Neumann (Lagrange-multiplier) problems are not implemented")
        self._matrix = _assemble(self.mesh, 0, models, materials, is_inner, is_symmetric, part)
        return self._matrix

    def matrix(self) -> finite_element_matrix:
        return self._matrix


class stiffness_matrix:
    """Plane-elasticity stiffness assembler (stiffness_matrix<T, I, J>), 2 DoF per node."""

    def __init__(self, mesh: mesh_2d):
        self.mesh = mesh
        self._matrix = None

    def compute(self, models, materials, is_inner, is_symmetric=True, part=NLF_PART_FULL):
        self._matrix = _assemble(self.mesh, 1, models, materials, is_inner, is_symmetric, part)
        return self._matrix

    def matrix(self) -> finite_element_matrix:
        return self._matrix


class preconditioner:
    """preconditioner_base implementation on the device: IDENTITY, JACOBI or ILLT (IC(0))."""

    def __init__(self, matrix: finite_element_matrix | None, kind=NLF_PRECOND_IDENTITY, ctx=None):
        self.ctx = matrix.ctx if matrix is not None else (ctx or default_context())
        h = C.c_void_p()
        self.ctx.check(load().nlf_precond_create(matrix.h if matrix is not None else None, kind, C.byref(h)))
        self.h = h
        self.kind = kind

    def __del__(self, _destroy=_destroy):      # bound now: module globals go first at exit
        if getattr(self, "h", None):
            _destroy("nlf_precond_destroy", self.h)
            self.h = None

    def info(self):
        ok, shift, fw, bw, sec = C.c_int(), C.c_double(), C.c_int(), C.c_int(), C.c_double()
        self.ctx.check(load().nlf_precond_info(self.h, C.byref(ok), C.byref(shift), C.byref(fw), C.byref(bw),
                                               C.byref(sec)))
        return dict(ok=bool(ok.value), shift=shift.value, fw_levels=fw.value, bw_levels=bw.value, seconds=sec.value)

    def sweep_info(self):
        """ILLT sweep layout (nlf_precond_sweep_info): per sweep the levels as swept, the compute
        width, the levels of the LDS dependency window and the kernel (0 packed, 1 ELL one level at a
        time, 2 ELL with prefetch)."""
        out = (C.c_int32 * 8)()
        self.ctx.check(load().nlf_precond_sweep_info(self.h, out))
        keys = ("levels", "width", "window_levels", "kernel")
        return {"forward": dict(zip(keys, out[0:4])), "backward": dict(zip(keys, out[4:8]))}

    def blocks(self):
        """Independent diagonal blocks of an ILLT's matrix, each factored and swept by a workgroup
        of its own (nlf_precond_blocks)."""
        out = C.c_int32()
        self.ctx.check(load().nlf_precond_blocks(self.h, C.byref(out)))
        return out.value

    def computation_info(self):
        return self.info()["ok"]

    def solve(self, r: torch.Tensor):
        z = torch.empty_like(r)
        self.ctx.check(load().nlf_precond_apply(self.h, _dev(r), _dev(z)))
        return z


class conjugate_gradient:
    """conjugate_gradient<T, I>: holds a non-owning reference to the matrix, owns its preconditioner."""

    def __init__(self, matrix: finite_element_matrix):
        self._matrix = matrix
        self._precond = None
        self._tol = 1e-15
        self._max_it = 10000
        self.stats = _abi.SolveStats()

    def init_preconditioner(self, kind, matrix: finite_element_matrix | None = None):
        self._precond = preconditioner(matrix if matrix is not None else
                                       (self._matrix if kind == NLF_PRECOND_JACOBI else None), kind,
                                       ctx=self._matrix.ctx)
        return self._precond

    def preconditioner(self):
        return self._precond

    def set_tolerance(self, tol):
        self._tol = float(tol)

    def set_max_interations(self, it):   # reference spelling (iterative_solver_base.hpp:49-51)
        self._max_it = int(it)

    set_max_iterations = set_max_interations

    def tolerance(self):
        return self._tol

    def max_iterations(self):
        return self._max_it

    def solve(self, b: torch.Tensor, x0: torch.Tensor | None = None, mode=NLF_MODE_FP64):
        x = torch.zeros_like(b) if x0 is None else x0.clone()
        st = _abi.SolveStats()
        self._matrix.ctx.check(load().nlf_cg_solve(self._matrix.h, self._precond.h if self._precond else None, _dev(b),
                                                   _dev(x), self._tol, self._max_it, mode, C.byref(st)))
        self.stats = st
        return x

    def iterations(self):
        return self.stats.iterations

    def residual(self):
        return self.stats.residual


@dataclass
class Solution:
    x: torch.Tensor
    iterations: int
    residual: float
    true_residual: float
    matrix: finite_element_matrix = None
    local_matrix: finite_element_matrix = None
    rhs: torch.Tensor = None
    timings: dict = field(default_factory=dict)
    stats: object = None
    solution: object = None        # heat_equation_solution_2d / mechanical_solution_2d (post-processing on demand)


class heat_equation_solution_2d:
    """Temperature and heat flux of a thermal solve (solvers/solver_2d/thermal/
    heat_equation_solution_2d.hpp:10-137), device resident.

    calc_flux() computes, on the GPU, the flux -lambda grad T at the quadrature points, its nonlocal
    average p1 q + (1 - p1) int phi q over each element's neighbours for nonlocal groups, and the
    node values as the area-weighted mean of the nearest quadrature points; it returns the pair
    (flux_x, flux_y) of node vectors and keeps it, like the reference's calc_flux.  The reference's
    solver calls calc_flux before returning; here the thermal solve returns this object with the
    flux not yet computed, so a solve that does not need the flux does not pay for it.
    """

    def __init__(self, mesh: mesh_2d, models, materials, temperature: torch.Tensor):
        n = mesh.nodes_count
        if temperature.numel() < n:
            raise NLFError(_abi.NLF_ERR_ARGUMENT, "The temperature vector is shorter than the number of nodes.")
        self.mesh = mesh
        self._models = list(models)
        self._materials = list(materials)
        self._T = temperature.reshape(-1)[:n].to(device=mesh.ctx.device, dtype=torch.float64).contiguous()
        self._flux = None
        self._qflux = None

    def temperature(self) -> torch.Tensor:
        return self._T

    def is_flux_calculated(self) -> bool:
        return self._flux is not None

    def flux(self):
        if not self.is_flux_calculated():
            raise RuntimeError("Flux wasn't calculated")
        return self._flux

    def flux_in_qnodes(self):
        """Quadrature-point fluxes (x, y), index e*nq + q, after the nonlocal averaging."""
        if not self.is_flux_calculated():
            raise RuntimeError("Flux wasn't calculated")
        return self._qflux

    def calc_flux(self):
        if self.is_flux_calculated():
            return self.flux()
        mesh = self.mesh
        if any(m.local_weight < 0.999 for m in self._models) and mesh.influences is None:
            _search(mesh, self._models)
        n, E = mesh.nodes_count, mesh.elements_count
        nq = 4 if mesh.mesh.nn == 4 else 9
        out = torch.empty(2 * n, dtype=torch.float64, device=mesh.ctx.device)
        q = torch.empty(2 * E * nq, dtype=torch.float64, device=mesh.ctx.device)
        marr = _abi.array_of(_abi.Model, self._models)
        matarr = _abi.array_of(_abi.ThermalMaterial, self._materials)
        mesh.ctx.check(load().nlf_heat_flux(mesh.h, marr, matarr, _dev(self._T), _dev(out), _dev(q)))
        self._flux = (out[:n], out[n:])
        self._qflux = (q[:E * nq], q[E * nq:])
        return self.flux()


class mechanical_solution_2d:
    """Displacement, strains and stresses of an elastic solve (solvers/solver_2d/mechanical/
    mechanical_solution_2d.hpp:11-232), device resident.

    calc_strain_and_stress() computes on the GPU the strains (u_x,x, u_y,y, (u_x,y + u_y,x)/2) at the
    quadrature points, the stresses C eps and their nonlocal average p1 s + (1 - p1) int phi s for
    nonlocal groups, and the node values as the area-weighted mean of the nearest quadrature points.
    Thermal strains (a temperature change) are not included.
    """

    def __init__(self, mesh: mesh_2d, models, materials, displacement: torch.Tensor, delta_temperature=None):
        n = mesh.nodes_count
        if displacement.numel() < 2 * n:
            raise NLFError(_abi.NLF_ERR_ARGUMENT, "The displacement vector is shorter than twice the number of nodes.")
        self.mesh = mesh
        self._models = list(models)
        self._materials = list(materials)
        self._u = displacement.reshape(-1)[:2 * n].to(device=mesh.ctx.device, dtype=torch.float64).contiguous()
        self._dT = _node_vector(mesh, delta_temperature)
        self._strain = self._stress = self._qstress = None

    def delta_temperature(self):
        """The node temperature change whose thermal strains the stresses exclude (None: no thermal strain)."""
        return self._dT

    def displacement(self):
        """(u_x, u_y) node vectors."""
        return self._u[0::2], self._u[1::2]

    def is_strain_and_stress_calculated(self) -> bool:
        return self._stress is not None

    def strain(self):
        """(xx, yy, xy) node vectors; empty before calc_strain_and_stress, like the reference."""
        return self._strain if self._strain is not None else ()

    def stress(self):
        return self._stress if self._stress is not None else ()

    def stress_in_qnodes(self):
        if self._qstress is None:
            raise RuntimeError("Stress wasn't calculated")
        return self._qstress

    def calc_strain_and_stress(self):
        if self.is_strain_and_stress_calculated():
            return
        mesh = self.mesh
        if any(m.local_weight < 0.999 for m in self._models) and mesh.influences is None:
            _search(mesh, self._models)
        n, E = mesh.nodes_count, mesh.elements_count
        nq = 4 if mesh.mesh.nn == 4 else 9
        dev = mesh.ctx.device
        strain = torch.empty(3 * n, dtype=torch.float64, device=dev)
        stress = torch.empty(3 * n, dtype=torch.float64, device=dev)
        q = torch.empty(3 * E * nq, dtype=torch.float64, device=dev)
        marr = _abi.array_of(_abi.Model, self._models)
        matarr = _abi.array_of(_abi.ElasticMaterial, self._materials)
        mesh.ctx.check(load().nlf_strain_stress(mesh.h, marr, matarr, _dev(self._u),
                                                _dev(self._dT) if self._dT is not None else None, _dev(strain),
                                                _dev(stress), _dev(q)))
        self._strain = tuple(strain[k * n:(k + 1) * n] for k in range(3))
        self._stress = tuple(stress[k * n:(k + 1) * n] for k in range(3))
        self._qstress = tuple(q[k * E * nq:(k + 1) * E * nq] for k in range(3))


def _node_vector(mesh: mesh_2d, v):
    """A node vector (array, tensor or None) as an FP64 device tensor of length N (None stays None)."""
    if v is None:
        return None
    t = v if isinstance(v, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64))
    if t.numel() != mesh.nodes_count:
        raise NLFError(_abi.NLF_ERR_ARGUMENT, "It is impossible to calculate temperature deformations due to "
                                              "different dimensions of the mesh and the delta temperature vector.")
    return t.reshape(-1).to(device=mesh.ctx.device, dtype=torch.float64).contiguous()


def temperature_condition(mesh: mesh_2d, models, materials, delta_temperature, f: torch.Tensor):
    """f (2N, x and y per node) += the thermal load of the node temperature change
    (temperature_condition, mechanical/temperature_condition_2d.hpp:128-170)."""
    dT = _node_vector(mesh, delta_temperature)
    if any(m.local_weight < 0.999 for m in models) and mesh.influences is None:
        _search(mesh, models)
    marr = _abi.array_of(_abi.Model, models)
    matarr = _abi.array_of(_abi.ElasticMaterial, materials)
    mesh.ctx.check(load().nlf_temperature_condition(mesh.h, marr, matarr, _dev(dT), _dev(f)))
    return f


def boundary_data(mesh: mesh_2d, dof, bcs, values=None):
    """Host boundary data (f, x_D, is_inner) of one boundary condition per boundary group; values is
    an optional callable (x, y, group, dof) -> value for the first-kind and second-kind groups, as
    the reference's conditions are functions of the point."""
    rows = dof * mesh.nodes_count
    f = np.zeros(rows)
    xD = np.zeros(rows)
    inner = np.ones(rows, dtype=np.uint8)
    arr = _abi.array_of(_abi.Boundary, bcs)
    fn = _abi.boundary_callback(values)
    mesh.ctx.check(load().nlf_boundary_data_fn(mesh.h, dof, arr, C.cast(fn, C.c_void_p) if fn else None, None,
                                               _ptr(f, _f64p), _ptr(xD, _f64p), _ptr(inner, _u8p)))
    return f, xD, inner


def _distances(models):
    """The distance function of each group's influence (read_influences, read_model.hpp:226-244)."""
    return [_abi.distance_of(m) for m in models]


def _search(mesh: mesh_2d, models):
    """Neighbour search of the models' groups: their search radii and distances."""
    return find_neighbours(mesh, _radii(models), distances=_distances(models))


def _radii(models):
    out = []
    for m in models:
        if m.local_weight >= 0.999:
            out.append([0.0, 0.0])
        else:
            s0 = m.search_radius[0] or m.radius[0]
            s1 = m.search_radius[1] or m.radius[1]
            out.append([s0, s1])
    return out


def _solve_driver(mesh: mesh_2d, physics, models, materials, bcs, precond, tol, max_it, mode, weights=None,
                  assembly32=False, coarse=None, values=None, source=None, delta_temperature=None, energy=0.0,
                  coefficients=None, smoother=None):
    import time
    ctx = mesh.ctx
    dof = 1 if physics == 0 else 2
    tm = {}
    t0 = time.perf_counter()
    if any(m.local_weight < 0.999 for m in models):
        _search(mesh, models)
    tm["neighbours"] = time.perf_counter() - t0
    f, xD, inner = boundary_data(mesh, dof, bcs, values)
    neumann = physics == 0 and is_neumann(bcs)
    if neumann and precond == _abi.NLF_PRECOND_LEARNED:
        raise NLFError(_abi.NLF_ERR_UNSUPPORTED, "the learned preconditioner does not take the bordered Neumann system")
    t1 = time.perf_counter()
    K = _assemble(mesh, physics, models, materials, inner, neumann=neumann, coefficients=coefficients)
    if physics == 0:
        K.add_convection(bcs)
    fd = torch.from_numpy(f).to(ctx.device)
    if source is not None:
        right_part(mesh, dof, source, fd)
    if delta_temperature is not None and physics == 1:
        temperature_condition(mesh, models, materials, delta_temperature, fd)
    if neumann:
        # init_right_part (stationary_heat_equation_solver_2d.hpp:28-44): solvable only when the
        # fluxes balance; the Lagrange row's right-hand side is the energy
        if abs(float(fd.sum())) > NEUMANN_PROBLEM_ERROR_THRESHOLD:
            raise NLFError(_abi.NLF_ERR_ARGUMENT, "It's unsolvable Neumann problem.")
        fd = torch.cat([fd, torch.tensor([float(energy)], dtype=torch.float64, device=ctx.device)])
        xD = np.append(xD, 0.0)
    xDd = torch.from_numpy(xD).to(ctx.device)
    K.apply_first_kind(xDd, fd)
    ctx.synchronize()
    tm["assembly"] = time.perf_counter() - t1
    t2 = time.perf_counter()
    KL = None
    cg = conjugate_gradient(K)
    cg.set_tolerance(tol)
    cg.set_max_interations(max_it)
    learned_pc = None
    if precond == NLF_PRECOND_ILLT:
        KL = _assemble(mesh, physics, models, materials, inner, part=NLF_PART_LOCAL, neumann=neumann,
                       coefficients=coefficients)
        P = cg.init_preconditioner(NLF_PRECOND_ILLT, KL)
        if not P.computation_info():          # fall back to identity like the reference
            cg.init_preconditioner(NLF_PRECOND_IDENTITY)
    elif precond == _abi.NLF_PRECOND_LEARNED:
        # the learned preconditioner is taken first; Jacobi is the fallback of the flexible CG
        from . import learned
        KL = _assemble(mesh, physics, models, materials, inner, part=NLF_PART_LOCAL)
        w, _ = learned.load_weights(weights or learned.DEFAULT_WEIGHTS)
        learned_pc = learned.LearnedPreconditioner(K, KL, models, w)
        if smoother is not None:
            learned_pc.set_apply(smoother)
        learned_pc.set_coarse(learned.DEFAULT_AGGREGATES if coarse is None else coarse)
        fallback = preconditioner(K, NLF_PRECOND_JACOBI)
    else:
        cg.init_preconditioner(precond)
    ctx.synchronize()
    tm["preconditioner"] = time.perf_counter() - t2
    if mode != NLF_MODE_FP64:
        t4 = time.perf_counter()
        if mode == _abi.NLF_MODE_SPLIT:
            K.make_split()
        elif assembly32:
            K.assemble_fp32()
        else:
            K.make_fp32()
        ctx.synchronize()
        tm["fp32_values"] = time.perf_counter() - t4
    t3 = time.perf_counter()
    if learned_pc is not None:
        x, st = learned.fcg_solve(K, learned_pc, fd, fallback=fallback, tol=tol, max_it=max_it, mode=mode)
        cg.stats = st
    else:
        x = cg.solve(fd, mode=mode)
    ctx.synchronize()
    tm["solve"] = time.perf_counter() - t3
    st = cg.stats
    if physics == 0:
        post = heat_equation_solution_2d(mesh, models, materials, x)
    else:
        post = mechanical_solution_2d(mesh, models, materials, x, delta_temperature)
    return Solution(x=x, iterations=st.iterations, residual=st.residual, true_residual=st.true_residual, matrix=K,
                    local_matrix=KL, rhs=fd, timings=tm, stats=st, solution=post)


def stationary_heat_equation_solver_2d(mesh: mesh_2d, models, materials, bcs, precond=NLF_PRECOND_ILLT, tol=1e-12,
                                       max_it=100000, mode=NLF_MODE_FP64, weights=None, assembly32=False,
                                       coarse=None, values=None, source=None, energy=0.0,
                                       coefficients=None, smoother=None) -> Solution:
    """K T = f with K = p1 K_L + (1 - p1) K_NL (solved directly; the reference's correction form
    T = T0 - K^-1 (K T0 - f) with T0 = 0 is the same system).  precond NLF_PRECOND_LEARNED runs
    flexible CG with the learned preconditioner (weights file `weights`, default the packaged one)
    and Jacobi-PCG as its fallback; tol is then on the FP64 true residual.  In the mixed mode the
    inner FP32 operator is the rounded FP64 matrix, or with assembly32 an FP32 assembly.  coarse sets
    the learned preconditioner's aggregation coarse correction (about that many aggregates; 0 = one
    level; default 4096).  values: optional callable (x, y, group, dof) -> boundary value at a point
    for the first-kind and flux groups.  source: optional heat source s(x, y) (arrays in, array out),
    integrated as int s N_i.  A problem whose conditions are all fluxes is a Neumann problem: it is
    solved with the Lagrange row (sum of T weighted by int N_i = energy) and its solution vector has
    N + 1 entries (the temperature, then the multiplier).  coefficients: per element group None or its
    conductivity at its quadrature points (all groups local).  smoother "jacobi" (learned only) keeps
    the coarse part and replaces the network by its fixed Jacobi term, the form that shows what the
    trained network adds."""
    return _solve_driver(mesh, 0, models, materials, bcs, precond, tol, max_it, mode, weights, assembly32, coarse,
                         values, source, energy=energy, coefficients=coefficients, smoother=smoother)


def equilibrium_equation(mesh: mesh_2d, models, materials, bcs, precond=None, tol=1e-12, max_it=100000,
                         mode=NLF_MODE_FP64, assembly32=False, values=None, source=None,
                         delta_temperature=None, coefficients=None) -> Solution:
    """Plane-stress static elasticity (equilibrium_equation, mechanical/equilibrium_equation_2d.hpp:20-62);
    ILLT of the local stiffness when nonlocal, identity otherwise.  materials: nlf_elastic_material per
    group (isotropic, orthotropic or anisotropic Hooke, optional thermal expansion).  values: optional
    callable (x, y, group, dof) -> displacement or pressure of a boundary group at a point.
    delta_temperature: node temperature changes; their thermal load enters the right-hand side and
    their thermal strains are left out of the stresses.  The solution's `solution` is a
    mechanical_solution_2d."""
    if precond is None:
        precond = NLF_PRECOND_ILLT if any(m.local_weight < 0.999 for m in models) else NLF_PRECOND_IDENTITY
    return _solve_driver(mesh, 1, models, materials, bcs, precond, tol, max_it, mode, assembly32=assembly32,
                         values=values, source=source, delta_temperature=delta_temperature,
                         coefficients=coefficients)


class nonstationary_heat_equation_solver_2d:
    """Implicit Euler heat conduction (nonstationary_heat_equation_solver_2d, thermal/
    nonstationary_heat_equation_solver_2d.hpp:17-106):

        (C + tau K) T_{n+1} = C T_n + tau (f_boundary + f_source),

    K the conductivity matrix (with convection), C the heat capacity matrix, both on the device.
    compute() assembles tau K + C once (nlf_matrix_combine); each calc_step() forms the right-hand
    side (second-kind conditions and the source at this step, plus C T_n), applies the first-kind
    conditions and runs CG from T_n.  As in the reference the CG has no preconditioner unless one is
    asked for; precond=NLF_PRECOND_JACOBI takes the diagonal of tau K + C.
    """

    def __init__(self, mesh: mesh_2d, time_step: float):
        self.mesh = mesh
        self._tau = float(time_step)
        self._T = None
        self._K = self._C = self._cg = None

    def time_step(self):
        return self._tau

    def temperature(self) -> torch.Tensor:
        return self._T

    def compute(self, models, materials, capacities, bcs, init_dist, values=None, precond=NLF_PRECOND_IDENTITY,
                tol=1e-15, max_it=10000):
        """init_dist: callable (x, y) on arrays of node coordinates, or an array of N temperatures."""
        mesh = self.mesh
        if any(m.local_weight < 0.999 for m in models):
            _search(mesh, models)
        _, _, inner = boundary_data(mesh, 1, bcs, values)
        self._inner = inner
        K = _assemble(mesh, 0, models, materials, inner)
        K.add_convection(bcs)
        Cm = heat_capacity_matrix_2d(mesh).calc_matrix(capacities, inner)
        self._K = combine(K, self._tau, Cm, 1.0)
        self._C = Cm
        xy = mesh.mesh.xy
        T0 = init_dist(xy[:, 0], xy[:, 1]) if callable(init_dist) else init_dist
        T0 = np.array(np.broadcast_to(np.asarray(T0, dtype=np.float64), (mesh.nodes_count,)))
        self._T = torch.from_numpy(T0).to(mesh.ctx.device)
        self._cg = conjugate_gradient(self._K)
        self._cg.set_tolerance(tol)
        self._cg.set_max_interations(max_it)
        if precond != NLF_PRECOND_IDENTITY:
            self._cg.init_preconditioner(precond)

    def calc_step(self, bcs, values=None, source=None):
        """One step; bcs / values / source may change from step to step (the first-kind node set was
        fixed by compute)."""
        mesh = self.mesh
        f, xD, inner = boundary_data(mesh, 1, bcs, values)
        if not np.array_equal(inner, self._inner):
            raise NLFError(_abi.NLF_ERR_ARGUMENT, "the first-kind boundary groups changed after compute()")
        fd = torch.from_numpy(f).to(mesh.ctx.device)
        if source is not None:
            right_part(mesh, 1, source, fd)
        rhs = fd.mul_(self._tau).add_(self._C.spmv(self._T))
        self._K.apply_first_kind(torch.from_numpy(xD).to(mesh.ctx.device), rhs)
        self._T = self._cg.solve(rhs, self._T)
        return self._T

    def stats(self):
        return self._cg.stats
