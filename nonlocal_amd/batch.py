"""Batched nonlocal heat solves: many uneven instances assembled as one block-diagonal system and
solved by lockstep Jacobi-PCG under a captured hipGraph (nlf_batch_* in libnlf.so).

The instances are concatenated into one mesh: node and element ids are offset, every instance
keeps its own element groups (models, materials) and boundary groups (conditions).  Groups never
couple in the neighbour search (find_neighbours.hpp:129-157), so one neighbour search, one
pattern build and one assembly launch cover the whole batch and produce a block-diagonal matrix.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _abi, api, meshgen
from ._lib import load
from .api import _destroy

_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


def concatenate(problems):
    """One Mesh2D plus per-group models/materials and per-boundary-group conditions for a list of
    thermal problems; returns (mesh, models, materials, bcs, node_begin)."""
    xy, elems, gb, gnames, be, bb, bnames = [], [], [0], [], [], [0], []
    models, mats, bcs = [], [], []
    node_begin = [0]
    e_off = 0
    nn = problems[0].mesh.nn
    for p in problems:
        m = p.mesh
        assert p.physics == 0 and m.nn == nn, "a batch holds thermal problems of one element type"
        off = node_begin[-1]
        xy.append(m.xy)
        elems.append(m.elems + off)
        gb.extend((m.group_begin[1:] + e_off).tolist())
        gnames.extend(m.group_names)
        be.append(m.bedges + off)
        bb.extend((m.bgroup_begin[1:] + bb[-1]).tolist())
        bnames.extend(m.bgroup_names)
        models.extend(p.models)
        mats.extend(p.materials)
        bcs.extend(p.bcs)
        e_off += m.n_elems
        node_begin.append(off + m.n_nodes)
    mesh = meshgen.Mesh2D(xy=np.concatenate(xy), elems=np.concatenate(elems).astype(np.int32), nn=nn,
                          group_begin=np.asarray(gb, dtype=np.int64), group_names=gnames,
                          bedges=np.concatenate(be).astype(np.int32), bgroup_begin=np.asarray(bb, dtype=np.int64),
                          bgroup_names=bnames)
    return mesh, models, mats, bcs, np.asarray(node_begin, dtype=np.int64)


class BatchSolver:
    """Assemble the batch once, then solve every instance together."""

    def __init__(self, problems, ctx: api.Context | None = None, precond=_abi.NLF_PRECOND_JACOBI):
        """precond NLF_PRECOND_JACOBI (of the matrix) or NLF_PRECOND_ILLT (IC(0) of the batch's local
        matrix: block diagonal, one sweep applies every instance's factor)."""
        ctx = ctx or api.default_context()
        self.ctx = ctx
        self.problems = problems
        t0 = time.perf_counter()
        mesh, models, mats, bcs, node_begin = concatenate(problems)
        self.node_begin = node_begin
        self.mesh = api.mesh_2d(mesh, ctx)
        radii = api._radii(models)
        if any(m.local_weight < 0.999 for m in models):
            api.find_neighbours(self.mesh, radii, distances=api._distances(models))
        rows = mesh.n_nodes
        f = np.zeros(rows)
        xD = np.zeros(rows)
        inner = np.ones(rows, dtype=np.uint8)
        ctx.check(load().nlf_boundary_data(self.mesh.h, 1, _abi.array_of(_abi.Boundary, bcs), f.ctypes.data_as(_f64p),
                                           xD.ctypes.data_as(_f64p), inner.ctypes.data_as(_u8p)))
        self.K = api._assemble(self.mesh, 0, models, mats, inner)
        self.K.add_convection(bcs)
        self.b = torch.from_numpy(f).to(ctx.device)
        self.K.apply_first_kind(torch.from_numpy(xD).to(ctx.device), self.b)
        h = C.c_void_p()
        rb = np.ascontiguousarray(node_begin, dtype=np.int64)
        ctx.check(load().nlf_batch_create(self.K.h, len(problems), rb.ctypes.data_as(_i64p), C.byref(h)))
        self.h = h
        self.KL = self.P = None
        if precond == _abi.NLF_PRECOND_ILLT:
            self.KL = api._assemble(self.mesh, 0, models, mats, inner, part=_abi.NLF_PART_LOCAL)
            self.P = api.preconditioner(self.KL, _abi.NLF_PRECOND_ILLT)
            ctx.check(load().nlf_batch_set_preconditioner(self.h, self.P.h))
        ctx.synchronize()
        self.setup_seconds = time.perf_counter() - t0
        self.target = None

    def __del__(self):
        if getattr(self, "h", None):
            _destroy("nlf_batch_destroy", self.h)
            self.h = None

    def autotune(self, cache_path: str | None = None):
        t = C.c_double()
        self.ctx.check(load().nlf_batch_autotune(self.h, cache_path.encode() if cache_path else None, C.byref(t)))
        self.target = t.value
        return t.value

    def solve(self, tol=1e-10, max_it=100000, mode=_abi.NLF_MODE_FP64, use_graph=True, x0=None, assembly32=False):
        """mode NLF_MODE_MIXED reads FP32 values: the rounded FP64 ones, or with assembly32 an FP32
        assembly of the whole batch."""
        if mode != _abi.NLF_MODE_FP64:
            if assembly32:
                self.K.assemble_fp32()
            else:
                self.K.make_fp32()
        x = torch.zeros_like(self.b) if x0 is None else x0.clone()
        count = len(self.problems)
        its = np.zeros(count, dtype=np.int64)
        res = np.zeros(count)
        st = _abi.SolveStats()
        self.ctx.check(load().nlf_batch_solve(self.h, api._dev(self.b), api._dev(x), tol, max_it, mode, int(use_graph),
                                              its.ctypes.data_as(_i64p), res.ctypes.data_as(_f64p), C.byref(st)))
        return x, its, res, st

    def split(self, x):
        """Per-instance views of a batch vector."""
        return [x[self.node_begin[i]:self.node_begin[i + 1]] for i in range(len(self.problems))]
