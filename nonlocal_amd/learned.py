"""Learned preconditioner: weights file, training data from FP64 solves, GPU training, flexible CG.

The network and its forward/backward/Adam kernels live in libnlf.so (nlf_learned_*, nlf_trainer_*,
nlf_fcg_solve); this module orchestrates them.  Weights are stored in a small versioned binary file:

    b"NLFW" | u32 version | u32 n_params | u64 seed | 32-byte SHA-256 of the training config (JSON) |
    n_params float32 (little endian)
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import struct
import time

import numpy as np
import torch

from . import _abi
from . import api, configs
from ._lib import load
from .api import _destroy

MAGIC = b"NLFW"
VERSION = 1
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "mpnn_thermal.bin")

_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)


def num_params() -> int:
    return load().nlf_learned_num_params()


def init_weights(seed: int) -> np.ndarray:
    w = np.zeros(num_params(), dtype=np.float32)
    load().nlf_learned_init_weights(seed, w.ctypes.data_as(_f32p))
    return w


def save_weights(path, w: np.ndarray, seed: int, config: dict):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    h = hashlib.sha256(json.dumps(config, sort_keys=True).encode()).digest()
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<IIQ", VERSION, w.size, seed) + h)
        f.write(np.ascontiguousarray(w, dtype="<f4").tobytes())


def load_weights(path=DEFAULT_WEIGHTS):
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != MAGIC:
        raise ValueError(f"{path}: not a weights file")
    version, n, seed = struct.unpack("<IIQ", data[4:20])
    if version != VERSION:
        raise ValueError(f"{path}: unsupported version {version}")
    if n != num_params():
        raise ValueError(f"{path}: {n} parameters, the network has {num_params()}")
    w = np.frombuffer(data[52:52 + 4 * n], dtype="<f4").astype(np.float32)
    return w, dict(seed=seed, config_sha256=data[20:52].hex())


DEFAULT_AGGREGATES = 4096


class LearnedPreconditioner:
    """preconditioner_base implementation backed by the message-passing network."""

    def __init__(self, matrix: api.finite_element_matrix, local_matrix: api.finite_element_matrix, models,
                 weights: np.ndarray | None = None):
        self.ctx = matrix.ctx
        self.matrix = matrix
        h = C.c_void_p()
        self.ctx.check(load().nlf_learned_create(matrix.h, local_matrix.h, _abi.array_of(_abi.Model, models),
                                                 C.byref(h)))
        self.h = h
        self.kind = _abi.NLF_PRECOND_LEARNED
        if weights is not None:
            self.set_weights(weights)

    def __del__(self):
        if getattr(self, "h", None):
            _destroy("nlf_precond_destroy", self.h)
            self.h = None

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self.ctx.check(load().nlf_learned_set_weights(self.h, w.ctypes.data_as(_f32p)))

    def info(self):
        """(compiled to an explicit sparse operator, its nonzeros, compile seconds)."""
        c, nnz, sec = C.c_int(), C.c_int64(), C.c_double()
        self.ctx.check(load().nlf_learned_info(self.h, C.byref(c), C.byref(nnz), C.byref(sec)))
        return bool(c.value), nnz.value, sec.value

    def set_apply(self, network):
        """Apply through the network kernels (True), the compiled sparse operator (False, default), or
        "jacobi": the network's fixed term diag^-1 r alone (learned correction off, coarse part kept)."""
        mode = 2 if network == "jacobi" else int(bool(network))
        self.ctx.check(load().nlf_learned_set_apply(self.h, mode))

    def set_coarse(self, aggregates: int = DEFAULT_AGGREGATES, matrix: api.finite_element_matrix | None = None):
        """Two-level form B r = M r + P A_c^-1 P^T r with about `aggregates` aggregates of the system
        matrix (default: the one the preconditioner was built for); 0 removes the coarse part."""
        m = matrix if matrix is not None else self.matrix
        self.ctx.check(load().nlf_learned_set_coarse(self.h, m.h, int(aggregates)))

    def coarse_info(self):
        """(coarse rows, set-up seconds); (0, 0.0) without the coarse part."""
        rows, sec = C.c_int64(), C.c_double()
        self.ctx.check(load().nlf_learned_coarse_info(self.h, C.byref(rows), C.byref(sec)))
        return rows.value, sec.value

    def get_weights(self):
        w = np.zeros(num_params(), dtype=np.float32)
        self.ctx.check(load().nlf_learned_get_weights(self.h, w.ctypes.data_as(_f32p)))
        return w

    def graph(self):
        info = self.matrix.info()
        n = info["rows"]
        outer = np.zeros(n + 1, dtype=np.int64)
        self.ctx.check(load().nlf_learned_export_graph(self.h, outer.ctypes.data_as(C.POINTER(C.c_int64)), None, None,
                                                       None))
        nnz = int(outer[-1])
        cols = np.zeros(nnz, dtype=np.int32)
        feat = np.zeros((nnz, 4), dtype=np.float32)
        s = np.zeros(n, dtype=np.float32)
        self.ctx.check(load().nlf_learned_export_graph(self.h, outer.ctypes.data_as(C.POINTER(C.c_int64)),
                                                       cols.ctypes.data_as(C.POINTER(C.c_int32)),
                                                       feat.ctypes.data_as(_f32p), s.ctypes.data_as(_f32p)))
        return outer, cols, feat, s

    def solve(self, r: torch.Tensor) -> torch.Tensor:
        z = torch.empty_like(r)
        self.ctx.check(load().nlf_precond_apply(self.h, api._dev(r), api._dev(z)))
        return z

    def loss_grad(self, r: torch.Tensor, e: torch.Tensor):
        loss = C.c_double()
        g = np.zeros(num_params(), dtype=np.float32)
        self.ctx.check(load().nlf_learned_loss_grad(self.h, self.matrix.h, api._dev(r), api._dev(e), C.byref(loss),
                                                    g.ctypes.data_as(_f32p)))
        return loss.value, g


class Trainer:
    def __init__(self, ctx: api.Context, w0: np.ndarray):
        self.ctx = ctx
        h = C.c_void_p()
        w0 = np.ascontiguousarray(w0, dtype=np.float32)
        ctx.check(load().nlf_trainer_create(ctx.h, w0.ctypes.data_as(_f32p), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            _destroy("nlf_trainer_destroy", self.h)
            self.h = None

    def step(self, pre: LearnedPreconditioner, r: torch.Tensor, e: torch.Tensor, lr: float) -> float:
        loss = C.c_double()
        self.ctx.check(load().nlf_trainer_step(self.h, pre.h, pre.matrix.h, api._dev(r), api._dev(e), C.c_float(lr),
                                               C.byref(loss)))
        return loss.value

    def weights(self):
        w = np.zeros(num_params(), dtype=np.float32)
        self.ctx.check(load().nlf_trainer_get_weights(self.h, w.ctypes.data_as(_f32p)))
        return w


def fcg_solve(matrix: api.finite_element_matrix, learned: LearnedPreconditioner, b: torch.Tensor,
              fallback: api.preconditioner | None = None, tol=1e-10, max_it=20000, mode=_abi.NLF_MODE_FP64,
              x0: torch.Tensor | None = None):
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    st = _abi.SolveStats()
    matrix.ctx.check(load().nlf_fcg_solve(matrix.h, learned.h, fallback.h if fallback is not None else None,
                                          api._dev(b), api._dev(x), tol, max_it, mode, C.byref(st)))
    return x, st


# ---------------------------------------------------------------------------------------------
# training data: (residual, error) pairs anchored on FP64 solves
# ---------------------------------------------------------------------------------------------
class TrainingProblem:
    """A seeded composite problem on the GPU: K, K_L, its graph and FP64 (r, e) pairs."""

    def __init__(self, n: int, seed: int, ctx: api.Context, count: int = 16, r: float | None = None):
        self.problem = configs.composite(n, r=r, seed=seed, count=count)
        p = self.problem
        self.mesh = api.mesh_2d(p.mesh, ctx)
        api.find_neighbours(self.mesh, [[m.radius[0], m.radius[1]] for m in p.models], distances=api._distances(p.models))
        rows = p.mesh.n_nodes
        f = np.zeros(rows)
        xD = np.zeros(rows)
        inner = np.ones(rows, dtype=np.uint8)
        ctx.check(load().nlf_boundary_data(self.mesh.h, 1, _abi.array_of(_abi.Boundary, p.bcs),
                                           f.ctypes.data_as(_f64p), xD.ctypes.data_as(_f64p),
                                           inner.ctypes.data_as(C.POINTER(C.c_uint8))))
        self.K = api._assemble(self.mesh, 0, p.models, p.materials, inner)
        self.KL = api._assemble(self.mesh, 0, p.models, p.materials, inner, part=_abi.NLF_PART_LOCAL)
        self.b = torch.from_numpy(f).to(ctx.device)
        self.K.apply_first_kind(torch.from_numpy(xD).to(ctx.device), self.b)
        self.inner = torch.from_numpy(inner.astype(bool)).to(ctx.device)
        self.learned = LearnedPreconditioner(self.K, self.KL, p.models)
        self.ctx = ctx
        self.pairs = []

    def _solve(self, b, max_it, tol=1e-13):
        cg = api.conjugate_gradient(self.K)
        cg.init_preconditioner(_abi.NLF_PRECOND_JACOBI)
        cg.set_tolerance(tol)
        cg.set_max_interations(max_it)
        return cg.solve(b)

    def make_pairs(self, rng: np.random.Generator, count: int = 8):
        """(residual, error) pairs anchored on this problem's FP64 solution x*: start points
        x0 = x* - xi with white-noise xi (zero on Dirichlet rows) give r = b - K x0 and e = x* - x0 = xi.
        White errors weight every part of the spectrum alike, which is what the preconditioner must
        cover; the residual is evaluated in FP64 with the assembled K."""
        n = self.b.numel()
        xs = self._solve(self.b, 100000)
        for _ in range(count):
            xi = torch.from_numpy(rng.standard_normal(n)).to(self.ctx.device) * self.inner
            x0 = xs - xi
            r = self.b - self.K.spmv(x0.contiguous())
            self.pairs.append((r.contiguous(), (xs - x0).contiguous()))
        return self.pairs


def train(problems, steps=1500, lr=2e-3, seed=0, log=print):
    """Adam over all (r, e) pairs of `problems`; returns the trained weights and the loss history."""
    ctx = problems[0].ctx
    w0 = init_weights(seed)
    tr = Trainer(ctx, w0)
    rng = np.random.default_rng(seed)
    items = [(p, k) for p in problems for k in range(len(p.pairs))]
    hist = []
    t0 = time.perf_counter()
    for step in range(steps):
        p, k = items[rng.integers(len(items))]
        r, e = p.pairs[k]
        lr_t = lr * (0.5 * (1 + np.cos(np.pi * step / steps)) * 0.9 + 0.1)
        loss = tr.step(p.learned, r, e, lr_t)
        hist.append(loss)
        if log and (step % max(1, steps // 10) == 0 or step == steps - 1):
            log(f"[train] step {step} loss {np.mean(hist[-50:]):.4f} ({time.perf_counter() - t0:.1f} s)")
    return tr.weights(), hist
