"""The symmetric SpMV (the upper triangle read once, in 64-row slices with wave-private LDS windows
for the transposed sums; independent_symmetric_matrix_vector_product.hpp:55-79) against the oracle
and against the full-storage product, through the C ABI.

* y = A x within 1e-13 relative of the oracle's upper-CSR product and within 1e-14 of the full-storage
  kernel, on thermal Q4, elastic Q4 and a multi-group composite; bitwise reproducible; matrices
  where the slice layout would pad too much (Q8, a small rotated-ellipse mesh, a Neumann-bordered
  system) keep the full storage;
* the storage follows the values: after nlf_matrix_add_convection changes them, the next product is
  the new matrix's;
* CG and flexible CG on it reach the oracle's direct solution (1e-10), with the symmetric storage in use.
"""
import dataclasses

import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, configs, meshgen
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _problems():
    out = [configs.cfg0()]
    p = configs.cfg1(scale=0.1)
    p.models = [configs.bell(0.3, 0.15)]
    out.append(p)
    q = configs.cfg2(scale=0.1)
    q.models = [configs.bell(0.3, 0.1)]
    out.append(q)
    out.append(configs.composite(48, r=0.08, count=4))
    mesh = meshgen.structured_q4(24, jitter=0.2, seed=7)
    mesh = dataclasses.replace(mesh, xy=mesh.xy - 0.5)
    mat = A.thermal_material(A.NLF_CONDUCTIVITY_ANISOTROPIC, (1.0, 0.25, 0.1))
    out.append(configs.Problem("q4_rotation", mesh, 0,
                               [A.model(0.5, (0.15, 0.08), distance=A.NLF_DISTANCE_ELLIPSE_WITH_ROTATION, p=2.0, q=2.0)],
                               [mat], configs.thermal_bcs(robin_top=0.5), {}))
    return out


PROBLEMS = _problems()
IDS = [p.name for p in PROBLEMS]
# where the slice layout saves traffic (the others -- Q8, whose corner and midside rows have different
# offset patterns, and a small rotated-ellipse mesh -- pad too much and keep the full storage)
APPLIES = {"cfg0_thermal_q4_40", "cfg2_elastic_q4_30", "composite_q4_48"}


def _assemble(problem, ctx):
    om, K, f, xD, inner = H.oracle_assembly(problem)
    g = api.mesh_2d(problem.mesh, ctx)
    api._search(g, problem.models)
    Kg = api._assemble(g, problem.physics, problem.models, problem.materials, inner)
    if problem.physics == 0:
        Kg.add_convection(problem.bcs)
    return K, f, g, Kg


def _spmv(Kg, x, ctx, symmetric):
    ctx.set_option(A.NLF_OPT_SPMV_SYMMETRIC, int(symmetric))
    try:
        return Kg.spmv(torch.from_numpy(x).cuda()).cpu().numpy()
    finally:
        ctx.set_option(A.NLF_OPT_SPMV_SYMMETRIC, 1)


@pytest.mark.parametrize("problem", PROBLEMS, ids=IDS)
def test_symmetric_spmv(gpu_ctx, problem):
    print(f"[symv] {problem.name}", flush=True)
    K, f, g, Kg = _assemble(problem, gpu_ctx)
    info = Kg.symv_info()
    assert (info is not None) == (problem.name in APPLIES)
    if info is not None:
        nnz_up = Kg.info()["nnz_upper"]
        print(f"  slots {info[0]} (x64 = {64 * info[0]}, upper nnz {nnz_up}), strips {info[1]}", flush=True)
        assert 64 * info[0] >= nnz_up
    x = np.random.default_rng(3).standard_normal(f.size)
    y = _spmv(Kg, x, gpu_ctx, True)
    y_full = _spmv(Kg, x, gpu_ctx, False)
    y_ref = K.spmv(x)
    e_ref, e_full = H.rel_l2(y, y_ref), H.rel_l2(y, y_full)
    print(f"  vs oracle {e_ref:.2e}, vs full storage {e_full:.2e}", flush=True)
    assert e_ref <= 1e-13 and e_full <= 1e-14
    assert np.array_equal(y, _spmv(Kg, x, gpu_ctx, True)), "the symmetric product must be bitwise reproducible"


def test_storage_follows_the_values(gpu_ctx):
    problem = configs.thermal_q4(30, r=0.12)
    problem.bcs = configs.thermal_bcs(robin_top=0.5)
    om, K, f, xD, inner = H.oracle_assembly(problem)       # with the Robin term
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    api._search(g, problem.models)
    Kg = api._assemble(g, 0, problem.models, problem.materials, inner)
    x = np.random.default_rng(4).standard_normal(f.size)
    before = _spmv(Kg, x, gpu_ctx, True)                   # builds the storage without the Robin term
    assert Kg.symv_info() is not None
    Kg.add_convection(problem.bcs)
    after = _spmv(Kg, x, gpu_ctx, True)
    assert not np.array_equal(before, after)
    assert H.rel_l2(after, K.spmv(x)) <= 1e-13


def test_neumann_bordered(gpu_ctx):
    mesh = meshgen.structured_q4(20, jitter=0.2, seed=3)
    models = [configs.bell(0.5, 0.12)]
    mats = [A.thermal_material()]
    inner = np.ones(mesh.n_nodes, dtype=np.uint8)
    g = api.mesh_2d(mesh, gpu_ctx)
    api._search(g, models)
    Kg = api._assemble(g, 0, models, mats, inner, neumann=True)
    # the border column adds a band of 64 offsets to every slice for one entry per row: the layout would
    # pad it 64-fold, so the bordered system keeps the full storage
    assert Kg.symv_info() is None
    x = np.random.default_rng(5).standard_normal(mesh.n_nodes + 1)
    assert H.rel_l2(_spmv(Kg, x, gpu_ctx, True), _spmv(Kg, x, gpu_ctx, False)) <= 1e-14


@pytest.mark.parametrize("problem", [PROBLEMS[0], PROBLEMS[2], PROBLEMS[3]], ids=[IDS[0], IDS[2], IDS[3]])
def test_cg_on_symmetric_storage(gpu_ctx, problem):
    print(f"[symv cg] {problem.name}", flush=True)
    om, K, f, xD, inner = H.oracle_assembly(problem)
    x_ref = H.direct_solution(K, f)
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    fn = api.stationary_heat_equation_solver_2d if problem.physics == 0 else api.equilibrium_equation
    sol = fn(g, problem.models, problem.materials, problem.bcs, precond=A.NLF_PRECOND_ILLT, tol=1e-14,
             mode=A.NLF_MODE_FP64)
    err = H.rel_l2(sol.x.cpu().numpy(), x_ref)
    print(f"  iterations {sol.iterations} rel err {err:.2e}", flush=True)
    assert err <= 1e-10
