"""GPU heat capacity matrix, source integrals, matrix combination and implicit Euler time stepping
against the oracle; the reference's source-term test through the GPU solve."""
import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, configs, meshgen
from nonlocal_amd._lib import load
from oracle import oracle as O
from tests import helpers as H
from tests.test_transient_oracle import source_test_problem

pytestmark = pytest.mark.gpu


def _full(csr, n):
    U = sp.csr_matrix((csr.val, csr.inner, csr.outer), shape=(n, n))
    return (U + sp.triu(U, 1).T).tocsr()


@pytest.mark.parametrize("nn", [4, 8])
def test_capacity_matrix_matches_oracle(gpu_ctx, nn):
    p = configs.cfg0(scale=0.4) if nn == 4 else configs.cfg1(scale=0.06)
    print(f"[capacity] {p.name}", flush=True)
    om = O.OracleMesh(p.mesh)
    f, xD, inner = om.boundary(p.bcs)
    ref = O.capacity_matrix(om, [2.5], inner)
    g = api.mesh_2d(p.mesh, gpu_ctx)
    Cm = api.heat_capacity_matrix_2d(g).calc_matrix([2.5], inner)
    for part in (0, 1):
        a, b = Cm.export(part), ref.export(part)
        assert np.array_equal(a.outer, b.outer) and np.array_equal(a.inner, b.inner)
        assert np.abs(a.val - b.val).max() <= 1e-14 * max(np.abs(b.val).max(), 1e-300)


@pytest.mark.parametrize("nn,dof", [(4, 1), (8, 1), (4, 2)])
def test_right_part_matches_oracle(gpu_ctx, nn, dof):
    mesh = meshgen.structured_q4(11, jitter=0.2, seed=3) if nn == 4 else meshgen.structured_q8(7, jitter=0.2, seed=3)
    if dof == 1:
        def src(x, y):
            return np.sin(4 * x) * np.exp(y)
    else:
        def src(x, y):
            return np.sin(4 * x) * np.exp(y), x * y - 0.5
    ref = O.OracleMesh(mesh).right_part(dof, lambda x, y: src(np.float64(x), np.float64(y)))
    g = api.mesh_2d(mesh, gpu_ctx)
    f = torch.zeros(dof * mesh.n_nodes, dtype=torch.float64, device="cuda")
    api.right_part(g, dof, src, f)
    assert np.abs(f.cpu().numpy() - ref).max() <= 1e-15 * np.abs(ref).max() * 4


def test_combine_matches_host_algebra(gpu_ctx):
    p = configs.cfg0(scale=0.5)
    g = api.mesh_2d(p.mesh, gpu_ctx)
    api.find_neighbours(g, H.radii(p.models))
    _, _, inner = api.boundary_data(g, 1, p.bcs)
    K = api._assemble(g, 0, p.models, p.materials, inner)
    Cm = api.heat_capacity_matrix_2d(g).calc_matrix([1.5], inner)
    k0, k1, c0 = K.export(2), K.export(1), Cm.export(2)
    n = p.mesh.n_nodes
    expect = (0.25 * sp.csr_matrix((k0.val, k0.inner, k0.outer), shape=(n, n)) +
              sp.csr_matrix((c0.val, c0.inner, c0.outer), shape=(n, n))).tolil()
    for r in np.flatnonzero(inner == 0):
        expect[r, r] = 1.0
    expect = expect.tocsr()
    api.combine(K, 0.25, Cm, 1.0)
    got = K.export(2)
    assert np.array_equal(got.outer, k0.outer) and np.array_equal(got.inner, k0.inner)
    dense_got = sp.csr_matrix((got.val, got.inner, got.outer), shape=(n, n))
    assert abs(dense_got - expect).max() <= 1e-15 * abs(expect).max()
    assert np.array_equal(K.export(1).val, 0.25 * k1.val)
    assert load().nlf_matrix_assemble_fp32(K.h) == A.NLF_ERR_STATE


@pytest.mark.parametrize("model", [A.model(1.0, 0.0), configs.bell(0.5, 0.15)], ids=["local", "nonlocal"])
def test_transient_matches_oracle(gpu_ctx, model):
    """Robin top, unit flux on the right, T = 0 on the left, a source, T0 = 1 - y: six steps."""
    print(f"[transient] model p1={model.local_weight}", flush=True)
    p = configs.cfg1(scale=0.06)                          # 12 x 12 Q8, jittered, anisotropic lambda, Robin
    p.models = [model]
    tau, steps = 0.02, 6

    def src(x, y):
        return 3.0 * np.cos(2 * x) * y

    T0 = 1.0 - p.mesh.xy[:, 1]
    ref = O.transient(p.mesh, p.models, p.materials, [2.0], p.bcs, tau, T0, steps,
                      source=lambda x, y: src(np.float64(x), np.float64(y)))
    g = api.mesh_2d(p.mesh, gpu_ctx)
    solver = api.nonstationary_heat_equation_solver_2d(g, tau)
    solver.compute(p.models, p.materials, [2.0], p.bcs, lambda x, y: 1.0 - y, precond=A.NLF_PRECOND_JACOBI,
                   tol=1e-15, max_it=20000)
    for k in range(steps):
        T = solver.calc_step(p.bcs, source=src).cpu().numpy()
        err = H.rel_l2(T, ref[k])
        assert err <= 1e-10, (k, err)
    print(f"  step {steps}: relative error {err:.2e}, CG iterations of the last step {solver.stats().iterations}",
          flush=True)


def test_transient_eigenmode_decay(gpu_ctx):
    """K v = lambda C v: one step divides v by 1 + tau lambda."""
    mesh = meshgen.structured_q4(10, jitter=0.15, seed=6)
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0), A.boundary(A.NLF_BC_FIRST_KIND, 0.0), A.boundary(), A.boundary()]
    models = [configs.bell(0.5, 0.25)]
    materials = [A.thermal_material(A.NLF_CONDUCTIVITY_ISOTROPIC, (1.3, 0, 0))]
    om = O.OracleMesh(mesh)
    om.find_neighbours(H.radii(models))
    f, xD, inner = om.boundary(bcs)
    n = mesh.n_nodes
    K = _full(om.assemble(0, models, materials, inner).export(0), n).toarray()
    Cm = _full(O.capacity_matrix(om, [2.0], inner).export(0), n).toarray()
    idx = np.flatnonzero(inner)
    lam, vec = la.eigh(K[np.ix_(idx, idx)], Cm[np.ix_(idx, idx)])
    v = np.zeros(n)
    v[idx] = vec[:, 3]
    tau, steps = 0.01, 5
    g = api.mesh_2d(mesh, gpu_ctx)
    solver = api.nonstationary_heat_equation_solver_2d(g, tau)
    solver.compute(models, materials, [2.0], bcs, v, tol=1e-15, max_it=20000)
    for _ in range(steps):
        T = solver.calc_step(bcs)
    expect = v / (1 + tau * lam[3]) ** steps
    assert np.abs(T.cpu().numpy() - expect).max() <= 1e-11 * np.abs(v).max()


def test_source_term_known_answer(gpu_ctx):
    mesh, models, materials, bcs, values, source, ref = source_test_problem()
    g = api.mesh_2d(mesh, gpu_ctx)
    sol = api.stationary_heat_equation_solver_2d(g, models, materials, bcs, tol=1e-14, values=values, source=source)
    exact = ref(mesh.xy[:, 0], mesh.xy[:, 1])
    err = float(np.sum((sol.x.cpu().numpy() - exact) ** 2) / np.sum(exact ** 2))
    print(f"[source term] relative squared error {err:.3e} (< 1e-3)", flush=True)
    assert err < 1e-3
