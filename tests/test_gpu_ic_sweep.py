"""ILLT sweeps on the GPU against the oracle's IC(0) apply (eigen_ILLT_preconditioner,
slae/eigen_preconditioner.hpp:56-57, restated in the oracle), at the shapes that select each sweep
layout, with nlf_precond_sweep_info naming the kernel that ran:

- the 1000 x 1000 composite of the benchmark: its level 0 (the rows with no dependency, the
  first-kind column among them) is twice as wide as any other level and is split, so both sweeps
  take the prefetching kernel at a width of 512;
- a square with first-kind conditions on all four sides, whose level 0 is wider than the largest
  workgroup (split into several levels);
- Q8 and plane-stress rows (up to 16 entries below the diagonal), which take shallower prefetch.

The arithmetic of every record is the same in every kernel, so z is held to 1e-12 relative to the
oracle as for the other layouts."""
import numpy as np
import pytest
import torch

from nonlocal_amd import _abi as A
from nonlocal_amd import api, configs, meshgen
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _local_problem_all_first_kind(n):
    mesh = meshgen.structured_q4(n)
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0) for _ in range(4)]
    return configs.Problem(f"q4_{n}_first_kind_all_sides", mesh, 0, [A.model(1.0, 0.0)], [A.thermal_material()], bcs,
                           dict(n=n))


# (name, problem, compute width of both sweeps or None, whether the widest level is split)
CASES = [
    ("composite_1000", lambda: configs.cfg4(1000), 512, True),
    ("first_kind_all_sides_300", lambda: _local_problem_all_first_kind(300), 512, True),
    ("q8_anisotropic_40", lambda: configs.cfg1(0.2), None, False),
    ("elastic_q4_30", lambda: configs.cfg2(0.1), None, False),
]


@pytest.mark.parametrize("name,build,width,split", CASES, ids=[c[0] for c in CASES])
def test_ic_sweep_matches_oracle(gpu_ctx, name, build, width, split):
    problem = build()
    print(f"[ic sweep] {name}: {problem.mesh.n_nodes} nodes", flush=True)
    om, KL, f, xD, inner = H.oracle_assembly(problem, part=A.NLF_PART_LOCAL)
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    KLg = api._assemble(g, problem.physics, problem.models, problem.materials, inner, part=A.NLF_PART_LOCAL)
    P = O.OraclePrecond(A.NLF_PRECOND_ILLT, KL)
    Pg = api.preconditioner(KLg, A.NLF_PRECOND_ILLT)
    assert Pg.info()["ok"] and P.ok()[0]
    info = Pg.sweep_info()
    print(f"  sweeps {info}", flush=True)
    for sweep in ("forward", "backward"):
        assert info[sweep]["kernel"] == 2, f"{sweep} sweep did not take the prefetching kernel"
        if width is not None:
            assert info[sweep]["width"] == width
    if split:
        assert info["forward"]["levels"] > Pg.info()["fw_levels"]
    r = np.random.default_rng(7).standard_normal(f.size)
    z_ref = P.apply(r)
    z = Pg.solve(torch.from_numpy(r).to(gpu_ctx.device)).cpu().numpy()
    err = H.rel_l2(z, z_ref)
    print(f"  z rel err {err:.2e}", flush=True)
    assert err <= 1e-12
    # a second apply gives the same bits (the sweep is deterministic)
    z2 = Pg.solve(torch.from_numpy(r).to(gpu_ctx.device)).cpu().numpy()
    assert np.array_equal(z, z2)


def test_ic_of_independent_blocks(gpu_ctx):
    """A block-diagonal matrix (four thermal instances of 1089 to 1681 nodes concatenated as in a
    batch) is factored and swept block by block, one workgroup per block: its IC(0) is the blocks'
    own, so z is every instance's oracle apply in turn."""
    from nonlocal_amd import batch
    problems = [configs.thermal_q4(n, 0.6, 2.0 / n, 1.0 + 0.1 * k, name=f"i{k}") for k, n in enumerate([35, 32, 40, 33])]
    mesh, models, mats, bcs, node_begin = batch.concatenate(problems)
    print(f"[ic sweep] independent blocks: {mesh.n_nodes} nodes in {len(problems)} instances", flush=True)
    z_ref, r_all = [], []
    rng = np.random.default_rng(11)
    for p in problems:
        om, KL, f, xD, inner = H.oracle_assembly(p, part=A.NLF_PART_LOCAL)
        P = O.OraclePrecond(A.NLF_PRECOND_ILLT, KL)
        assert P.ok()[0] and P.ok()[1] == 0.0
        r = rng.standard_normal(f.size)
        r_all.append(r)
        z_ref.append(P.apply(r))
    z_ref, r_all = np.concatenate(z_ref), np.concatenate(r_all)
    g = api.mesh_2d(mesh, gpu_ctx)
    inner = np.concatenate([H.oracle_assembly(p, part=A.NLF_PART_LOCAL)[4] for p in problems])
    KLg = api._assemble(g, 0, models, mats, inner, part=A.NLF_PART_LOCAL)
    Pg = api.preconditioner(KLg, A.NLF_PRECOND_ILLT)
    assert Pg.info()["ok"] and Pg.info()["shift"] == 0.0
    assert Pg.blocks() == len(problems)
    info = Pg.sweep_info()
    print(f"  blocks {Pg.blocks()} sweeps {info}", flush=True)
    assert info["forward"]["kernel"] == 2 and info["backward"]["kernel"] == 2
    z = Pg.solve(torch.from_numpy(r_all).to(gpu_ctx.device)).cpu().numpy()
    err = H.rel_l2(z, z_ref)
    print(f"  z rel err {err:.2e}", flush=True)
    assert err <= 1e-12
    for i in range(len(problems)):
        a, b = node_begin[i], node_begin[i + 1]
        assert H.rel_l2(z[a:b], z_ref[a:b]) <= 1e-12, f"instance {i}"
    assert np.array_equal(z, Pg.solve(torch.from_numpy(r_all).to(gpu_ctx.device)).cpu().numpy())


def test_connected_matrix_is_one_block(gpu_ctx):
    problem = configs.thermal_q4(40, 0.6, 0.05)
    om, KL, f, xD, inner = H.oracle_assembly(problem, part=A.NLF_PART_LOCAL)
    g = api.mesh_2d(problem.mesh, gpu_ctx)
    KLg = api._assemble(g, 0, problem.models, problem.materials, inner, part=A.NLF_PART_LOCAL)
    assert api.preconditioner(KLg, A.NLF_PRECOND_ILLT).blocks() == 1
