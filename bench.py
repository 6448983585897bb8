"""Benchmark of the 2D nonlocal hot path on one MI355X.

Workload: composite 2D stationary heat conduction on 1000 x 1000 bilinear quads (~1M nodes), 16
circular inclusions (lambda = 10) in a lambda = 1 matrix, p1 = 0.5, bell influence r = 0.01, T = 0 on
x = 0 and unit inward flux on x = 1.  A "step" is the whole hot path on that input, already
resident in HBM: neighbour search, CSR pattern, FP64 element-pair assembly, boundary conditions,
preconditioner set-up and the solve to <= 1e-10 FP64 true residual.

value = assembly+solve seconds of the fastest accurate GPU method (normally flexible CG with the
learned preconditioner: FP32 network, FP64 Krylov vectors, SpMV and true residual); the plain FP64
ILLT-CG time and the other methods are reported beside it.  roofline = the dominant kernel (CSR SpMV inside the
solve) against 8 TB/s HBM, timed live with HIP events on the solver stream.  cpu_baseline = the
oracle (the reference algorithm restated in C++/OpenMP) timed on this host on the full workload
(neighbours, assembly and preconditioner set-up whole; the ILLT-CG for a bounded number of
iterations, scaled by the iteration count); the same oracle system checks the GPU result.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "spmv_composite1000_hbm_counters.txt")
TRAFFIC_PROFILE_SYM = os.path.join(ROOT, "profiles", "symv_composite1000_hbm_counters.txt")


def counter_traffic(n, mode_fp64, symmetric=False):
    """HBM bytes per SpMV launch from the committed FETCH_SIZE / WRITE_SIZE summary of the FP64
    SpMV on the 1000x1000 composite (the symmetric slice kernels or the full-storage kernel; None for
    other workloads or value types)."""
    import re
    path = TRAFFIC_PROFILE_SYM if symmetric else TRAFFIC_PROFILE
    if n != 1000 or not mode_fp64 or not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"SpMV traffic per launch:.*=\s*([0-9.]+e[+0-9]+)\s*B", f.read())
    return float(m.group(1)) if m else None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1000, help="mesh side (1000 = the benchmark workload)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-cg-iterations", type=int, default=40,
                    help="ILLT-CG iterations the oracle runs on the full workload for its per-iteration time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-precond", choices=("jacobi", "illt"), default="illt",
                    help="preconditioner of the batched solves (batch64)")
    ap.add_argument("--no-batch", action="store_true",
                    help="leave out the batch64 figures (64 uneven instances) of the default line")
    ap.add_argument("--methods", default="fp64_jacobi,mixed_jacobi,learned_fcg_onelevel,coarse_jacobi,learned_fcg")
    ap.add_argument("--workload", default="composite1000",
                    choices=["composite1000", "batch64", "precision", "spmv", "flux", "assembly"],
                    help="composite1000: the 1000x1000 composite (the metric's workload); batch64: 64 uneven "
                         "instances solved together vs one after another; precision: FP32 assembly and "
                         "mixed-precision errors against the FP64 path on every workload; spmv: the FP64 "
                         "SpMV of the composite alone, plus a device copy of known size (HBM counter "
                         "calibration); flux: the heat-flux post-processing (calc_flux) of the composite's "
                         "solution; assembly: the composite's pattern and FP64 assembly of K alone, plus the "
                         "same copy (counters of the value pass)")
    return ap.parse_args()


def run_step(api, A, mesh, problem, precond, mode, tol, coarse=None, smoother=None):
    """One pass of the hot path; returns (solution, timings)."""
    torch = api.torch
    t0 = time.perf_counter()
    sol = api._solve_driver(mesh, problem.physics, problem.models, problem.materials, problem.bcs, precond, tol,
                            30000, mode, coarse=coarse, smoother=smoother)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return sol, wall


def host_description():
    """The host the CPU leg runs on: logical CPUs, the ones this process may use, the CPU model, the
    NUMA nodes with their CPU lists, and OMP_NUM_THREADS."""
    import glob
    d = {"nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)),
         "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        with open("/proc/cpuinfo") as f:
            d["cpu_model"] = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        d["cpu_model"] = None
    numa = {}
    for node in sorted(glob.glob("/sys/devices/system/node/node[0-9]*")):
        try:
            with open(os.path.join(node, "cpulist")) as f:
                numa[os.path.basename(node)] = f.read().strip()
        except OSError:
            pass
    d["numa_nodes"] = numa
    return d


def cpu_baseline(problem, plain_iterations, check, cg_iterations):
    """The oracle (the reference algorithm restated in C++/OpenMP, with its O(N^2) flag resets) on
    the full workload: neighbour search, assembly of K, boundary conditions, assembly of the local
    matrix and its IC(0) factor are timed whole; the ILLT-CG is timed for `cg_iterations`
    iterations and its per-iteration time is multiplied by the iteration count the same IC(0)-CG
    needs on this system (the GPU's plain FP64 ILLT-CG count).  The same oracle system then checks
    the GPU result: the residual of the GPU solution on the oracle's K and f, the oracle's K v
    against the GPU's for one random v, and sampled rows of the two upper CSRs."""
    from nonlocal_amd import _abi as A
    from oracle import oracle as O
    cores = O.set_threads(0)
    radii = [[m.radius[0], m.radius[1]] for m in problem.models]
    t0 = time.perf_counter()
    om = O.OracleMesh(problem.mesh)
    om.find_neighbours(radii)
    t1 = time.perf_counter()
    log(f"[cpu] neighbours {t1 - t0:.2f} s ({cores} threads)")
    f, xD, inner = om.boundary(problem.bcs, 1)
    K = om.assemble(0, problem.models, problem.materials, inner, faithful_reset=True)
    om.convection(K, problem.bcs, inner)
    f = O.first_kind_rhs(K, inner, xD, f)
    t2 = time.perf_counter()
    log(f"[cpu] assembly {t2 - t1:.2f} s")
    KL = om.assemble(0, problem.models, problem.materials, inner, part=A.NLF_PART_LOCAL, faithful_reset=True)
    P = O.OraclePrecond(A.NLF_PRECOND_ILLT, KL)
    del KL
    t3 = time.perf_counter()
    log(f"[cpu] local matrix and IC(0) {t3 - t2:.2f} s")
    _, it, _ = O.cg(K, P, f, tol=1e-300, max_it=cg_iterations)
    t4 = time.perf_counter()
    log(f"[cpu] {it} ILLT-CG iterations {t4 - t3:.2f} s")
    per_it = (t4 - t3) / max(it, 1)
    value = (t1 - t0) + (t2 - t1) + (t3 - t2) + per_it * plain_iterations
    # the GPU result against the oracle's system
    x = check["x"]
    res = float(np.linalg.norm(f - K.spmv(x)) / np.linalg.norm(f))
    kv = K.spmv(check["v"])
    kv_err = float(np.linalg.norm(check["Kv"] - kv) / np.linalg.norm(kv))
    worst = 0.0
    nnz_checked = 0
    for r0, r1, g in check["rows"]:
        o = K.export_rows(r0, r1)
        if not (np.array_equal(o.outer, g.outer) and np.array_equal(o.inner, g.inner)):
            raise RuntimeError(f"GPU pattern of rows [{r0}, {r1}) differs from the oracle's")
        worst = max(worst, float(np.abs(o.val - g.val).max() / np.abs(o.val).max()))
        nnz_checked += o.val.size
    sample = (f"oracle on the full {problem.mesh.n_nodes}-node composite, {cores} OpenMP threads: neighbours "
              f"{t1 - t0:.2f} s, assembly of K with boundary conditions {t2 - t1:.2f} s, local matrix and IC(0) "
              f"{t3 - t2:.2f} s, ILLT-CG {t4 - t3:.2f} s for {it} iterations ({per_it * 1e3:.1f} ms each) "
              f"scaled to the {plain_iterations} iterations of the same IC(0)-CG")
    return dict(value=round(value, 3), unit="s", cores=cores, kind="port", sample=sample,
                measured_seconds=round(t4 - t0, 3), host=host_description(),
                phases={"neighbours": round(t1 - t0, 3), "assembly": round(t2 - t1, 3),
                        "preconditioner": round(t3 - t2, 3), "cg_per_iteration": round(per_it, 5),
                        "cg_estimated": round(per_it * plain_iterations, 3)},
                oracle_check={"gpu_solution_residual_on_oracle_system": res,
                              "Kv_rel_diff_gpu_vs_oracle": kv_err,
                              "sampled_rows_max_rel_entry_diff": worst, "sampled_nonzeros": int(nnz_checked)})


def bench_batch(args):
    """64 uneven thermal instances (Q4 sides 32-256 log-uniform, r in [0.02, 0.1], p1 in [0.2, 1],
    lambda in [0.5, 2], seed 3): one batched assembly + lockstep mixed-precision Jacobi-PCG under a
    captured hipGraph, against 64 sequential plain FP64 ILLT-CG solves of the same instances.  (The
    batch's ILLT option cuts the iterations 2.7x here but its single-workgroup sweep runs the 64
    instances' levels one after another, 22 ms per iteration: slower than Jacobi on this batch.)"""
    print(json.dumps(batch64_record(args.tol, args.steps, args.warmup, args.batch_precond)), flush=True)


def batch64_record(tol, steps, warmup, precond="illt"):
    """The batch64 measurement as one JSON-ready dict (the --workload batch64 line, and the batch64
    key of the default line)."""
    import numpy as np
    import torch
    from nonlocal_amd import _abi as A
    from nonlocal_amd import api, batch, configs
    problems = configs.cfg3_instances(64, seed=3)
    ctx = api.default_context()
    log(f"[bench] batch64: {sum(p.mesh.n_nodes for p in problems)} nodes in {len(problems)} instances")
    # sequential plain FP64 ILLT-CG, one instance after the other
    t0 = time.perf_counter()
    seq = []
    for p in problems:
        mesh = api.mesh_2d(p.mesh, ctx)
        sol = api.stationary_heat_equation_solver_2d(mesh, p.models, p.materials, p.bcs, precond=A.NLF_PRECOND_ILLT,
                                                     tol=tol, max_it=100000)
        seq.append((sol.x.clone(), int(sol.iterations)))
        del mesh, sol
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    log(f"[bench] sequential FP64 ILLT-CG: {t_seq:.3f} s")
    cache = os.path.join(ROOT, "gpurun_out", "batch_autotune.txt") if os.path.isdir(os.path.join(ROOT, "gpurun_out")) \
        else None
    mode = A.NLF_MODE_MIXED
    times = []
    for k in range(warmup + steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bs = batch.BatchSolver(problems, ctx, precond=A.NLF_PRECOND_ILLT if precond == "illt" else A.NLF_PRECOND_JACOBI)
        t_asm = time.perf_counter() - t0
        target = bs.autotune(cache)
        t1 = time.perf_counter()
        x, its, res, st = bs.solve(tol=tol, max_it=100000, mode=mode, use_graph=True)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        log(f"[bench] batch step {k}: assembly {t_asm:.3f} s, autotune {t1 - t0 - t_asm:.3f} s, solve {t2 - t1:.3f} s, "
            f"max iterations {int(its.max())}, graph replays {st.outer_iterations}")
        if k >= warmup:
            times.append(t2 - t0)
    errs = [float(torch.linalg.norm(xi - xs) / torch.linalg.norm(xs)) for xi, (xs, _) in zip(bs.split(x), seq)]
    value = sum(times) / len(times)
    line = {
        "metric": "batched assembly+solve s, 64 uneven instances, mixed vs 64 sequential plain FP64",
        "value": round(value, 4), "unit": "s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(value * 1e3, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "64 thermal Q4 instances, sides 32-256 log-uniform, r U(0.02,0.1), p1 U(0.2,1), "
                               "lambda U(0.5,2), seed 3, solve to %g true residual" % tol,
                   "nodes": int(sum(p.mesh.n_nodes for p in problems)),
                   "nnz_full": int(bs.K.info()["nnz_full"]), "autotune_target_nnz_per_lane": target,
                   "preconditioner": "ILLT of the local matrix" if precond == "illt" else "Jacobi"},
        "sequential_fp64_illt_s": round(t_seq, 4), "speedup_vs_sequential": round(t_seq / value, 3),
        "max_rel_err_vs_fp64": max(errs), "max_true_residual": float(res.max()),
        "iterations": {"batch_max": int(its.max()), "batch_sum": int(its.sum()),
                       "sequential_illt_sum": int(sum(i for _, i in seq))},
    }
    return line


def _fp32_entry_report(K, t_fp32):
    d = K.compare_fp32()
    return dict(fp32_assembly_s=round(t_fp32, 4), entry_max_err_rel_to_max_entry=d["max_abs"] / d["max_entry"],
                entry_frobenius_rel_err=d["l2"] / d["norm"])


def bench_precision(args):
    """FP32 assembly (FP32 element-pair contractions and accumulation) and mixed-precision solves
    against the FP64 GPU path on the full-size workloads: 40x40 Q4, 200x200 jittered Q8,
    300x300 elastic Q4, 64 batched instances and the 1000x1000 composite.  The FP64 GPU path is the
    reference here (it matches the oracle to 1e-12 in the entries and 1e-10 in the solution on the
    sizes the oracle solves directly)."""
    import torch
    from nonlocal_amd import _abi as A
    from nonlocal_amd import api, batch, configs
    ctx = api.default_context()
    out = {}

    def solve(problem, mesh, precond, mode, tol, assembly32=False):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol = api._solve_driver(mesh, problem.physics, problem.models, problem.materials, problem.bcs, precond, tol,
                                100000, mode, assembly32=assembly32)
        torch.cuda.synchronize()
        return sol, time.perf_counter() - t0

    cases = [("thermal_q4_40", configs.cfg0(), A.NLF_PRECOND_ILLT),
             ("thermal_q8_200_jittered", configs.cfg1(), A.NLF_PRECOND_ILLT),
             ("elastic_q4_300", configs.cfg2(), A.NLF_PRECOND_ILLT),
             ("composite_q4_%d" % args.n, configs.cfg4(args.n), A.NLF_PRECOND_LEARNED)]
    for name, problem, precond in cases:
        log(f"[precision] {name}: {problem.mesh.n_nodes} nodes")
        mesh = api.mesh_2d(problem.mesh, ctx)
        tol = args.tol
        ref, t64 = solve(problem, mesh, A.NLF_PRECOND_ILLT, A.NLF_MODE_FP64, tol)
        x64 = ref.x.clone()
        K = ref.matrix
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        K.assemble_fp32()
        torch.cuda.synchronize()
        rec = _fp32_entry_report(K, time.perf_counter() - t0)
        rec.update(nodes=int(problem.mesh.n_nodes), nnz_full=int(K.info()["nnz_full"]),
                   fp64_values_s=round(K.timings()[1], 4), fp64_illt_s=round(t64, 4),
                   fp64_illt_iterations=int(ref.iterations))
        del ref, K
        runs = [("mixed_fp32_assembly", precond, A.NLF_MODE_MIXED, True)]
        if precond == A.NLF_PRECOND_LEARNED:
            runs.append(("learned_fcg_fp64_spmv", precond, A.NLF_MODE_FP64, False))
        for label, pc, mode, a32 in runs:
            sol, t = solve(problem, mesh, pc, mode, tol, assembly32=a32)
            err = float(torch.linalg.norm(sol.x - x64) / torch.linalg.norm(x64))
            rec[label] = dict(seconds=round(t, 4), iterations=int(sol.iterations), true_residual=sol.true_residual,
                              rel_err_vs_fp64=err, precond=int(pc), phases={k: round(v, 4) for k, v in sol.timings.items()})
            log(f"[precision] {name} {label}: {t:.3f} s, it {sol.iterations}, rel err {err:.2e}")
            del sol
        log(f"[precision] {name}: {rec}")
        out[name] = rec
        del mesh
    # 64 batched instances: one FP32 assembly of the block-diagonal batch, lockstep mixed solve
    problems = configs.cfg3_instances(64, seed=3)
    log("[precision] batch64")
    seq = []
    for p in problems:
        mesh = api.mesh_2d(p.mesh, ctx)
        sol = api.stationary_heat_equation_solver_2d(mesh, p.models, p.materials, p.bcs, precond=A.NLF_PRECOND_ILLT,
                                                     tol=tol, max_it=100000)
        seq.append(sol.x.clone())
        del mesh, sol
    bs = batch.BatchSolver(problems, ctx)
    bs.autotune(None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bs.K.assemble_fp32()
    torch.cuda.synchronize()
    rec = _fp32_entry_report(bs.K, time.perf_counter() - t0)
    x, its, res, st = bs.solve(tol=args.tol, max_it=100000, mode=A.NLF_MODE_MIXED, use_graph=True, assembly32=True)
    errs = [float(torch.linalg.norm(xi - xs) / torch.linalg.norm(xs)) for xi, xs in zip(bs.split(x), seq)]
    rec.update(nodes=int(sum(p.mesh.n_nodes for p in problems)), nnz_full=int(bs.K.info()["nnz_full"]),
               mixed_fp32_assembly=dict(max_iterations=int(its.max()), max_true_residual=float(res.max()),
                                        max_rel_err_vs_fp64=max(errs)))
    out["batch64"] = rec
    log(f"[precision] batch64: {rec}")
    line = {"metric": "FP32 assembly entry error and mixed-precision solution error vs the FP64 path",
            "unit": "relative", "n_gpus": 1, "dtype": "f32 assembly / f32 inner, f64 refinement",
            "data": "synthetic", "tol": args.tol, "configs": out}
    print(json.dumps(line), flush=True)


def bench_spmv(args):
    """The composite's FP64 SpMV alone (--steps launches after --warmup), then one device-to-device
    copy of 2^28 doubles (2 GiB read, 2 GiB written): a kernel of known traffic against which the
    HBM counters (FETCH_SIZE, WRITE_SIZE) of the SpMV are scaled."""
    import torch
    from nonlocal_amd import _abi as A
    from nonlocal_amd import api, configs
    ctx = api.default_context()
    problem = configs.cfg4(args.n)
    mesh = api.mesh_2d(problem.mesh, ctx)
    api.find_neighbours(mesh, api._radii(problem.models))
    inner = np.ones(problem.mesh.n_nodes, dtype=np.uint8)
    f = np.zeros(problem.mesh.n_nodes)
    xD = np.zeros(problem.mesh.n_nodes)
    arr = A.array_of(A.Boundary, problem.bcs)
    ctx.check(api.load().nlf_boundary_data(mesh.h, 1, arr, api._ptr(f, api._f64p), api._ptr(xD, api._f64p),
                                           api._ptr(inner, api._u8p)))
    K = api._assemble(mesh, 0, problem.models, problem.materials, inner)
    info = K.info()
    x = torch.randn(info["rows"], dtype=torch.float64, device=ctx.device)
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        K.spmv(x, y)
    ctx.synchronize()
    log(f"[spmv] {args.steps} launches, nnz_full {info['nnz_full']}")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        K.spmv(x, y)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    a = torch.empty(1 << 28, dtype=torch.float64, device=ctx.device)
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    rows = info["rows"]
    algo = info["nnz_upper"] * 12 + rows * 24
    line = {"metric": "FP64 SpMV of the composite, ms per launch (host clock over back-to-back launches)",
            "value": round(ms, 4), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "higher_is_better": False, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"composite {args.n}x{args.n} Q4, full CSR SpMV", "rows": rows,
                       "nnz_full": int(info["nnz_full"]), "nnz_upper": int(info["nnz_upper"])},
            "algorithmic_bytes": int(algo), "algorithmic_GBps": round(algo / ms * 1e-6, 1),
            "calibration_copy_bytes": {"read": int(a.numel() * 8), "written": int(b.numel() * 8)}}
    print(json.dumps(line), flush=True)


FP64_PEAK_TFLOPS = 78.6


def bench_assembly(args):
    """The composite's pattern and FP64 assembly of K alone (--steps assemblies after --warmup, each
    matrix released before the next), then one device-to-device copy of 2^28 doubles: the kernel of
    known traffic that scales the HBM counters of the value pass (k_pairs_mfma, k_assemble)."""
    import torch
    from nonlocal_amd import _abi as A
    from nonlocal_amd import api, configs
    ctx = api.default_context()
    problem = configs.cfg4(args.n)
    mesh = api.mesh_2d(problem.mesh, ctx)
    api.find_neighbours(mesh, api._radii(problem.models))
    inner = np.ones(problem.mesh.n_nodes, dtype=np.uint8)
    f = np.zeros(problem.mesh.n_nodes)
    xD = np.zeros(problem.mesh.n_nodes)
    arr = A.array_of(A.Boundary, problem.bcs)
    ctx.check(api.load().nlf_boundary_data(mesh.h, 1, arr, api._ptr(f, api._f64p), api._ptr(xD, api._f64p),
                                           api._ptr(inner, api._u8p)))
    times, info = [], None
    for k in range(args.warmup + args.steps):
        ctx.synchronize()
        t0 = time.perf_counter()
        K = api._assemble(mesh, 0, problem.models, problem.materials, inner)
        ctx.synchronize()
        if k >= args.warmup:
            times.append(time.perf_counter() - t0)
        info = K.info()
        del K
        log(f"[assembly] step {k}: {times[-1] if k >= args.warmup else float('nan'):.4f} s")
    a = torch.empty(1 << 28, dtype=torch.float64, device=ctx.device)
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    value = sum(times) / len(times)
    line = {"metric": "pattern + FP64 assembly of K, s per assembly", "value": round(value, 4), "unit": "s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": False, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"composite {args.n}x{args.n} Q4, pattern and values of K", "rows": info["rows"],
                       "nnz_full": int(info["nnz_full"]), "nnz_upper": int(info["nnz_upper"])},
            "calibration_copy_bytes": {"read": int(a.numel() * 8), "written": int(b.numel() * 8)}}
    print(json.dumps(line), flush=True)


def bench_flux(args):
    """Heat flux of the composite's temperature: gradient and -lambda grad T at the 4M quadrature points,
    the nonlocal average over every element's neighbours (the dominant kernel: one influence
    evaluation and two FMAs per pair of quadrature points), node values.  value = seconds per
    nlf_heat_flux call (device buffers allocated by the call included)."""
    import torch
    from nonlocal_amd import _abi as A
    from nonlocal_amd import api, configs
    from nonlocal_amd._lib import load
    import ctypes as C
    api.torch = torch
    problem = configs.cfg4(args.n)
    ctx = api.default_context()
    mesh = api.mesh_2d(problem.mesh, ctx)
    log(f"[flux] solving the {args.n}x{args.n} composite for its temperature")
    sol, _ = run_step(api, A, mesh, problem, A.NLF_PRECOND_LEARNED, A.NLF_MODE_FP64, args.tol)
    total = C.c_int64()
    ctx.check(load().nlf_neighbours_count(mesh.h, C.byref(total)))
    nq = 4
    pairs = int(total.value) * nq * nq                 # (qL, qNL) pairs: every group is nonlocal here
    for _ in range(args.warmup):
        api.heat_equation_solution_2d(mesh, problem.models, problem.materials, sol.x).calc_flux()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h = api.heat_equation_solution_2d(mesh, problem.models, problem.materials, sol.x)
        h.calc_flux()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    fx, fy = h.flux()
    flops = pairs * 12.0                              # phi (8) + two multiply-adds (4) per pair
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = O.set_threads(0)
        n_s = 300
        ps = configs.composite(n_s, r=10.0 / n_s)
        om = O.OracleMesh(ps.mesh)
        off, ids, _ = om.find_neighbours([[m.radius[0], m.radius[1]] for m in ps.models])
        T = np.random.default_rng(0).standard_normal(ps.mesh.n_nodes)
        t1 = time.perf_counter()
        om.heat_flux(ps.models, ps.materials, T)
        tc = time.perf_counter() - t1
        scale = pairs / (ids.size * nq * nq)
        cpu = dict(value=tc * scale, unit="s", cores=cores, kind="port",
                   sample=f"oracle calc_flux on the {n_s}x{n_s} composite with the same r/h = 10 "
                          f"({ids.size * nq * nq:.3g} quadrature-point pairs, {tc:.2f} s), scaled by pairs x{scale:.1f}")
    line = {
        "metric": "heat-flux post-processing seconds (calc_flux) on the 1000x1000 composite",
        "value": round(sec, 5), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"composite {args.n}x{args.n} Q4, 16 inclusions, p1=0.5, bell r=0.01: flux of the "
                               f"solved temperature", "quadrature_point_pairs": pairs},
        "pairs_per_s": round(pairs / sec, 1),
        "flux_abs_max": float(torch.maximum(fx.abs().max(), fy.abs().max())),
        "fp64_rate_upper_bound": {"achieved_TFLOPs_whole_call": round(flops / sec * 1e-12, 2),
                                  "peak": FP64_PEAK_TFLOPS, "flops_per_pair": 12},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this "
              f"build", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch64":
        bench_batch(args)
        return
    if args.workload == "precision":
        bench_precision(args)
        return
    if args.workload == "spmv":
        bench_spmv(args)
        return
    if args.workload == "flux":
        bench_flux(args)
        return
    if args.workload == "assembly":
        bench_assembly(args)
        return
    import torch
    from nonlocal_amd import _abi as A
    from nonlocal_amd import api, configs
    api.torch = torch

    log(f"[bench] building workload n={args.n}")
    problem = configs.cfg4(args.n)
    ctx = api.default_context()
    mesh = api.mesh_2d(problem.mesh, ctx)
    torch.cuda.synchronize()

    methods = {
        "plain_fp64_illt": (A.NLF_PRECOND_ILLT, A.NLF_MODE_FP64),
        "mixed_illt": (A.NLF_PRECOND_ILLT, A.NLF_MODE_MIXED),
        "mixed_jacobi": (A.NLF_PRECOND_JACOBI, A.NLF_MODE_MIXED),
        "fp64_jacobi": (A.NLF_PRECOND_JACOBI, A.NLF_MODE_FP64),
        "learned_fcg": (A.NLF_PRECOND_LEARNED, A.NLF_MODE_FP64),
        "learned_fcg_split": (A.NLF_PRECOND_LEARNED, A.NLF_MODE_SPLIT),
        "learned_fcg_mixed": (A.NLF_PRECOND_LEARNED, A.NLF_MODE_MIXED),
        "split_jacobi": (A.NLF_PRECOND_JACOBI, A.NLF_MODE_SPLIT),
        "learned_fcg_onelevel": (A.NLF_PRECOND_LEARNED, A.NLF_MODE_FP64),
        # the flagship's two-level form with the network switched off (its fixed Jacobi term and the
        # same coarse space): the difference to learned_fcg is what the trained network adds
        "coarse_jacobi": (A.NLF_PRECOND_LEARNED, A.NLF_MODE_FP64),
    }
    coarse_of = {"learned_fcg_onelevel": 0}
    smoother_of = {"coarse_jacobi": "jacobi"}
    # plain FP64 ILLT-CG: the baseline the metric compares against; also the FP64 reference solution
    log("[bench] plain FP64 ILLT-CG")
    ref, t_plain = run_step(api, A, mesh, problem, *methods["plain_fp64_illt"], args.tol)
    x_ref = ref.x.clone()
    log(f"[bench] plain: {t_plain:.3f} s, it {ref.iterations}, true res {ref.true_residual:.2e}, "
        f"timings {ref.timings}")
    results = {"plain_fp64_illt": dict(seconds=t_plain, iterations=int(ref.iterations),
                                       true_residual=ref.true_residual,
                                       phases={k: round(v, 4) for k, v in ref.timings.items()})}
    del ref                                   # its matrices and factor are several GB: free them now
    best = None
    for name in [m for m in args.methods.split(",") if m]:
        pc, mode = methods[name]
        sol, t = run_step(api, A, mesh, problem, pc, mode, args.tol, coarse_of.get(name), smoother_of.get(name))
        err = float(torch.linalg.norm(sol.x - x_ref) / torch.linalg.norm(x_ref))
        log(f"[bench] {name}: {t:.3f} s, it {sol.iterations}, outer {sol.stats.outer_iterations}, "
            f"true res {sol.true_residual:.2e}, rel err vs FP64 {err:.2e}, timings {sol.timings}")
        results[name] = dict(seconds=t, iterations=int(sol.iterations), true_residual=sol.true_residual,
                             rel_err_vs_fp64=err, phases={k: round(v, 4) for k, v in sol.timings.items()})
        del sol
        if err <= 1e-8 and (best is None or t < results[best]["seconds"]):
            best = name
    if best is None:
        best = "plain_fp64_illt"
    pc, mode = methods[best]
    log(f"[bench] flagship method: {best}; warmup {args.warmup}, steps {args.steps}")
    for _ in range(args.warmup):
        run_step(api, A, mesh, problem, pc, mode, args.tol, coarse_of.get(best), smoother_of.get(best))
    torch.cuda.synchronize()
    # Each step's Solution (matrix, local matrix, preconditioner) is released before the next step
    # starts, so every step runs with the same free device memory; only the last step's results are
    # kept for the line.
    step_s = []
    last_sol = None
    t0 = time.perf_counter()
    for k in range(args.steps):
        last_sol = None
        sol, t_step = run_step(api, A, mesh, problem, pc, mode, args.tol, coarse_of.get(best), smoother_of.get(best))
        step_s.append(t_step)
        if k == args.steps - 1:
            last_sol = sol
        del sol
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sec_per_step = elapsed / args.steps
    # after the timed region: the last step's results, and for the oracle check K v for one random v
    # and a few sampled row ranges of the upper CSR
    sol = last_sol
    last = dict(x=sol.x.clone(), info=sol.matrix.info(), stats=sol.stats, timings=dict(sol.timings),
                iterations=int(sol.iterations), true_residual=sol.true_residual,
                symv=sol.matrix.symv_info() if mode == A.NLF_MODE_FP64 else None)
    if not args.no_cpu:
        n_rows = last["info"]["rows"]
        v = torch.from_numpy(np.random.default_rng(5).standard_normal(n_rows)).to(sol.x.device)
        starts = np.linspace(0, n_rows - 64, 48).astype(np.int64)
        last["check"] = dict(v=v.cpu().numpy(), Kv=sol.matrix.spmv(v).cpu().numpy(), x=sol.x.cpu().numpy(),
                             rows=[(int(r), int(r) + 64, sol.matrix.export_rows(int(r), int(r) + 64)) for r in starts])
    del sol, last_sol
    err = float(torch.linalg.norm(last["x"] - x_ref) / torch.linalg.norm(x_ref))
    info = last["info"]
    st = last["stats"]
    rows = info["rows"]
    sv = {A.NLF_MODE_MIXED: 4, A.NLF_MODE_SPLIT: 6}.get(mode, 8)      # matrix value bytes (vectors stay FP64)
    sx = 8
    algo_bytes = info["nnz_upper"] * (sv + 4) + rows * (8 + 2 * sx)     # upper-symmetric CSR formula
    achieved = algo_bytes / (st.spmv_ms * 1e-3) / 1e9 if st.spmv_ms > 0 else 0.0
    symmetric = sv == 8 and last.get("symv") is not None
    moved_gbs = st.spmv_bytes / (st.spmv_ms * 1e-3) / 1e9 if st.spmv_ms > 0 else 0.0
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(problem, results["plain_fp64_illt"]["iterations"], last["check"], args.cpu_cg_iterations)
        log(f"[bench] CPU baseline: {cpu['value']} s ({cpu['sample']}); oracle check {cpu['oracle_check']}")
    line = {
        "metric": "assembly+solve s & %roofline, learned/mixed vs plain FP64 GPU at <=1e-8 rel err",
        "value": round(sec_per_step, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_per_step * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if (mode != A.NLF_MODE_FP64 or pc == A.NLF_PRECOND_LEARNED) else "f64",
        "data": "synthetic",
        "config": {"workload": f"composite 2D stationary heat, {args.n}x{args.n} Q4, 16 inclusions, p1=0.5, "
                               f"bell r=0.01, solve to {args.tol:g} true residual",
                   "method": best, "nodes": int(problem.mesh.n_nodes), "nnz_upper": int(info["nnz_upper"]),
                   "nnz_full": int(info["nnz_full"])},
        "step_seconds": {"mean": round(sec_per_step, 4), "median": round(float(np.median(step_s)), 4),
                         "min": round(min(step_s), 4), "max": round(max(step_s), 4)},
        "plain_fp64_illt_s": round(t_plain, 4),
        "speedup_vs_plain_fp64_illt": round(t_plain / sec_per_step, 3),
        "rel_err_vs_fp64": err,
        "true_residual": last["true_residual"],
        "iterations": last["iterations"],
        "methods": results,
        "phase_seconds": {k: round(v, 4) for k, v in last["timings"].items()},
        "precond_apply_ms": round(st.precond_ms, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": counter_traffic(args.n, sv == 8, symmetric),
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE_SYM if symmetric else TRAFFIC_PROFILE, ROOT),
                     "kernel": ("k_symv_slices + k_symv_reduce (fp64 upper triangle read once, 64-row slices)"
                                if symmetric else
                                f"k_spmv2 ({ {4: 'fp32', 6: 'fp32 + 16-bit remainder'}.get(sv, 'fp64')} values, full CSR)"),
                     "avg_launch_ms": round(st.spmv_ms, 4), "samples": int(st.spmv_samples),
                     "algorithmic_bytes_per_launch": int(algo_bytes),
                     "moved_bytes_per_launch": int(st.spmv_bytes), "moved_GBps": round(moved_gbs, 1)},
        "cpu_baseline": cpu,
    }
    if not args.no_batch and args.n == 1000:
        # the batching workload beside the headline (after its timed region; not part of `value`)
        log("[bench] batch64")
        b = batch64_record(args.tol, 2, 1, args.batch_precond)
        line["batch64"] = {k: b[k] for k in ("value", "unit", "sequential_fp64_illt_s", "speedup_vs_sequential",
                                             "max_rel_err_vs_fp64", "max_true_residual", "iterations")}
        line["batch64"]["config"] = b["config"]["workload"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
