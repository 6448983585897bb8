"""Train the learned preconditioner on the GPU and write a weights file.

    python -m nonlocal_amd.train_learned --sizes 64 96 128 192 256 --steps 3000 --out weights.bin

Training data: seeded composite meshes (matrix lambda = 1, circular inclusions lambda = 10, p1 = 0.5,
bell influence with r = 10 h); (residual, error) pairs anchored on each problem's FP64 solution
(white-noise errors around x*, residuals evaluated in FP64).  Forward,
backward and Adam run in libnlf.so.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import api, learned


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--sizes", type=int, nargs="+", default=[64, 96, 128, 192, 256])
    ap.add_argument("--seeds-per-size", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=8, help="(residual, error) pairs per problem")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=learned.DEFAULT_WEIGHTS)
    ap.add_argument("--log", default=None, help="JSON file with the loss history")
    a = ap.parse_args(argv)

    ctx = api.default_context()
    rng = np.random.default_rng(a.seed)
    problems = []
    t0 = time.perf_counter()
    for n in a.sizes:
        for k in range(a.seeds_per_size):
            seed = 100 + 7 * n + k
            tp = learned.TrainingProblem(n, seed=seed, ctx=ctx)
            tp.make_pairs(rng, count=a.pairs)
            problems.append(tp)
            print(f"[train] problem n={n} seed={seed}: {tp.b.numel()} nodes, {len(tp.pairs)} pairs "
                  f"({time.perf_counter() - t0:.1f} s)", flush=True)
    w, hist = learned.train(problems, steps=a.steps, lr=a.lr, seed=a.seed,
                            log=lambda m: print(m, flush=True))
    config = dict(sizes=a.sizes, seeds_per_size=a.seeds_per_size, pairs=a.pairs, steps=a.steps,
                  lr=a.lr, seed=a.seed)
    learned.save_weights(a.out, w, a.seed, config)
    print(f"[train] wrote {a.out} ({time.perf_counter() - t0:.1f} s)", flush=True)
    if a.log:
        os.makedirs(os.path.dirname(os.path.abspath(a.log)), exist_ok=True)
        with open(a.log, "w") as f:
            json.dump(dict(config=config, loss=[float(x) for x in hist]), f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
