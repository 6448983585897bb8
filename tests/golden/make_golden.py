"""Writes tests/golden/golden.npz from the oracle (run from the repository root).

Fixtures: the 40 x 40 thermal workload's neighbour lists (SHA-256 + sampled rows), its upper CSR
pattern (SHA-256) and sampled entries, its solution vector; the full conductivity matrix of a 4 x 4
mesh for every influence family; jittered Q8 centres.
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from nonlocal_amd import _abi as A  # noqa: E402
from nonlocal_amd import configs, meshgen  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests import helpers as H  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
FAMILIES = {"fast_polynomial": (A.NLF_INFLUENCE_POLYNOMIAL, 2, 1), "bell": (A.NLF_INFLUENCE_POLYNOMIAL, 2, 2),
            "polynomial_p4_q3": (A.NLF_INFLUENCE_POLYNOMIAL, 4, 3), "constant": (A.NLF_INFLUENCE_CONSTANT, 2, 1),
            "exponential": (A.NLF_INFLUENCE_EXPONENTIAL, 2, 0.5)}
SAMPLE_ELEMS = np.array([0, 1, 39, 40, 41, 799, 820, 1200, 1559, 1599])


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def cfg0_data():
    p = configs.cfg0()
    res = O.thermal_problem(p.mesh, p.models, p.materials, p.bcs, precond=A.NLF_PRECOND_ILLT, tol=1e-14)
    off, ids, exp = res["neighbours"]
    K = res["K"].export(0)
    x = H.direct_solution(res["K"], res["f"])
    rng = np.random.default_rng(0)
    pick = np.sort(rng.choice(K.val.size, 500, replace=False))
    rows = np.searchsorted(K.outer, pick, side="right") - 1
    samples = {f"nb_{e}": ids[off[e]:off[e + 1]] for e in SAMPLE_ELEMS}
    return dict(cfg0_nb_offsets=off, cfg0_nb_sha=np.array(sha(off, ids)), cfg0_expanded=exp,
                cfg0_pattern_sha=np.array(sha(K.outer, K.inner)), cfg0_entry_pos=pick, cfg0_entry_row=rows,
                cfg0_entry_col=K.inner[pick], cfg0_entry_val=K.val[pick], cfg0_f=res["f"], cfg0_solution=x,
                **{f"cfg0_{k}": v for k, v in samples.items()})


def small_matrices():
    out = {}
    mesh = meshgen.structured_q4(4, 0.1, seed=11)
    for name, (kind, p, q) in FAMILIES.items():
        om = O.OracleMesh(mesh)
        model = A.model(0.4, 0.45, influence=kind, p=p, q=q)
        om.find_neighbours([[0.45, 0.45]])
        K = om.assemble(0, [model], [A.thermal_material()], np.ones(mesh.n_nodes, np.uint8)).export(0)
        out[f"k4_{name}_outer"] = K.outer
        out[f"k4_{name}_inner"] = K.inner
        out[f"k4_{name}_val"] = K.val
    return out


def main():
    data = {}
    data.update(cfg0_data())
    data.update(small_matrices())
    q8 = meshgen.structured_q8(6, 0.2, seed=12)
    data["q8_centres"] = O.OracleMesh(q8).centres()
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
