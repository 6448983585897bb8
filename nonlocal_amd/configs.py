"""The benchmark workloads of the 2D nonlocal hot path, as synthetic problem descriptions.

Each builder returns a Problem: mesh, per-group models and materials, boundary conditions (one per
boundary group: left x=0, right x=1, bottom y=0, top y=1) and the physics.  `scale` shrinks the
mesh for quick parity runs while keeping every other parameter.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _abi as A
from . import meshgen


@dataclass
class Problem:
    name: str
    mesh: meshgen.Mesh2D
    physics: int          # 0 thermal, 1 elastic
    models: list
    materials: list
    bcs: list
    extra: dict

    @property
    def dof(self):
        return 1 if self.physics == 0 else 2


def bell(p1, r):
    """Polynomial bell phi = c (1 - rho^2/r^2)^2 (n = p = 2, q = 2), normalised to unit integral."""
    return A.model(p1, r, influence=A.NLF_INFLUENCE_POLYNOMIAL, p=2.0, q=2.0)


def thermal_bcs(robin_top=None):
    """T = 0 on x = 0, unit inward flux on x = 1, insulated elsewhere (optional Robin h on y = 1)."""
    top = A.boundary() if robin_top is None else A.boundary(A.NLF_BC_CONVECTION, heat_transfer=robin_top, ambient=0.0)
    return [A.boundary(A.NLF_BC_FIRST_KIND, 0.0), A.boundary(A.NLF_BC_FLUX, 1.0), A.boundary(), top]


def thermal_q4(n=40, p1=0.5, r=0.1, lam=1.0, name="thermal_q4"):
    """Structured n x n bilinear quads on the unit square, isotropic lambda."""
    mesh = meshgen.structured_q4(n)
    return Problem(name, mesh, 0, [bell(p1, r)], [A.thermal_material(A.NLF_CONDUCTIVITY_ISOTROPIC, (lam, 0, 0))],
                   thermal_bcs(), dict(n=n, p1=p1, r=r))


def cfg0(scale=1.0):
    """40 x 40 Q4, lambda = 1, p1 = 0.5, bell r = 0.1."""
    return thermal_q4(max(4, int(round(40 * scale))), name="cfg0_thermal_q4_40")


def cfg1(scale=1.0, seed=1):
    """200 x 200 Q8 (p = 2/9), Gauss 3x3, jittered interior nodes, anisotropic lambda, p1 = 0.3,
    r = 0.05, Robin h = 0.5 on y = 1."""
    n = max(4, int(round(200 * scale)))
    mesh = meshgen.structured_q8(n, jitter=0.2, seed=seed)
    mat = A.thermal_material(A.NLF_CONDUCTIVITY_ANISOTROPIC, (1.0, 0.25, 0.1))
    return Problem(f"cfg1_thermal_q8_{n}", mesh, 0, [bell(0.3, 0.05)], [mat], thermal_bcs(robin_top=0.5),
                   dict(n=n, p1=0.3, r=0.05, seed=seed))


def cfg2(scale=1.0):
    """300 x 300 Q4 plane stress, E = 1, nu = 0.3, p1 = 0.3, r = 0.03; clamped x = 0, traction (1, 0) on x = 1."""
    n = max(4, int(round(300 * scale)))
    mesh = meshgen.structured_q4(n)
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0, A.NLF_BC_FIRST_KIND, 0.0),
           A.boundary(A.NLF_BC_FLUX, 1.0, A.NLF_BC_FLUX, 0.0), A.boundary(), A.boundary()]
    return Problem(f"cfg2_elastic_q4_{n}", mesh, 1, [bell(0.3, 0.03)], [A.elastic_material(1.0, 0.3)], bcs,
                   dict(n=n, p1=0.3, r=0.03))


def cfg3_instances(count=64, seed=3, nmin=32, nmax=256):
    """Uneven batch: side n ~ exp U(ln nmin, ln nmax), r ~ U(0.02, 0.1), p1 ~ U(0.2, 1), lambda ~ U(0.5, 2)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        r = float(rng.uniform(0.02, 0.1))
        p1 = float(rng.uniform(0.2, 1.0))
        lam = float(rng.uniform(0.5, 2.0))
        out.append(thermal_q4(n, p1, r, lam, name=f"cfg3_{k}_n{n}"))
    return out


def cfg4(n=1000, seed=4, count=16):
    """Composite: n x n Q4, lambda = 1 matrix with `count` circular inclusions of lambda = 10 (radii
    U(0.03, 0.08), seeded, non-overlapping), each inclusion its own group; p1 = 0.5, r = 0.01."""
    mesh, circles = meshgen.composite_q4(n, count=count, seed=seed)
    G = mesh.n_groups
    models = [bell(0.5, 0.01) for _ in range(G)]
    mats = [A.thermal_material(A.NLF_CONDUCTIVITY_ISOTROPIC, (1.0 if g == 0 else 10.0, 0, 0)) for g in range(G)]
    return Problem(f"cfg4_composite_q4_{n}", mesh, 0, models, mats, thermal_bcs(), dict(n=n, circles=circles))


def composite(n, r=None, seed=4, count=16):
    """Composite of another size (training meshes for the learned preconditioner, parity tests);
    the radius defaults to 10 h as at n = 1000."""
    p = cfg4(n, seed, count)
    r = 10.0 / n if r is None else r
    p.models = [bell(0.5, r) for _ in p.models]
    p.name = f"composite_q4_{n}"
    return p
