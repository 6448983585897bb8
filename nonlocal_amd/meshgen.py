"""Synthetic 2D meshes for the nonlocal hot path.

The reference reads SU2 files (libraries/mesh/mesh_2d/su2_parser.hpp:165-195) whose element
numbering follows std::unordered_set iteration order.  Here meshes are built directly with a
documented numbering so that neighbour lists and sparsity patterns are reproducible:

* Q4: node id = j*(n+1) + i (x fastest), element id = j*n + i, local node order
  (-1,-1), (1,-1), (1,1), (-1,1) as in serendipity<T,1> (serendipity_2d.hpp:47-56).
* Q8: nodes on the half-step grid (2n+1)^2 without cell centres, numbered by a row-major sweep;
  local order c0, m01, c1, m12, c2, m23, c3, m30 as in serendipity<T,2> (serendipity_2d.hpp:84-99).
* Element groups are contiguous element ranges (mesh_container_2d.hpp:30-32); boundary groups
  hold 1D edge elements, linear [a, b] or quadratic [a, mid, b] (local nodes -1, 0, 1,
  lagrangian_element_1d<T,N> with uniform_partition nodes).

Boundary groups are always, in this order: left (x=0), right (x=1), bottom (y=0), top (y=1).
"""
from __future__ import annotations

from dataclasses import dataclass, field
import numpy as np

Q4 = 4
Q8 = 8
BOUNDARY_NAMES = ("left", "right", "bottom", "top")


@dataclass
class Mesh2D:
    xy: np.ndarray                 # (N, 2) float64
    elems: np.ndarray              # (E, nn) int32
    nn: int                        # 4 (bilinear, Gauss 2x2) or 8 (serendipity, Gauss 3x3)
    group_begin: np.ndarray        # (G+1,) int64, contiguous element ranges
    group_names: list
    bedges: np.ndarray             # (B, 2|3) int32
    bgroup_begin: np.ndarray       # (4+1,) int64
    bgroup_names: list = field(default_factory=lambda: list(BOUNDARY_NAMES))
    n: int = 0                     # elements per side (structured generators)

    @property
    def n_nodes(self) -> int:
        return int(self.xy.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.elems.shape[0])

    @property
    def n_groups(self) -> int:
        return len(self.group_names)

    def boundary_nodes(self, name: str) -> np.ndarray:
        k = self.bgroup_names.index(name)
        e = self.bedges[self.bgroup_begin[k]:self.bgroup_begin[k + 1]]
        return np.unique(e.ravel())


def _boundary_edges_q4(n: int):
    def v(i, j):
        return j * (n + 1) + i
    left = [[v(0, j), v(0, j + 1)] for j in range(n)]
    right = [[v(n, j), v(n, j + 1)] for j in range(n)]
    bottom = [[v(i, 0), v(i + 1, 0)] for i in range(n)]
    top = [[v(i, n), v(i + 1, n)] for i in range(n)]
    edges = left + right + bottom + top
    begin = np.array([0, n, 2 * n, 3 * n, 4 * n], dtype=np.int64)
    return np.asarray(edges, dtype=np.int32), begin


def structured_q4(n: int, jitter: float = 0.0, seed: int = 0) -> Mesh2D:
    """n x n bilinear quads on the unit square; optional interior-node jitter U(-a, a)^2, a = jitter*h."""
    i = np.arange(n + 1, dtype=np.float64)
    xs = i / n
    X, Y = np.meshgrid(xs, xs, indexing="xy")          # X[j, i] = i/n
    xy = np.stack([X.ravel(), Y.ravel()], axis=1).copy()
    if jitter > 0.0:
        h = 1.0 / n
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter * h, jitter * h, size=xy.shape)
        ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
        interior = ((ii > 0) & (ii < n) & (jj > 0) & (jj < n)).ravel()
        xy[interior] += d[interior]
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    base = jj * (n + 1) + ii
    elems = np.stack([base, base + 1, base + n + 2, base + n + 1], axis=1).astype(np.int32)
    bedges, bbegin = _boundary_edges_q4(n)
    return Mesh2D(xy=xy, elems=elems, nn=Q4, group_begin=np.array([0, n * n], dtype=np.int64),
                  group_names=["material"], bedges=bedges, bgroup_begin=bbegin, n=n)


def structured_q8(n: int, jitter: float = 0.0, seed: int = 0) -> Mesh2D:
    """n x n serendipity quads; jitter amplitude is jitter * (node spacing h/2) on interior nodes."""
    m = 2 * n + 1
    ids = -np.ones((m, m), dtype=np.int64)             # ids[J, I]
    coords = []
    nid = 0
    for J in range(m):
        for I in range(m):
            if (I % 2 == 1) and (J % 2 == 1):
                continue
            ids[J, I] = nid
            coords.append((I / (2.0 * n), J / (2.0 * n)))
            nid += 1
    xy = np.asarray(coords, dtype=np.float64)
    if jitter > 0.0:
        hs = 0.5 / n
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter * hs, jitter * hs, size=xy.shape)
        interior = np.zeros(nid, dtype=bool)
        for J in range(1, m - 1):
            for I in range(1, m - 1):
                if ids[J, I] >= 0:
                    interior[ids[J, I]] = True
        xy[interior] += d[interior]
    el = []
    for j in range(n):
        for i in range(n):
            I, J = 2 * i, 2 * j
            el.append([ids[J, I], ids[J, I + 1], ids[J, I + 2], ids[J + 1, I + 2],
                       ids[J + 2, I + 2], ids[J + 2, I + 1], ids[J + 2, I], ids[J + 1, I]])
    elems = np.asarray(el, dtype=np.int32)
    left = [[ids[2 * j, 0], ids[2 * j + 1, 0], ids[2 * j + 2, 0]] for j in range(n)]
    right = [[ids[2 * j, m - 1], ids[2 * j + 1, m - 1], ids[2 * j + 2, m - 1]] for j in range(n)]
    bottom = [[ids[0, 2 * i], ids[0, 2 * i + 1], ids[0, 2 * i + 2]] for i in range(n)]
    top = [[ids[m - 1, 2 * i], ids[m - 1, 2 * i + 1], ids[m - 1, 2 * i + 2]] for i in range(n)]
    bedges = np.asarray(left + right + bottom + top, dtype=np.int32)
    bbegin = np.array([0, n, 2 * n, 3 * n, 4 * n], dtype=np.int64)
    return Mesh2D(xy=xy, elems=elems, nn=Q8, group_begin=np.array([0, n * n], dtype=np.int64),
                  group_names=["material"], bedges=bedges, bgroup_begin=bbegin, n=n)


def sample_inclusions(count: int, rmin: float, rmax: float, seed: int):
    """Non-overlapping circles inside the unit square, rejection-sampled from one seed."""
    rng = np.random.default_rng(seed)
    circles = []
    tries = 0
    while len(circles) < count:
        tries += 1
        if tries > 100000:
            raise RuntimeError("could not place inclusions")
        r = rng.uniform(rmin, rmax)
        cx, cy = rng.uniform(r, 1.0 - r, size=2)
        if all((cx - x) ** 2 + (cy - y) ** 2 >= (r + q) ** 2 for x, y, q in circles):
            circles.append((cx, cy, r))
    return circles


def composite_q4(n: int, count: int = 16, rmin: float = 0.03, rmax: float = 0.08, seed: int = 4):
    """Structured Q4 mesh whose elements are regrouped into a matrix group and one group per inclusion.

    An element belongs to inclusion k when its vertex-mean centre lies strictly inside circle k.
    Group order: matrix, inclusion 0..count-1; each group keeps ascending structured element order.
    Returns (mesh, circles).
    """
    base = structured_q4(n)
    circles = sample_inclusions(count, rmin, rmax, seed)
    c = base.xy[base.elems].mean(axis=1)
    label = np.zeros(base.n_elems, dtype=np.int64)
    for k, (x, y, r) in enumerate(circles):
        inside = (c[:, 0] - x) ** 2 + (c[:, 1] - y) ** 2 < r * r
        label[inside] = k + 1
    order = np.argsort(label, kind="stable")
    counts = np.bincount(label, minlength=count + 1)
    begin = np.zeros(count + 2, dtype=np.int64)
    begin[1:] = np.cumsum(counts)
    mesh = Mesh2D(xy=base.xy, elems=base.elems[order].copy(), nn=Q4, group_begin=begin,
                  group_names=["matrix"] + [f"inclusion{k}" for k in range(count)],
                  bedges=base.bedges, bgroup_begin=base.bgroup_begin, n=n)
    return mesh, circles


def element_centres_numpy(mesh: Mesh2D) -> np.ndarray:
    """Centre c_e = sum_i x_i N_i(0,0) in the reference's accumulation order (0 + a0) + a1 + ...

    mesh_container_2d.hpp:130-139 evaluates N at xi0 = 0 for quads; N_i(0,0) is 1/4 for Q4 and
    1/16 (corners) / 3/16 (midsides) for Q8 with p = 2/9.
    """
    if mesh.nn == Q4:
        w = np.full(4, 0.25)
    else:
        p = 2.0 / 9.0
        wc = (9.0 * p - 1.0) / 16.0
        wm = -(9.0 * p - 5.0) / 16.0
        w = np.array([wc, wm, wc, wm, wc, wm, wc, wm])
    c = np.zeros((mesh.n_elems, 2))
    for i in range(mesh.nn):
        c = c + mesh.xy[mesh.elems[:, i]] * w[i]
    return c
