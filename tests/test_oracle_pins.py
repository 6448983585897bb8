"""Pins of oracle functions that the reference's own tests leave open: the diameter-expanding
strategies NO, MIN and MEAN of find_neighbours (find_neighbours.hpp:49-84) and the Robin
(convection) term with a nonzero ambient temperature (convection_condition_2d.hpp:11-31,
thermal_boundary_conditions_2d.hpp:54-56).  Each check is a closed form or an independent
computation, chosen so that a plausible slip in the oracle (a dropped term, a wrong sign, the
lower instead of the upper triangle, a pair counted twice) fails it.
"""
import math

import numpy as np
import pytest

from nonlocal_amd import _abi as A
from nonlocal_amd import configs, meshgen
from oracle import oracle as O
from tests import helpers as H


# ---- neighbour lists for a given expanded radius: all-pairs within the group -------------------
def lists_for_radius(mesh, expanded):
    """All-pairs search with the reference predicate ((0 + z0^2) + z1^2 <= 1, z = (c - c')/r) and
    the reference output order (x-cell, y-cell, id of the cell grid of side max(r))."""
    c = meshgen.element_centres_numpy(mesh)
    lists = [[] for _ in range(mesh.n_elems)]
    for g in range(mesh.n_groups):
        b, e = mesh.group_begin[g], mesh.group_begin[g + 1]
        r = np.asarray(expanded[g], dtype=np.float64)
        if r[0] <= 0 or r[1] <= 0:
            continue
        pts = mesh.xy[mesh.elems[b:e].ravel()]
        left, down = pts[:, 0].min(), pts[:, 1].min()
        R = max(r)
        ids = np.arange(b, e)
        cell = np.stack([((c[ids, 0] - left) / R).astype(np.int64), ((c[ids, 1] - down) / R).astype(np.int64)], 1)
        for el in ids:
            z0 = (c[el, 0] - c[ids, 0]) / r[0]
            z1 = (c[el, 1] - c[ids, 1]) / r[1]
            inside = ((0.0 + z0 * z0) + z1 * z1) <= 1.0
            cand = ids[inside]
            order = np.lexsort((cand, cell[inside, 1], cell[inside, 0]))
            lists[el] = list(cand[order])
    off = np.zeros(mesh.n_elems + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(x) for x in lists])
    ids = np.concatenate([np.asarray(x, dtype=np.int32) for x in lists]) if off[-1] else np.zeros(0, np.int32)
    return off, ids


def node_adjacent_pairs(mesh, g):
    """The distinct unordered pairs {e, a}, e in group g, a sharing a node with e (a in any group,
    e itself included), as a set: independent of any iteration order."""
    node_elems = [[] for _ in range(mesh.n_nodes)]
    for e, nodes in enumerate(mesh.elems):
        for v in nodes:
            node_elems[v].append(e)
    pairs = set()
    for e in range(mesh.group_begin[g], mesh.group_begin[g + 1]):
        for v in mesh.elems[e]:
            for a in node_elems[v]:
                pairs.add((min(e, a), max(e, a)))
    return pairs


MESHES = {
    "q4_jitter": lambda: (meshgen.structured_q4(14, 0.2, seed=3), [[0.15, 0.12]]),
    "q8_jitter": lambda: (meshgen.structured_q8(9, 0.2, seed=4), [[0.21, 0.21]]),
    "composite": lambda: (meshgen.composite_q4(24, count=3, seed=6)[0], None),
}


def _mesh(case):
    mesh, radii = MESHES[case]()
    if radii is None:
        radii = [[0.13, 0.13]] * mesh.n_groups
    return mesh, radii


@pytest.mark.parametrize("case", list(MESHES))
@pytest.mark.parametrize("strategy", [A.NLF_EXPAND_NO, A.NLF_EXPAND_MIN])
def test_no_and_min_expansion_add_nothing(case, strategy):
    """NO adds 0 by definition.  MIN takes the smallest centre distance over node-adjacent pairs, and
    every element is node-adjacent to itself (distance 0), so MIN adds exactly 0 as well.  The lists
    are then the all-pairs lists at the unexpanded radius."""
    mesh, radii = _mesh(case)
    off, ids, exp = O.OracleMesh(mesh).find_neighbours(radii, strategy)
    assert np.array_equal(exp, np.asarray(radii, dtype=np.float64))
    boff, bids = lists_for_radius(mesh, exp)
    assert np.array_equal(off, boff) and np.array_equal(ids, bids)


def test_no_expansion_lists_are_smaller_than_max():
    """Monotonicity in the radius: the NO lists are subsets of the MAX lists, strictly smaller somewhere."""
    mesh, radii = _mesh("q4_jitter")
    om = O.OracleMesh(mesh)
    off0, ids0, _ = om.find_neighbours(radii, A.NLF_EXPAND_NO)
    off1, ids1, _ = om.find_neighbours(radii, A.NLF_EXPAND_MAX)
    for e in range(mesh.n_elems):
        assert set(ids0[off0[e]:off0[e + 1]]) <= set(ids1[off1[e]:off1[e + 1]])
    assert off0[-1] < off1[-1]


@pytest.mark.parametrize("n", [1, 2, 7, 12])
def test_mean_expansion_closed_form_structured_q4(n):
    """n x n uniform grid, spacing h: the node-adjacent pairs are n^2 self pairs (distance 0),
    2n(n-1) edge pairs (distance h) and 2(n-1)^2 diagonal pairs (distance h sqrt 2)."""
    mesh = meshgen.structured_q4(n)
    r = 0.3
    _, _, exp = O.OracleMesh(mesh).find_neighbours([[r, r]], A.NLF_EXPAND_MEAN)
    h = 1.0 / n
    pairs = n * n + 2 * n * (n - 1) + 2 * (n - 1) ** 2
    mean = (2 * n * (n - 1) * h + 2 * (n - 1) ** 2 * h * math.sqrt(2.0)) / pairs
    assert abs(exp[0, 0] - (r + mean)) <= 2e-15 * (r + mean)          # a running sum of up to ~600 terms
    assert exp[0, 0] == exp[0, 1]


def test_mean_expansion_closed_form_structured_q8():
    """Q8 elements of a uniform grid share nodes exactly when the Q4 cells do: same pairs, same mean."""
    n = 6
    mesh = meshgen.structured_q8(n)
    _, _, exp = O.OracleMesh(mesh).find_neighbours([[0.25, 0.25]], A.NLF_EXPAND_MEAN)
    h = 1.0 / n
    pairs = n * n + 2 * n * (n - 1) + 2 * (n - 1) ** 2
    mean = (2 * n * (n - 1) * h + 2 * (n - 1) ** 2 * h * math.sqrt(2.0)) / pairs
    assert abs(exp[0, 0] - (0.25 + mean)) <= 2e-15 * (0.25 + mean)


@pytest.mark.parametrize("case", list(MESHES))
def test_mean_expansion_equals_set_mean(case):
    """On jittered and multi-group meshes: the mean centre distance over the set of distinct pairs
    (computed from a Python set, so a pair counted twice or a cross-group pair dropped would show),
    then the lists at that radius equal the all-pairs lists."""
    mesh, radii = _mesh(case)
    off, ids, exp = O.OracleMesh(mesh).find_neighbours(radii, A.NLF_EXPAND_MEAN)
    c = meshgen.element_centres_numpy(mesh)
    for g in range(mesh.n_groups):
        pairs = node_adjacent_pairs(mesh, g)
        d = np.array([math.hypot(*(c[a] - c[b])) for a, b in pairs])
        mean = math.fsum(d) / len(pairs)
        assert abs(exp[g, 0] - (radii[g][0] + mean)) <= 1e-15 * exp[g, 0]
        assert abs(exp[g, 1] - (radii[g][1] + mean)) <= 1e-15 * exp[g, 1]
    boff, bids = lists_for_radius(mesh, exp)
    assert np.array_equal(off, boff) and np.array_equal(ids, bids)


# ---- the Robin term --------------------------------------------------------------------------
def robin_problem(mesh, h, ambient):
    """T = 0 on x = 0, convection h (T_a) on x = 1, insulated elsewhere."""
    return [A.boundary(A.NLF_BC_FIRST_KIND, 0.0), A.boundary(A.NLF_BC_CONVECTION, heat_transfer=h, ambient=ambient),
            A.boundary(), A.boundary()]


@pytest.mark.parametrize("maker", ["q4", "q8"])
@pytest.mark.parametrize("h,ambient", [(0.7, 2.0), (3.0, -0.5)])
def test_robin_patch_linear_solution(maker, h, ambient):
    """-T'' = 0, T(0) = 0, -T'(1) = h (T(1) - T_a): T = h T_a x / (1 + h), linear, so bilinear and
    serendipity elements (jittered) reproduce it to rounding."""
    mesh = meshgen.structured_q4(9, 0.2, seed=11) if maker == "q4" else meshgen.structured_q8(6, 0.2, seed=12)
    bcs = robin_problem(mesh, h, ambient)
    res = O.thermal_problem(mesh, [A.model(1.0, 0.0)], [A.thermal_material()], bcs, precond=A.NLF_PRECOND_ILLT,
                            tol=1e-15)
    T = res["T"]
    exact = h * ambient * mesh.xy[:, 0] / (1.0 + h)
    x = H.direct_solution(res["K"], res["f"])
    assert np.abs(x - exact).max() < 1e-12 * max(1.0, abs(ambient))
    assert np.abs(T - exact).max() < 1e-10 * max(1.0, abs(ambient))


@pytest.mark.parametrize("maker", ["q4", "q8"])
def test_robin_terms_of_a_nonlocal_operator(maker):
    """With no first-kind group the nonlocal operator annihilates constants (sum_j grad N_j = 0), so
    (K 1)_i is the convection row sum h int_G N_i: its total is h |G| = h (the edge x = 1 has unit
    length), it is zero off the Robin edge and positive on it; the right-hand side adds h T_a int N_i,
    total h T_a.  The symmetric (upper) storage is read as the full matrix."""
    mesh = meshgen.structured_q4(12, 0.2, seed=13) if maker == "q4" else meshgen.structured_q8(7, 0.2, seed=14)
    h, ambient = 0.8, 1.75
    bcs = [A.boundary(), A.boundary(A.NLF_BC_CONVECTION, heat_transfer=h, ambient=ambient), A.boundary(),
           A.boundary()]
    om = O.OracleMesh(mesh)
    model = configs.bell(0.4, 0.25)
    om.find_neighbours([[0.25, 0.25]])
    f, xD, inner = om.boundary(bcs, 1)
    assert inner.all()
    K = om.assemble(0, [model], [A.thermal_material()], inner)
    before = H.upper_to_full(K.export(0), mesh.n_nodes) @ np.ones(mesh.n_nodes)
    om.convection(K, bcs, inner)
    full = H.upper_to_full(K.export(0), mesh.n_nodes)
    row_sums = full @ np.ones(mesh.n_nodes)
    scale = np.abs(full).max()
    assert np.abs(before).max() < 1e-13 * scale
    on_edge = np.zeros(mesh.n_nodes, dtype=bool)
    on_edge[mesh.boundary_nodes("right")] = True
    assert abs(row_sums.sum() - h) < 1e-13
    assert np.abs(row_sums[~on_edge]).max() < 1e-13 * scale
    assert (row_sums[on_edge] > 0).all()
    # the convection block of a straight two-node edge: the reference integrates linear edges with one
    # Gauss point (vtk_elements_set.hpp:43-51), N_a = N_b = 1/2 there, so every entry is h * length / 4
    # (the exact integral would give length/6 off the diagonal)
    if maker == "q4":
        diff = (full - H.upper_to_full(om.assemble(0, [model], [A.thermal_material()], inner).export(0),
                                       mesh.n_nodes)).toarray()
        right = mesh.boundary_nodes("right")
        ys = mesh.xy[right, 1]
        order = right[np.argsort(ys)]
        for a, b in zip(order[:-1], order[1:]):
            length = abs(mesh.xy[b, 1] - mesh.xy[a, 1])
            assert abs(diff[a, b] - h * length / 4.0) < 1e-15
            assert abs(diff[b, a] - h * length / 4.0) < 1e-15
    assert abs(f.sum() - h * ambient) < 1e-13
    assert np.abs(f[~on_edge]).max() == 0.0
