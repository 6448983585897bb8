"""Oracle pins for the distances other than the Euclidean one: powered_distance with n = 1, integer
n, real n and the maximum norm (search_function.hpp:28-73), and powered_distance_with_rotation
(search_function.hpp:75-90), in the neighbour search (find_neighbours.hpp:152-155) and in the influence
functions with their normalisations (influence_functions_2d.hpp:39-125; read_model.hpp:97-165).

What fixes each part independently of the restatement:
- every influence integrates to 1 over the plane (what the normalisations of
  influence_functions_2d.hpp:44-48, 70-75, 103-109 are for), by a midpoint sum on a grid, for each kind
  and each distance, the rotated ellipse at a point away from the origin;
- the shape: phi(x, y) / phi(x, x) = (1 - rho^p)^q (polynomial), exp(-q rho^p) (exponential) with rho
  the Lp norm of numpy.linalg.norm(z, ord=n);
- the rotated ellipse: the axis-aligned ellipse on the x axis (radii r0, r1), the swapped one on the y
  axis, the circle when r0 = r1, and in general (radial / r0)^2 + (tangential / r1)^2 from the unit
  vector x / |x|; NaN at the origin like the reference's 0 / 0;
- the search returns the brute-force set {e' : distance(c_e, c_e') <= 1} from numpy norms for every
  distance (pairs within 1e-12 of the boundary left out);
- integer and real exponents of the same value give the same lists (n = 1 / 1.0, 2 / 2.0).
"""
import dataclasses
import math

import numpy as np
import pytest

from nonlocal_amd import _abi as A
from nonlocal_amd import meshgen
from oracle import oracle as O

R = (0.7, 0.4)


def _grid(cx, cy, hx, hy, m):
    """Midpoints of an m x m grid over [cx - hx, cx + hx] x [cy - hy, cy + hy] and the cell area."""
    u = (np.arange(m) + 0.5) / m * 2.0 - 1.0
    gx, gy = np.meshgrid(cx + hx * u, cy + hy * u, indexing="ij")
    return np.stack([gx.ravel(), gy.ravel()], axis=1), (2 * hx / m) * (2 * hy / m)


def _integral(model, x, half, m):
    pts, area = _grid(x[0], x[1], half[0], half[1], m)
    return math.fsum(O.influence_points(model, x, pts)) * area


LP = [1, 3, 4, 1.5, "inf"]
CASES = []
for n in LP:
    CASES.append((f"constant_n{n}", dict(influence=A.NLF_INFLUENCE_CONSTANT), n, 2e-3))
    CASES.append((f"poly_p2q1_n{n}", dict(p=2.0, q=1.0), n, 2e-4))
    CASES.append((f"poly_p3q2_n{n}", dict(p=3.0, q=2.0), n, 2e-4))
    CASES.append((f"exp_p2_n{n}", dict(influence=A.NLF_INFLUENCE_EXPONENTIAL, p=2.0, q=0.5), n, 2e-4))
    CASES.append((f"exp_p1_n{n}", dict(influence=A.NLF_INFLUENCE_EXPONENTIAL, p=1.0, q=2.0), n, 2e-4))


@pytest.mark.parametrize("name,kw,n,tol", CASES, ids=[c[0] for c in CASES])
def test_lp_influence_integrates_to_one(name, kw, n, tol):
    model = A.model(0.5, R, n=n, **kw)
    exponential = kw.get("influence") == A.NLF_INFLUENCE_EXPONENTIAL
    # the exponential has no cutoff: its tail beyond 12 radii (1 / q^(1/p) units) is below 1e-12
    scale = 12.0 / (kw["q"] ** (1.0 / kw["p"])) if exponential else 1.0
    x = (0.3, -0.2)
    total = _integral(model, x, (R[0] * scale, R[1] * scale), 1500)
    assert abs(total - 1.0) <= tol, f"{name}: integral {total}"


ROT_CASES = [("constant", dict(influence=A.NLF_INFLUENCE_CONSTANT), 3e-3),
             ("fast_polynomial", dict(p=2.0, q=1.0), 2e-4),
             ("bell", dict(p=2.0, q=2.0), 2e-4),
             ("poly_p3q1", dict(p=3.0, q=1.0), 2e-4),
             ("exponential", dict(influence=A.NLF_INFLUENCE_EXPONENTIAL, p=2.0, q=0.5), 2e-4)]


@pytest.mark.parametrize("name,kw,tol", ROT_CASES, ids=[c[0] for c in ROT_CASES])
def test_rotated_influence_integrates_to_one(name, kw, tol):
    model = A.model(0.5, (0.5, 0.3), distance=A.NLF_DISTANCE_ELLIPSE_WITH_ROTATION, **kw)
    x = (1.3, 0.6)
    half = 0.5 * (12.0 / math.sqrt(kw["q"]) if kw.get("influence") == A.NLF_INFLUENCE_EXPONENTIAL else 1.0)
    total = _integral(model, x, (half, half), 1500)
    assert abs(total - 1.0) <= tol, f"{name}: integral {total}"


def _lp_norm(z, n):
    return np.linalg.norm(z, ord=np.inf if n == "inf" else float(n), axis=1)


@pytest.mark.parametrize("n", LP)
@pytest.mark.parametrize("p,q", [(2.0, 1.0), (3.0, 2.0), (1.0, 1.5)])
def test_lp_polynomial_shape(n, p, q):
    model = A.model(0.5, R, p=p, q=q, n=n)
    x = np.array([0.1, 0.2])
    ys = x + np.random.default_rng(3).uniform(-1.0, 1.0, (4000, 2)) * np.array(R)
    phi = O.influence_points(model, x, ys)
    phi0 = O.influence(model, x, x)
    rho = _lp_norm((x - ys) / np.array(R), n)
    want = np.where(rho < 1.0, np.clip(1.0 - rho ** p, 0.0, None) ** q, 0.0)
    keep = np.abs(rho - 1.0) > 1e-9
    assert np.allclose(phi[keep] / phi0, want[keep], rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("n", LP)
@pytest.mark.parametrize("p,q", [(2.0, 0.5), (1.0, 2.0), (3.0, 1.0)])
def test_lp_exponential_shape(n, p, q):
    model = A.model(0.5, R, influence=A.NLF_INFLUENCE_EXPONENTIAL, p=p, q=q, n=n)
    x = np.array([0.1, 0.2])
    ys = x + np.random.default_rng(4).uniform(-3.0, 3.0, (4000, 2)) * np.array(R)
    phi = O.influence_points(model, x, ys)
    rho = _lp_norm((x - ys) / np.array(R), n)
    assert np.allclose(phi / O.influence(model, x, x), np.exp(-q * rho ** p), rtol=1e-11, atol=1e-300)


def _rot(r):
    return A.distance(A.NLF_DISTANCE_ELLIPSE_WITH_ROTATION), np.asarray(r, dtype=np.float64)


def test_rotated_distance_closed_forms():
    d, r = _rot((0.5, 0.3))
    rng = np.random.default_rng(5)
    ys = rng.uniform(-2.0, 2.0, (2000, 2))
    # on the positive x axis: the axis-aligned ellipse with r0 along x
    a = 1.7
    got = O.distance_points(d, (a, 0.0), r, ys)
    want = ((ys[:, 0] - a) / r[0]) ** 2 + (ys[:, 1] / r[1]) ** 2
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15)
    # on the y axis: r0 along y
    got = O.distance_points(d, (0.0, -a), r, ys)
    want = ((ys[:, 1] + a) / r[0]) ** 2 + (ys[:, 0] / r[1]) ** 2
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15)
    # anywhere: radial and tangential parts of y - x along the unit vector x / |x|
    for x in rng.uniform(-2.0, 2.0, (20, 2)):
        u = x / np.hypot(*x)
        dlt = ys - x
        radial = dlt @ u
        tangential = dlt[:, 0] * u[1] - dlt[:, 1] * u[0]
        want = (radial / r[0]) ** 2 + (tangential / r[1]) ** 2
        assert np.allclose(O.distance_points(d, x, r, ys), want, rtol=1e-12, atol=1e-14)
    # equal radii: the circle
    dc, rc = _rot((0.4, 0.4))
    x = np.array([0.8, -1.1])
    assert np.allclose(O.distance_points(dc, x, rc, ys), ((ys - x) ** 2).sum(axis=1) / 0.16, rtol=1e-13, atol=1e-15)
    # the reference's 0 / 0 at the origin
    assert np.isnan(O.distance_points(d, (0.0, 0.0), r, ys[:3])).all()


def _brute_lists(centres, ranges, radii, dist_fn):
    """{e' in e's group : dist(c_e, c_e') <= 1} per element, pairs within 1e-12 of 1 reported apart."""
    E = centres.shape[0]
    lists, close = [set() for _ in range(E)], set()
    for g, (b, e) in enumerate(ranges):
        if radii[g][0] <= 0:
            continue
        idx = np.arange(b, e)
        for i in idx:
            v = dist_fn(centres[i], centres[idx], np.asarray(radii[g]))
            with np.errstate(invalid="ignore"):
                lists[i] = set(idx[v <= 1.0].tolist())
                close |= {(int(i), int(j)) for j in idx[np.abs(v - 1.0) <= 1e-12]}
    return lists, close


def _np_distance(n):
    if n == "rot":
        def f(x, ys, r):
            nx = np.hypot(*x)
            u = x / nx
            dlt = ys - x
            with np.errstate(invalid="ignore", divide="ignore"):
                return ((dlt @ u) / r[0]) ** 2 + ((dlt[:, 0] * u[1] - dlt[:, 1] * u[0]) / r[1]) ** 2 if nx > 0 \
                    else np.full(len(ys), np.nan)
        return f
    return lambda x, ys, r: _lp_norm((x - ys) / r, n) ** (1.0 if n == "inf" else float(n))


@pytest.mark.parametrize("n", [1, 3, 1.5, "inf", "rot"])
@pytest.mark.parametrize("shift", [0.0, -0.5])
def test_search_matches_brute_force(n, shift):
    mesh = meshgen.structured_q4(17, jitter=0.2, seed=2)
    mesh = dataclasses.replace(mesh, xy=mesh.xy + shift)
    om = O.OracleMesh(mesh)
    d = A.distance(A.NLF_DISTANCE_ELLIPSE_WITH_ROTATION) if n == "rot" else A.distance(A.NLF_DISTANCE_LP, n)
    radii = [[0.21, 0.13]]
    off, ids, expanded = om.find_neighbours(radii, strategy=A.NLF_EXPAND_MAX, distances=[d])
    centres = om.centres()
    ranges = [(int(mesh.group_begin[g]), int(mesh.group_begin[g + 1])) for g in range(len(mesh.group_begin) - 1)]
    lists, close = _brute_lists(centres, ranges, expanded, _np_distance(n))
    mismatched = 0
    for e in range(mesh.n_elems):
        got = ids[off[e]:off[e + 1]].tolist()
        assert len(got) == len(set(got))
        diff = set(got) ^ lists[e]
        assert all((e, j) in close for j in diff), f"element {e}: {sorted(diff)[:5]}"
        mismatched += len(diff)
    assert mismatched <= 2
    assert ids.size > 5 * mesh.n_elems


@pytest.mark.parametrize("a,b", [(1, 1.0), (2, 2.0)])
def test_integer_and_real_exponents_agree(a, b):
    mesh = meshgen.structured_q4(15, jitter=0.15, seed=1)
    om = O.OracleMesh(mesh)
    ra = om.find_neighbours([[0.17, 0.23]], distances=[A.distance(A.NLF_DISTANCE_LP, a)])
    rb = om.find_neighbours([[0.17, 0.23]], distances=[A.distance(A.NLF_DISTANCE_LP, b)])
    assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])
    # and the influences: integer p/n and the real p/n
    ma, mb = A.model(0.5, R, p=4.0, q=2.0, n=a), A.model(0.5, R, p=4.0, q=2.0, n=b)
    x = np.array([0.2, 0.1])
    ys = x + np.random.default_rng(6).uniform(-1, 1, (500, 2)) * np.array(R)
    pa, pb = O.influence_points(ma, x, ys), O.influence_points(mb, x, ys)
    # power by squaring and pow differ in the last bits only
    assert np.abs(pa - pb).max() <= 1e-14 * np.abs(pa).max()


def test_rotated_origin_element_has_no_neighbours():
    # an element centred at the origin: the rotated distance is 0 / 0 there, so its list is empty
    mesh = meshgen.structured_q4(9)
    mesh = dataclasses.replace(mesh, xy=mesh.xy - 0.5)
    om = O.OracleMesh(mesh)
    centres = om.centres()
    e0 = int(np.argmin(np.hypot(centres[:, 0], centres[:, 1])))
    assert np.all(centres[e0] == 0.0)
    off, ids, _ = om.find_neighbours([[0.3, 0.2]], distances=[A.distance(A.NLF_DISTANCE_ELLIPSE_WITH_ROTATION)])
    assert off[e0 + 1] == off[e0]
    assert all(off[e + 1] > off[e] for e in range(mesh.n_elems) if e != e0)


def test_invalid_exponents_raise():
    with pytest.raises(ValueError):
        A.distance_exponent(0)
    with pytest.raises(ValueError):
        A.distance_exponent(-1.5)
    with pytest.raises(ValueError):
        A.distance_exponent("two")
    bad = A.Distance()
    bad.kind, bad.n, bad.n_real = A.NLF_DISTANCE_LP, -3, 0.0
    om = O.OracleMesh(meshgen.structured_q4(4))
    with pytest.raises(RuntimeError):
        om.find_neighbours([[0.3, 0.3]], distances=[bad])
