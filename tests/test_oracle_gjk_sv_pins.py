"""Pins of the oracle's signed-volume GJK (closest_points_GJK_SV,
Geometry/src/GJK_SV.cpp:428-524, selected by CollisionDetection Type="GJK_SV",
Base/src/Grains.cpp:660-666).  None of them retypes the sub-algorithm:

* the distance between two polytopes equals the distance from the origin to
  the convex hull of their Minkowski difference, found by brute force over
  every vertex, edge and triangle of that point set (independent of GJK);
* the witness points are mutually closest: each is a point of its own body
  (inside every support plane), on the support plane facing the other body,
  and |pa - pb| is the distance (wrong weights, e.g. ones that do not sum to
  one, move them off the bodies);
* closed forms: two aligned cubes a gap apart, a tetrahedron's apex above a
  cube's face;
* overlapping bodies (the origin inside a tetrahedron of the difference, the
  3-simplex case) give a distance below EPSILON, as the Johnson GJK does;
* a whole convex step with either narrow phase gives the same contacts."""
import itertools
import math

import numpy as np
import pytest

from pacific_amd import workloads
from test_oracle_convex_pins import I3, ONE, rand_rot, shapes_std, small_wl


def _oracle(oracle_mod, sv=True):
    wl = small_wl(shapes_std(), {"x": [[0.01] * 3], "shape": [0]})
    wl["narrow_phase"] = "GJK_SV" if sv else "GJK"
    return oracle_mod.ConvexOracle(wl)


def _verts():
    cube = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * 1e-3
    return {0: cube, 1: workloads.tetrahedron_vertices()}


def _hull_distance(m):
    """Distance from the origin to conv(m).  Every vertex, edge and triangle of
    the point set lies in the hull; the nearest of them to the origin holds the
    hull's closest point c when the origin is outside, which is the case
    exactly when every point m satisfies m.c >= |c|^2 (c's plane separates);
    otherwise the origin is inside and the distance is 0."""
    cands = [m]
    a, b = (np.array(x) for x in zip(*itertools.combinations(m, 2)))
    e = b - a
    t = np.clip(-(a * e).sum(1) / (e * e).sum(1), 0.0, 1.0)
    cands.append(a + t[:, None] * e)
    a, b, c = (np.array(x) for x in zip(*itertools.combinations(m, 3)))
    n = np.cross(b - a, c - a)
    nn = (n * n).sum(1)
    ok = nn > 1e-30
    a, b, c, n, nn = a[ok], b[ok], c[ok], n[ok], nn[ok]
    p = n * ((a * n).sum(1) / nn)[:, None]
    v0, v1, v2 = b - a, c - a, p - a
    d00, d01, d11 = (v0 * v0).sum(1), (v0 * v1).sum(1), (v1 * v1).sum(1)
    d20, d21 = (v2 * v0).sum(1), (v2 * v1).sum(1)
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    inside = (v >= -1e-12) & (w >= -1e-12) & (1.0 - v - w >= -1e-12)
    cands.append(p[inside])
    pts = np.concatenate(cands)
    best = pts[np.argmin((pts * pts).sum(1))]
    d2 = best @ best
    if (m @ best).min() < d2 * (1.0 - 1e-9):
        return 0.0
    return math.sqrt(d2)


def _random_pair(rng, gap_lo=1.5e-3, gap_hi=4e-3):
    sa, sb = (int(k) for k in rng.integers(0, 2, 2))
    Ra, Rb = rand_rot(rng), rand_rot(rng)
    ca = rng.uniform(-1e-4, 1e-4, 3)
    dirn = rng.normal(size=3)
    dirn /= np.linalg.norm(dirn)
    cb = ca + dirn * rng.uniform(gap_lo, gap_hi)
    return sa, Ra, ca, sb, Rb, cb


def test_distance_is_the_minkowski_hull_distance(oracle_mod):
    o = _oracle(oracle_mod)
    verts = _verts()
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(200):
        sa, Ra, ca, sb, Rb, cb = _random_pair(rng)
        va = verts[sa] @ Ra.T + ca
        vb = verts[sb] @ Rb.T + cb
        want = _hull_distance((va[:, None, :] - vb[None, :, :]).reshape(-1, 3))
        d, pa, pb, it = o.closest_points(sa, Ra.reshape(-1), ca, ONE, sb, Rb.reshape(-1), cb, ONE)
        if want < 1e-7:
            assert d < 1e-7
            continue
        assert d == pytest.approx(want, rel=1e-8), (d, want)
        checked += 1
    assert checked > 110


def test_witness_points_are_on_their_bodies_and_mutually_closest(oracle_mod):
    o = _oracle(oracle_mod)
    verts = _verts()
    rng = np.random.default_rng(12)
    dirs = rng.normal(size=(40, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    checked = 0
    for _ in range(300):
        sa, Ra, ca, sb, Rb, cb = _random_pair(rng)
        va = verts[sa] @ Ra.T + ca
        vb = verts[sb] @ Rb.T + cb
        d, pa, pb, it = o.closest_points(sa, Ra.reshape(-1), ca, ONE, sb, Rb.reshape(-1), cb, ONE)
        if d < 1e-7:
            continue
        tol = 1e-9 * 3e-3
        assert np.linalg.norm(pa - pb) == pytest.approx(d, rel=1e-9)
        # inside every support plane of its own body
        assert ((dirs @ pa) - (va @ dirs.T).max(0) <= tol).all()
        assert ((dirs @ pb) - (vb @ dirs.T).max(0) <= tol).all()
        # on the support plane facing the other body
        n = (pa - pb) / np.linalg.norm(pa - pb)
        assert (va @ -n).max() - pa @ -n <= tol
        assert (vb @ n).max() - pb @ n <= tol
        # the symmetric call gives the same distance
        d2, *_ = o.closest_points(sb, Rb.reshape(-1), cb, ONE, sa, Ra.reshape(-1), ca, ONE)
        assert d2 == pytest.approx(d, rel=1e-8)
        checked += 1
    assert checked > 150


def test_known_distances(oracle_mod):
    o = _oracle(oracle_mod)
    for g in (1e-4, 3e-5, 1e-6):
        d, pa, pb, it = o.closest_points(0, I3, [0, 0, 0], ONE, 0, I3,
                                         [2e-3 + g, 0.3e-3, -0.2e-3], ONE)
        assert d == pytest.approx(g, rel=1e-9)
        assert pa[0] == pytest.approx(1e-3, abs=1e-15)
        assert pb[0] == pytest.approx(1e-3 + g, rel=1e-12)
    tv = workloads.tetrahedron_vertices()
    Rx = np.diag([1.0, -1.0, -1.0])
    apex = tv[3, 2]
    for h in (2e-4, 1e-5):
        c = np.array([0.1e-3, -0.2e-3, 1e-3 + apex + h])
        d, pa, pb, it = o.closest_points(1, Rx.reshape(-1), c, ONE, 0, I3, [0, 0, 0], ONE)
        assert d == pytest.approx(h, rel=1e-8)
        np.testing.assert_allclose(pa, [0.1e-3, -0.2e-3, 1e-3 + h], atol=1e-15)


def test_overlapping_bodies_give_zero_distance(oracle_mod):
    sv, jo = _oracle(oracle_mod), _oracle(oracle_mod, sv=False)
    rng = np.random.default_rng(13)
    for _ in range(100):
        sa, Ra, ca, sb, Rb, cb = _random_pair(rng, 0.0, 0.8e-3)
        args = (sa, Ra.reshape(-1), ca, ONE, sb, Rb.reshape(-1), cb, ONE)
        assert sv.closest_points(*args)[0] < 1e-10
        assert jo.closest_points(*args)[0] < 1e-10


def test_convex_step_matches_the_johnson_narrow_phase(oracle_mod):
    """1000 cubes and tetrahedra overlapping their neighbours from the start:
    both narrow phases find the same contacts, their overlaps agree to the GJK
    tolerance and the forces after 20 steps agree closely."""
    out = {}
    for name in ("GJK", "GJK_SV"):
        wl = workloads.convex_box(1000, seed=9, phi=0.18, height=0.015, overlap=4e-5)
        wl["narrow_phase"] = name
        o = oracle_mod.ConvexOracle(wl, log_capacity=20000)
        o.prime()
        c = o.contacts()
        o.step(20, wl["dt"])
        out[name] = (c, o.state())
    (cj, sj), (cs, ss) = out["GJK"], out["GJK_SV"]
    assert len(cj["id0"]) > 60
    kj = sorted(zip(cj["id0"].tolist(), cj["id1"].tolist()))
    ks = sorted(zip(cs["id0"].tolist(), cs["id1"].tolist()))
    assert kj == ks
    oj = dict(zip(zip(cj["id0"].tolist(), cj["id1"].tolist()), cj["overlap"].tolist()))
    for key, ov in zip(zip(cs["id0"].tolist(), cs["id1"].tolist()), cs["overlap"].tolist()):
        assert ov == pytest.approx(oj[key], rel=1e-5, abs=1e-13)
    scale = np.abs(sj["F"]).max()
    assert np.abs(ss["F"] - sj["F"]).max() <= 1e-4 * scale
    assert math.isfinite(scale) and scale > 0
