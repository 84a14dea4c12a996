"""Pins of the oracle's Rectangle obstacles (orc_add_rectangle_wall: the
reference's ClosestPointRECTANGLE, particle first) and of the
Rectangle-element cylindrical shell, each fixed by geometry or statics:

* a sphere resting on a horizontal rectangle at the static overlap m g / kn
  carries its weight: contact point midway between its lowest point and that
  point's projection on the plane, overlap vector (0, 0, delta);
* the same from below (the normal turns towards the sphere), and through a
  rotated sphere: the sphere's support point goes through its basis and back
  (B (B^T d)), so a turned sphere feels the unrotated one's force to
  rounding (a transposed basis would move the contact point by about r);
* a sphere whose centre projects outside the rectangle feels nothing even
  though it overlaps the plane within its circumscribed sphere (a box of the
  same face would touch it), and one just inside feels the full force;
* no contact beyond the circumscribed-sphere test or on the plane's far side;
* the Rectangle-element shell's rectangles lie at the inner radius, their
  normals point at the axis and their local x runs along it; the library's
  construction equals the oracle's bit for bit;
* a box of colliding spheres closed by six rectangles conserves momentum
  exactly as a box of walls does: each wall takes the reaction of its
  contacts (loads balance the particles' contact forces)."""
import math

import numpy as np
import pytest

from pacific_amd import gdem, workloads
from pacific_amd.gdem import MODEL_HOOKE
from test_oracle_composite_pins import rotation

HOOKE = (1.0e5, 0.5, -1.0, 0.3, 0.0)
R = 1.0e-3
M = workloads.DENSITY * 4.0 * math.pi * R ** 3 / 3.0
G = 9.81


def plate(z, x0=(0.0, 0.0), q=(0.0, 0.0, 0.0, 1.0), basis=workloads.IDENTITY, gravity=-G,
          size=(0.02, 0.02)):
    wl = workloads.sphere_on_plate((MODEL_HOOKE, HOOKE), x0=(x0[0], x0[1], z), gravity=gravity)
    wl["walls"] = []
    wl["rectangles"] = [(basis, (0.0, 0.0, 0.0), size[0], size[1], workloads.CRUST, 1)]
    wl["origin"] = (-0.02, -0.02, -0.005)
    wl["size"] = (0.04, 0.04, 0.01)
    wl["particles"]["q"] = np.array([q], float)
    return wl


def primed(oracle_mod, wl, log=True):
    o = oracle_mod.Oracle(wl, log_capacity=16) if log else oracle_mod.Oracle(wl)
    o.prime()
    return o


def test_sphere_resting_on_rectangle(oracle_mod):
    delta = M * G / HOOKE[0]
    o = primed(oracle_mod, plate(R - delta, x0=(0.003, -0.002)))
    F = o.state()["F"][0]
    assert abs(F[2]) <= 1e-9 * M * G          # the overlap is held as r - z
    np.testing.assert_allclose(F[:2], 0.0, atol=1e-15 * M * G)
    c = o.contacts()
    assert len(c["id0"]) == 1 and c["id1"][0] == -1
    np.testing.assert_allclose(c["point"][0], [0.003, -0.002, -delta / 2], rtol=0, atol=1e-18)
    np.testing.assert_allclose(c["ovec"][0], [0.0, 0.0, delta], rtol=1e-6, atol=1e-20)
    assert c["overlap"][0] == pytest.approx(-delta, rel=1e-6)
    # the wall carries the weight
    f, _ = o.wall_loads()
    assert f[0][2] == pytest.approx(-M * G, rel=1e-9)


def test_from_below_and_rotated_sphere(oracle_mod):
    delta = M * G / HOOKE[0]
    below = primed(oracle_mod, plate(-(R - delta), gravity=G))
    assert below.state()["F"][0][2] == pytest.approx(0.0, abs=1e-9 * M * G)
    assert below.contacts()["ovec"][0][2] < 0.0
    ref = primed(oracle_mod, plate(R - 0.5 * delta, x0=(0.001, 0.002))).state()["F"][0]
    for k, axis in enumerate(((1.0, 2.0, 3.0), (0.0, 1.0, 0.0), (-2.0, 0.5, 1.0))):
        a = np.asarray(axis) / np.linalg.norm(axis)
        th = 0.7 + k
        q = (*(math.sin(th / 2) * a), math.cos(th / 2))
        F = primed(oracle_mod, plate(R - 0.5 * delta, x0=(0.001, 0.002), q=q)).state()["F"][0]
        np.testing.assert_allclose(F, ref, rtol=1e-9, atol=1e-12 * M * G)


def test_projection_must_fall_inside(oracle_mod):
    delta = M * G / HOOKE[0]
    edge = 0.01                                         # half side of the 2 cm plate
    # centre 0.3 r beyond the edge and 0.9 r above the plane: the lowest
    # point projects outside, so no contact
    out = primed(oracle_mod, plate(0.9 * R, x0=(edge + 0.3 * R, 0.0)))
    assert len(out.contacts()["id0"]) == 0
    assert out.state()["F"][0][2] == pytest.approx(-M * G, rel=1e-15)
    # a box wall of the same top face touches that sphere at its edge
    # (distance sqrt(0.09 + 0.81) r < r)
    wl = plate(0.9 * R, x0=(edge + 0.3 * R, 0.0))
    wl["rectangles"] = []
    wl["walls"] = [(workloads.IDENTITY, (0.0, 0.0, -1.0e-3), (edge, edge, 1.0e-3), workloads.CRUST, 1)]
    assert len(primed(oracle_mod, wl).contacts()["id0"]) == 1
    # just inside: the full force
    inside = primed(oracle_mod, plate(R - delta, x0=(edge - 1.0e-6, 0.0)))
    assert len(inside.contacts()["id0"]) == 1
    assert abs(inside.state()["F"][0][2]) <= 1e-9 * M * G


def test_no_contact_cases(oracle_mod):
    # above the plane: no contact
    assert len(primed(oracle_mod, plate(R + 1.0e-6)).contacts()["id0"]) == 0
    # a small rectangle far to the side of an overlapping plane position:
    # the circumscribed-sphere test rejects it
    wl = plate(R - 1.0e-6, x0=(0.0, 0.0), size=(1.0e-4, 1.0e-4))
    wl["rectangles"][0] = (workloads.IDENTITY, (0.004, 0.0, 0.0), 1.0e-4, 1.0e-4,
                           workloads.CRUST, 1)
    assert len(primed(oracle_mod, wl).contacts()["id0"]) == 0


@pytest.mark.parametrize("shell", [
    (workloads.DRUM_BASIS, (0.0, 0.0, 0.0), 0.04, 0.01, 0.001, 24),
    (tuple(rotation((0.3, -0.5, 0.8), 1.1).ravel()), (0.01, -0.02, 0.03), 0.02, 0.05, 0.004, 7),
])
def test_rectangle_shell_geometry(oracle_mod, shell):
    basis, centre, height, radius, width, n = shell
    B0 = np.array(basis).reshape(3, 3)
    c0 = np.array(centre)
    axis = B0[:, 2]
    b, c, h = oracle_mod.cylindrical_shell_rectangles(basis, centre, height, radius, width, n)
    delta = 2 * math.pi / n
    ly = 2 * (radius + width) * (math.sin(delta) - (1 - math.cos(delta)) / math.tan(delta))
    np.testing.assert_allclose(h[:, 0], height / 2, rtol=1e-15)
    np.testing.assert_allclose(h[:, 1], ly / 2, rtol=1e-13)
    np.testing.assert_array_equal(h[:, 2], 0.0)
    for i in range(n):
        Bi = b[i].reshape(3, 3)
        np.testing.assert_allclose(Bi @ Bi.T, np.eye(3), atol=1e-15)
        np.testing.assert_allclose(Bi[:, 0], axis, atol=1e-15)           # local x = axis
        rel = c[i] - c0
        assert abs(np.dot(rel, axis)) <= 1e-15
        assert np.linalg.norm(rel) == pytest.approx(radius, rel=1e-14)   # at the inner radius
        np.testing.assert_allclose(Bi[:, 2], -rel / np.linalg.norm(rel), atol=1e-14)  # towards the axis
        ang = (0.5 + i) * delta                                          # at angle (i + 1/2) delta
        np.testing.assert_allclose(rel, B0 @ (radius * np.array([math.cos(ang), math.sin(ang), 0.0])),
                                   rtol=0, atol=1e-15 * radius)
    lb, lc, lh = gdem.cylindrical_shell_rectangles(basis, centre, height, radius, width, n)
    for x, y in ((lb, b), (lc, c), (lh, h)):
        np.testing.assert_array_equal(x, y)


def test_momentum_balance_with_rectangles(oracle_mod):
    from test_oracle_wall_load_pins import _balances
    wl = workloads.rectangle_box(1000, seed=1)
    wl["particles"]["radius"] = wl["particles"]["radius"] * 1.08
    wl["cell_size"] = 2.0 * wl["particles"]["radius"].max()
    o = oracle_mod.Oracle(wl)
    o.prime()
    centres = [r[1] for r in wl["rectangles"]]
    wl_b = dict(wl, walls=[(r[0], r[1], None, None, None) for r in wl["rectangles"]])
    for _ in range(3):
        o.step(20, wl["dt"])
        f, t = o.wall_loads()
        assert np.abs(f).sum() > 0.0
        lin, ang, sf, stq = _balances(wl_b, o.state(), f, t)
        assert np.abs(lin).max() <= 1e-13 * sf
        assert np.abs(ang).max() <= 1e-13 * stq
    assert len(centres) == 6


def test_cube_on_rectangle_touches_at_one_corner(oracle_mod):
    """A convex particle meets a rectangle at its support point alone: a cube
    lying flat with its lower face 2e-6 below the plane has one contact, at
    the corner the box support picks for (0, 0, -1) (Box::support,
    Box.cpp:307-316: +e for a zero component), midway between that corner
    and its projection."""
    e = workloads.CUBE_HALF[0]
    depth = 2.0e-6
    x = np.array([[0.001, -0.002, e - depth]])
    parts = {"x": x, "v": np.zeros((1, 3)), "w": np.zeros((1, 3)),
             "q": np.array([[0.0, 0.0, 0.0, 1.0]]), "basis": np.array([workloads.IDENTITY]),
             "shape": np.zeros(1, np.int32), "density": np.full(1, workloads.DENSITY),
             "crust": np.full(1, workloads.CRUST), "material": np.zeros(1, np.int32)}
    wl = {"origin": (-0.02, -0.02, -0.005), "size": (0.04, 0.04, 0.02), "periodic": (0, 0, 0),
          "cell_size": 4.0 * e, "gravity": (0.0, 0.0, 0.0), "dt": 1.0e-6,
          "models": [(0, 0, MODEL_HOOKE, HOOKE), (0, 1, MODEL_HOOKE, HOOKE)],
          "walls": [], "shapes": [("box", workloads.CUBE_HALF)], "particles": parts,
          "rectangles": [(workloads.IDENTITY, (0.0, 0.0, 0.0), 0.02, 0.02, workloads.CRUST, 1)]}
    o = oracle_mod.ConvexOracle(wl, log_capacity=16)
    o.prime()
    c = o.contacts()
    assert len(c["id0"]) == 1 and c["id1"][0] == -1
    corner = x[0] + np.array([e, e, -e])
    np.testing.assert_allclose(c["point"][0], [corner[0], corner[1], -depth / 2], rtol=0,
                               atol=1e-18)
    np.testing.assert_allclose(c["ovec"][0], [0.0, 0.0, depth], rtol=1e-9, atol=1e-21)
    # the single off-centre contact turns the cube: torque (corner - x) x F
    F, T = o.state()["F"][0], o.state()["T"][0]
    np.testing.assert_allclose(T, np.cross(c["point"][0] - x[0], F), rtol=1e-12, atol=1e-20)
    assert np.abs(T[:2]).max() > 0.0
