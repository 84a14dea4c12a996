"""Pins of the oracle's composite obstacles and box-element cylindrical
shells (oracle/gdem_oracle.c: orc_add_cylindrical_shell, composite moves),
each fixed by geometry or kinematics:

* the shell's inner faces lie at the inner radius from its axis, each box's
  local x is radial and local z the axis, and the outer corners of adjacent
  boxes coincide (the outer faces form a regular polygon of inradius
  radius + width: ly / 2 = (radius + width) tan(pi / n));
* the library's own construction (gdem_cylindrical_shell_boxes, host code,
  callable without a GPU) equals the oracle's bit for bit;
* a shell rotating as a composite about its axis moves every box rigidly:
  centres on C + R(omega t)(c0 - C), bases R(omega t) B0 (Rodrigues), each box
  with the composite's angular velocity;
* a composite translating at V carries a member that also spins about its
  own centre: the member's centre moves by V t, its basis turns by its own
  rotation, its velocity is V and its angular velocity its own;
* nested composites compose: a wall in a child of a rotating composite turns
  with it, to second order in omega dt per step (the reference takes a
  nested member's lever from its composite's centre after that composite has
  moved, CompositeObstacle.cpp:167-178, so the member drifts from the rigid
  path by about (omega dt)^2 |lever| per step);
* a sphere resting at the bottom of a drum spinning at omega about its axis
  feels, before any step, the full Coulomb force along omega x (contact
  point - axis): the composed velocity of the box under it."""
import math

import numpy as np
import pytest

from pacific_amd import gdem, workloads
from pacific_amd.gdem import MODEL_HOOKE, MOTION_CONSTANT_ROTATION, MOTION_CONSTANT_TRANSLATION

HOOKE = (1.0e5, 0.5, -1.0, 0.3, 0.0)


def rotation(omega, t):
    w = np.asarray(omega, float)
    a = np.linalg.norm(w) * t
    k = w / np.linalg.norm(w)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * (K @ K)


SHELLS = [
    (workloads.DRUM_BASIS, (0.0, 0.0, 0.0), 0.04, 0.01, 0.001, 24),
    (tuple(rotation((0.3, -0.5, 0.8), 1.1).ravel()), (0.01, -0.02, 0.03), 0.02, 0.05, 0.004, 7),
]


@pytest.mark.parametrize("shell", SHELLS)
def test_shell_geometry(oracle_mod, shell):
    basis, centre, height, radius, width, n = shell
    B0 = np.array(basis).reshape(3, 3)
    c0 = np.array(centre)
    b, c, h = oracle_mod.cylindrical_shell_boxes(basis, centre, height, radius, width, n)
    axis = B0[:, 2]
    np.testing.assert_allclose(h[:, 0], width / 2, rtol=1e-15)
    np.testing.assert_allclose(h[:, 2], height / 2, rtol=1e-15)
    np.testing.assert_allclose(h[:, 1], (radius + width) * math.tan(math.pi / n), rtol=1e-13)
    outer = []
    for i in range(n):
        Bi = b[i].reshape(3, 3)
        np.testing.assert_allclose(Bi @ Bi.T, np.eye(3), atol=1e-15)
        np.testing.assert_allclose(Bi[:, 2], axis, atol=1e-15)          # local z = axis
        rel = c[i] - c0
        radial = rel - np.dot(rel, axis) * axis
        assert abs(np.dot(rel, axis)) <= 1e-15
        np.testing.assert_allclose(Bi[:, 0], radial / np.linalg.norm(radial), atol=1e-14)
        inner = c[i] - h[i, 0] * Bi[:, 0]
        assert np.linalg.norm(inner - c0) == pytest.approx(radius, rel=1e-14)
        outer.append((c[i] + h[i, 0] * Bi[:, 0] - h[i, 1] * Bi[:, 1],
                      c[i] + h[i, 0] * Bi[:, 0] + h[i, 1] * Bi[:, 1]))
    for i in range(n):
        np.testing.assert_allclose(outer[i][1], outer[(i + 1) % n][0], rtol=0, atol=1e-15 * (radius + width) * 10)
    # the library's construction is the same arithmetic
    lb, lc, lh = gdem.cylindrical_shell_boxes(basis, centre, height, radius, width, n)
    for x, y in ((lb, b), (lc, c), (lh, h)):
        np.testing.assert_array_equal(x, y)


def drum_walls(oracle_mod, wl, nsteps, dt):
    o = oracle_mod.Oracle(wl)
    o.prime()
    out = []
    for _ in range(nsteps):
        o.step(1, dt)
        out.append([o.wall(k) for k in range(len(wl["walls"]) + 24)])
    return o, out


def far_sphere(wl):
    wl = dict(wl)
    wl["gravity"] = (0.0, 0.0, 0.0)
    wl["particles"] = workloads._particles(np.array([[0.0, 0.0, 0.0]]), np.zeros((1, 3)),
                                           np.array([1.0e-3]))
    return wl


def test_rotating_shell_moves_rigidly(oracle_mod):
    omega = np.array([0.0, 5.0, 0.0])
    dt = 2.0 ** -10
    wl = far_sphere(workloads.rotating_drum(omega=5.0))
    o0 = oracle_mod.Oracle(wl)
    o0.prime()
    start = [o0.wall(k) for k in range(26)]
    C = o0.composite(0)[0]
    o, traj = drum_walls(oracle_mod, wl, 60, dt)
    for k in (0, 29, 59):
        R = rotation(omega, (k + 1) * dt)
        for w in range(26):
            c, b, v, om = traj[k][w]
            np.testing.assert_allclose(c, C + R @ (start[w][0] - C), rtol=0, atol=1e-15)
            np.testing.assert_allclose(b.reshape(3, 3), R @ start[w][1].reshape(3, 3), rtol=0, atol=1e-13)
            np.testing.assert_array_equal(om, omega)
            prev = start[w][0] if k == 0 else traj[k - 1][w][0]
            np.testing.assert_allclose(v, np.cross(omega, prev - C), rtol=1e-13, atol=1e-18)
    cc, cv, cw = o.composite(0)
    np.testing.assert_allclose(cc, C, atol=1e-18)
    np.testing.assert_array_equal(cw, omega)


def test_translating_composite_with_spinning_member(oracle_mod):
    V = np.array([0.2, -0.1, 0.05])
    w_own = np.array([0.0, 0.0, 3.0])
    dt = 2.0 ** -12
    wl = workloads.sphere_on_plate((MODEL_HOOKE, HOOKE), x0=(0.0, 0.0, 5.0e-3), gravity=0.0, dt=dt)
    wl["composites"] = [(-1, [0])]
    wl["motions"] = [(0, MOTION_CONSTANT_ROTATION, 0.0, 1.0, (*w_own, 0.0, 0.0, 0.0, 1.0))]
    wl["composite_motions"] = [(0, MOTION_CONSTANT_TRANSLATION, 0.0, 1.0, tuple(V))]
    o = oracle_mod.Oracle(wl)
    o.prime()
    c0 = o.wall(0)[0]
    for k in range(1, 41):
        o.step(1, dt)
        c, b, v, om = o.wall(0)
        np.testing.assert_allclose(c, c0 + V * k * dt, rtol=0, atol=1e-17)
        np.testing.assert_allclose(b.reshape(3, 3), rotation(w_own, k * dt), rtol=0, atol=1e-14)
        np.testing.assert_array_equal(v, V)
        np.testing.assert_array_equal(om, w_own)
    np.testing.assert_allclose(o.composite(0)[0], c0 + V * 40 * dt, rtol=0, atol=1e-17)


def test_nested_composites_turn_together(oracle_mod):
    omega = np.array([1.0, 0.5, -2.0])
    dt = 2.0 ** -10
    wl = workloads.sphere_on_plate((MODEL_HOOKE, HOOKE), x0=(0.0, 0.0, 5.0e-3), gravity=0.0, dt=dt)
    wl["walls"] = wl["walls"] + [(workloads.IDENTITY, (0.01, 0.02, -0.004), (0.002, 0.003, 0.001),
                                  workloads.CRUST, 1)]
    wl["composites"] = [(-1, [0]), (0, [1])]    # composite 1 (wall 1) inside composite 0
    wl["composite_motions"] = [(0, MOTION_CONSTANT_ROTATION, 0.0, 1.0, (*omega, 0.0, 0.0, 0.0, 1.0))]
    o = oracle_mod.Oracle(wl)
    o.prime()
    C = o.composite(0)[0]
    c1 = o.wall(1)[0]
    o.step(50, dt)
    R = rotation(omega, 50 * dt)
    c, b, _, om = o.wall(1)
    drift = np.linalg.norm(c - (C + R @ (c1 - C)))
    bound = 50 * (np.linalg.norm(omega) * dt) ** 2 * np.linalg.norm(c1 - C)
    assert drift <= bound
    # a member of composite 0 itself moves rigidly
    np.testing.assert_allclose(o.wall(0)[1].reshape(3, 3), R, rtol=0, atol=1e-13)
    np.testing.assert_allclose(b.reshape(3, 3), R, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(om, omega)
    # composite 1 is a direct member of composite 0: its centre (initially
    # its wall's) moves rigidly
    np.testing.assert_allclose(o.composite(1)[0], C + R @ (c1 - C), rtol=0, atol=1e-15)


def test_sphere_in_spinning_drum_feels_wall_velocity(oracle_mod):
    omega = 5.0
    wl = workloads.rotating_drum(nbox=26, omega=omega)   # a face centred under the axis
    # already turning at the start (a motion starting at t = 0 is not active
    # over [-dt, 0], so the drum would start at rest)
    wl["composite_motions"] = [(0, MOTION_CONSTANT_ROTATION, -1.0, 10.0,
                                (0.0, omega, 0.0, 0.0, 0.0, 0.0, 1.0))]
    m = workloads.DENSITY * 4.0 * math.pi * 1e-9 / 3.0
    delta = m * 9.81 / 1.0e5
    wl["models"] = [(0, 0, MODEL_HOOKE, HOOKE), (0, 1, MODEL_HOOKE, HOOKE)]
    wl["particles"] = workloads._particles(np.array([[0.0, 0.0, -0.01 + 1.0e-3 - delta]]),
                                           np.zeros((1, 3)), np.array([1.0e-3]))
    o = oracle_mod.Oracle(wl)
    o.prime()
    F = o.state()["F"][0]
    # contact point (0, 0, -R): omega y x (0, 0, -R) = (-omega R, 0, 0)
    assert F[0] == pytest.approx(-0.3 * m * 9.81, rel=1e-12)
    assert abs(F[1]) <= 1e-12 * abs(F[0])
