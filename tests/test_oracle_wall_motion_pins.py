"""Pins of the oracle's imposed wall motions (oracle/gdem_oracle.c,
orc_move_walls and the wall velocity in the contact law), each fixed by
kinematics or mechanics rather than by the oracle's own formulas:

* a constant translation moves the wall by V (t - t_start) and gives it
  velocity V while active; the velocity falls to zero the step after the
  motion ends (the reset of AllComponents::Move), and the wall then stays;
* a sinusoidal translation integrates to (A P / 2 pi)(cos phi -
  cos(2 pi t / P + phi)) along the unit direction, its velocity is the
  derivative of that path, and the phase is read in degrees;
* a crenel translation integrates to a triangle wave;
* a multi-sine along x reproduces the sine translation bit for bit, and along
  (1, 1, 0)/sqrt(2) carries the reference's squared direction factor;
* a constant rotation turns the basis by R(omega t) (Rodrigues), about the
  wall's centre or, with a given centre C, moves the centre on
  C + R(omega t)(c0 - C);
* motions of one wall add up, in the order the reference keeps them (start
  time, a later one of equal start ahead of earlier ones);
* a sphere at rest on a conveyor plate moving at V ends rolling at 2 V / 7
  (V / (1 + m a^2 / I)); on a plate spinning at Omega it first feels the
  full Coulomb force along omega x lever, then rolls on a circle at 2 Omega / 7
  (the turntable result), which needs the plate's angular velocity in the
  contact law with the right lever."""
import math

import numpy as np
import pytest

from pacific_amd import workloads
from pacific_amd.gdem import (MODEL_HOOKE, MODEL_HOOKE_MEMORY, MOTION_CONSTANT_ROTATION,
                              MOTION_CONSTANT_TRANSLATION, MOTION_CRENEL_TRANSLATION,
                              MOTION_MULTISIN_TRANSLATION, MOTION_SIN_TRANSLATION)

HOOKE = (1.0e5, 0.5, -1.0, 0.3, 0.0)
DT = 2.0 ** -16          # exact in binary: step times k dt are exact
C0 = np.array([0.0, 0.0, -1.0e-3])


def rotation(omega, t):
    """Rodrigues: rotation by |omega| t about omega."""
    w = np.asarray(omega, float)
    a = np.linalg.norm(w) * t
    k = w / np.linalg.norm(w)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * (K @ K)


def walls_over(oracle_mod, motions, nsteps, dt=DT):
    """Plate trajectory (centre, basis, velocity, angular velocity after each
    step), the sphere far above it and no gravity."""
    wl = workloads.sphere_on_plate((MODEL_HOOKE, HOOKE), x0=(0.0, 0.0, 5.0e-3), gravity=0.0,
                                   dt=dt, motions=motions)
    o = oracle_mod.Oracle(wl)
    o.prime()
    out = []
    for _ in range(nsteps):
        o.step(1, dt)
        out.append(o.wall(0))
    return [np.array([w[i] for w in out]) for i in range(4)]


def test_constant_translation(oracle_mod):
    V = np.array([0.3, -0.2, 0.1])
    t0, t1 = 2.5 * DT, 40.5 * DT
    c, b, v, w = walls_over(oracle_mod, [(MOTION_CONSTANT_TRANSLATION, t0, t1, V)], 45)
    t = np.arange(1, 46) * DT
    expect = C0 + np.clip(t - t0, 0.0, t1 - t0)[:, None] * V
    np.testing.assert_allclose(c, expect, rtol=0, atol=1e-17)
    np.testing.assert_array_equal(v[:2], 0.0)          # not started
    np.testing.assert_array_equal(v[2:40], np.tile(V, (38, 1)))
    np.testing.assert_array_equal(v[40:], 0.0)         # completed at t > t1
    np.testing.assert_array_equal(b, np.tile(np.eye(3).ravel(), (45, 1)))
    np.testing.assert_array_equal(w, 0.0)


def test_velocity_reset_after_motion_ends(oracle_mod):
    """Active over [5 dt, 10 dt]: at step 11 nothing is active, but the motion
    was active the step before, so the kinematics are reset."""
    V = np.array([0.0, 0.0, 0.2])
    c, _, v, _ = walls_over(oracle_mod, [(MOTION_CONSTANT_TRANSLATION, 5 * DT, 10 * DT, V)], 14)
    np.testing.assert_array_equal(v[:5], 0.0)
    np.testing.assert_array_equal(v[5:10], np.tile(V, (5, 1)))
    np.testing.assert_array_equal(v[10:], 0.0)
    np.testing.assert_allclose(c[-1], C0 + 5 * DT * V, rtol=0, atol=1e-17)
    np.testing.assert_array_equal(c[9:], np.tile(c[9], (5, 1)))


def test_sin_translation(oracle_mod):
    A, P, phi = 0.2, 40 * DT, 30.0
    d = np.array([1.0, 2.0, 2.0])
    u = d / 3.0
    t0 = 3 * DT
    n = 160
    c, _, v, _ = walls_over(oracle_mod, [(MOTION_SIN_TRANSLATION, t0, 1.0, (*d, A, P, phi))], n)
    t = np.arange(1, n + 1) * DT
    ph = math.radians(phi)
    s = np.where(t > t0, (A * P / (2 * math.pi)) * (math.cos(ph) - np.cos(2 * math.pi * (t - t0) / P + ph)), 0.0)
    np.testing.assert_allclose(c, C0 + s[:, None] * u, rtol=0, atol=1e-17)
    # velocity = derivative of the path (central differences, O((2 pi dt / P)^2))
    fd = (c[2:] - c[:-2]) / (2 * DT)
    k = np.nonzero(t[1:-1] > t0 + DT)[0] + 1
    np.testing.assert_allclose(v[k], fd[k - 1], rtol=0, atol=5e-3 * A)
    # first active step: A sin(2 pi dt / P + 30 deg), the phase read in degrees
    np.testing.assert_allclose(v[3], A * math.sin(2 * math.pi * DT / P + ph) * u, rtol=1e-14)


def test_crenel_translation(oracle_mod):
    A, P = 0.05, 10 * DT
    c, _, v, _ = walls_over(oracle_mod, [(MOTION_CRENEL_TRANSLATION, 0.0, 1.0, (0, 0, 2.0, A, P))], 40)
    k = np.arange(1, 41)
    tri = np.where((k // 10) % 2 == 0, k % 10, 10 - k % 10) * DT * A
    np.testing.assert_allclose(c[:, 2] - C0[2], tri, rtol=0, atol=1e-17)
    np.testing.assert_array_equal(v[:, 2], np.where((k // 10) % 2 == 0, A, -A))


def test_multisin_reduces_to_sin(oracle_mod):
    A, P, phi = 0.1, 24 * DT, 45.0
    sin_ = walls_over(oracle_mod, [(MOTION_SIN_TRANSLATION, 0.0, 1.0, (1, 0, 0, A, P, phi))], 60)
    ms = walls_over(oracle_mod, [(MOTION_MULTISIN_TRANSLATION, 0.0, 1.0,
                                  (1, 0, 0, A, 7.0, 9.0, P, 3 * DT, 5 * DT, phi, 10.0, 20.0))], 60)
    for a, b in zip(sin_, ms):
        np.testing.assert_array_equal(a, b)
    # along (1, 1, 0)/sqrt(2) each component carries the direction twice
    # (amplitude stored times the direction, ObstacleImposedVelocity.cpp:249-254,
    # and the velocity multiplied by it again): half the sine's amplitude
    diag = walls_over(oracle_mod, [(MOTION_MULTISIN_TRANSLATION, 0.0, 1.0,
                                    (1, 1, 0, A, A, 0.0, P, P, P, phi, phi, 0.0))], 60)
    np.testing.assert_allclose(diag[0][:, 0] - C0[0], 0.5 * (sin_[0][:, 0] - C0[0]), rtol=1e-12,
                               atol=1e-19)
    np.testing.assert_allclose(diag[0][:, 1], diag[0][:, 0], rtol=0, atol=0)
    np.testing.assert_allclose(diag[2][:, 0], 0.5 * sin_[2][:, 0], rtol=1e-12, atol=1e-15)


def test_rotation_about_centre(oracle_mod):
    omega = np.array([0.3, -1.2, 2.0])
    dt = 2.0 ** -10
    n = 300
    c, b, v, w = walls_over(oracle_mod, [(MOTION_CONSTANT_ROTATION, 0.0, 10.0,
                                          (*omega, 0.0, 0.0, 0.0, 1.0))], n, dt)
    for k in (0, 99, 299):
        np.testing.assert_allclose(b[k].reshape(3, 3), rotation(omega, (k + 1) * dt), rtol=0, atol=1e-13)
    np.testing.assert_array_equal(c, np.tile(C0, (n, 1)))
    np.testing.assert_array_equal(w, np.tile(omega, (n, 1)))
    np.testing.assert_array_equal(v, 0.0)


def test_rotation_about_a_point(oracle_mod):
    omega = np.array([0.5, 0.2, 3.0])
    C = np.array([0.01, -0.02, 0.0])
    dt = 2.0 ** -10
    n = 300
    c, b, v, w = walls_over(oracle_mod, [(MOTION_CONSTANT_ROTATION, 0.0, 10.0,
                                          (*omega, *C, 0.0))], n, dt)
    for k in (0, 149, 299):
        R = rotation(omega, (k + 1) * dt)
        np.testing.assert_allclose(c[k], C + R @ (C0 - C), rtol=0, atol=1e-15)
        np.testing.assert_allclose(b[k].reshape(3, 3), R, rtol=0, atol=1e-13)
    # the velocity is the rate of the centre's motion to first order in dt
    fd = (c[1:] - c[:-1]) / dt
    r = np.linalg.norm(np.cross(omega, C0 - C))
    assert np.abs(v[1:] - fd).max() <= 0.6 * np.linalg.norm(omega) * r * np.linalg.norm(omega) * dt
    np.testing.assert_array_equal(w, np.tile(omega, (n, 1)))


def test_motions_add_in_reference_order(oracle_mod):
    V1, V2 = np.array([0.1, 0.0, 0.0]), np.array([0.0, 0.05, 0.02])
    c, _, v, _ = walls_over(oracle_mod, [(MOTION_CONSTANT_TRANSLATION, 0.0, 20 * DT, V1),
                                         (MOTION_CONSTANT_TRANSLATION, 10 * DT, 30 * DT, V2)], 32)
    np.testing.assert_array_equal(v[:10], np.tile(V1, (10, 1)))
    np.testing.assert_array_equal(v[10:20], np.tile(V1 + V2, (10, 1)))
    np.testing.assert_array_equal(v[20:30], np.tile(V2, (10, 1)))
    np.testing.assert_allclose(c[-1], C0 + 20 * DT * V1 + 20 * DT * V2, rtol=0, atol=1e-17)
    # three motions of equal start added as a, b, c are summed c, b, a
    # (ObstacleKinematicsVelocity::append): (1e-16 + 1e-16) + 1 rounds up,
    # (1 + 1e-16) + 1e-16 does not
    _, _, v3, _ = walls_over(oracle_mod, [(MOTION_CONSTANT_TRANSLATION, 0.0, 1.0, (1.0, 0, 0)),
                                          (MOTION_CONSTANT_TRANSLATION, 0.0, 1.0, (1e-16, 0, 0)),
                                          (MOTION_CONSTANT_TRANSLATION, 0.0, 1.0, (1e-16, 0, 0))], 2)
    assert v3[-1, 0] == 1.0000000000000002


def test_restricted_directions(oracle_mod):
    """Obstacle::setRestrictedGeomDirMotion: a wall translating at V with x
    restricted keeps its x while y and z move, and keeps the full velocity;
    a composite restricted along z keeps its centre's z, but its members get
    the full composed translation (the reference restricts only the
    composite's own centre, CompositeObstacle.cpp:153-160)."""
    V = np.array([0.3, -0.2, 0.1])
    wl = workloads.sphere_on_plate((MODEL_HOOKE, HOOKE), x0=(0.0, 0.0, 5.0e-3), gravity=0.0, dt=DT,
                                   motions=[(MOTION_CONSTANT_TRANSLATION, 0.0, 1.0, tuple(V))])
    wl["restrictions"] = [(0, -1, 1)]
    o = oracle_mod.Oracle(wl)
    o.prime()
    o.step(20, DT)
    c, _, v, _ = o.wall(0)
    np.testing.assert_allclose(c, C0 + np.array([0.0, 1.0, 1.0]) * V * 20 * DT, rtol=0, atol=1e-17)
    np.testing.assert_array_equal(v, V)
    wl = workloads.sphere_on_plate((MODEL_HOOKE, HOOKE), x0=(0.0, 0.0, 5.0e-3), gravity=0.0, dt=DT)
    wl["composites"] = [(-1, [0])]
    wl["composite_motions"] = [(0, MOTION_CONSTANT_TRANSLATION, 0.0, 1.0, tuple(V))]
    wl["restrictions"] = [(-1, 0, 4)]
    o = oracle_mod.Oracle(wl)
    o.prime()
    c0 = o.composite(0)[0]
    o.step(20, DT)
    np.testing.assert_allclose(o.composite(0)[0], c0 + np.array([1.0, 1.0, 0.0]) * V * 20 * DT,
                               rtol=0, atol=1e-17)
    np.testing.assert_allclose(o.wall(0)[0], C0 + V * 20 * DT, rtol=0, atol=1e-17)


def _mass_inertia(r=1.0e-3):
    m = workloads.DENSITY * 4.0 * math.pi * r ** 3 / 3.0
    return m, 0.4 * m * r * r


@pytest.mark.parametrize("kind,prm", [(MODEL_HOOKE, HOOKE),
                                      (MODEL_HOOKE_MEMORY, (1.0e5, 0.5, -1.0, 0.3, 1.0e4, 0.0, 0.0, 0.0))])
def test_conveyor_two_sevenths(oracle_mod, kind, prm):
    V = 0.1
    dt = 5.0e-6
    wl = workloads.sphere_on_plate((kind, prm), x0=(-0.02, 0.0, None), dt=dt,
                                   motions=[(MOTION_CONSTANT_TRANSLATION, -1.0, 1.0, (V, 0.0, 0.0))])
    o = oracle_mod.Oracle(wl)
    o.prime()
    o.step(8000, dt)  # 0.04 s; slipping ends after 2 V / (7 mu g) = 0.0097 s
    st = o.state()
    m, inertia = _mass_inertia()
    a = 1.0e-3
    assert st["v"][0, 0] == pytest.approx(V / (1.0 + m * a * a / inertia), rel=1e-4)
    # rolling on the belt: contact point moves with it, v - a w_y = V
    assert st["v"][0, 0] - a * st["w"][0, 1] == pytest.approx(V, rel=1e-4)


def test_spinning_plate_force_and_turntable_orbit(oracle_mod):
    Omega = 10.0
    d = 0.01
    dt = 5.0e-6
    wl = workloads.sphere_on_plate((MODEL_HOOKE, HOOKE), x0=(d, 0.0, None), dt=dt,
                                   motions=[(MOTION_CONSTANT_ROTATION, -1.0, 10.0,
                                             (0.0, 0.0, Omega, 0.0, 0.0, 0.0, 1.0))])
    o = oracle_mod.Oracle(wl)
    o.prime()
    m, _ = _mass_inertia()
    # initial velocities set before the first force: the plate moves at
    # Omega z x (d, 0, 1 mm) = (0, Omega d, 0) under the sphere; full slip
    F = o.state()["F"][0]
    assert F[1] == pytest.approx(0.3 * m * 9.81, rel=1e-12)
    assert abs(F[0]) <= 1e-12 * F[1]
    # rolling on a turntable: the centre circles at 2 Omega / 7
    o.step(10000, dt)  # 0.05 s, past the slip phase
    ang = []
    for _ in range(50):
        o.step(2000, dt)
        vx, vy = o.state()["v"][0, :2]
        ang.append(math.atan2(vy, vx))
    ang = np.unwrap(ang)
    rate = np.polyfit(np.arange(50) * 2000 * dt, ang, 1)[0]
    assert rate == pytest.approx(2.0 * Omega / 7.0, rel=2e-3)
