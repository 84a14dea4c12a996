"""Pins of the oracle's memory contact models (HookeMemory / HertzMemory,
oracle/gdem_oracle.c memory_force, contact maps): each check is fixed by
mathematics or mechanics, not by retyping the oracle's formulas.

* the contact-plane rotation equals Rodrigues' rotation about u x v;
* a sphere resting on a wall, pushed tangentially below 1 mm/s (spring
  regime of the tangent definition), oscillates with the period of the
  tangential spring acting on the contact point, 2 pi sqrt(m_eff / kt) with
  m_eff = 1 / (1/m + a^2/I) (translation and rotation both respond);
* a sphere launched sliding without spin ends rolling at v0 / (1 + I/(m a^2))
  (5/7 v0): friction acts at the contact point, so the angular momentum about
  it is conserved, whatever mu and kt;
* rolling friction (mur) decelerates a rolling sphere at mur R* g a / (r^2 + I/m)
  (R* = 1/(1/r + 1/R_wall)) once the spring torque saturates;
* each contact is remembered by both partners with opposite signs, and the
  memory of a contact that ended is erased (as Component::updateContactMap)."""
import math

import numpy as np
import pytest

from pacific_amd import workloads
from pacific_amd.gdem import MODEL_HOOKE_MEMORY

pytestmark = pytest.mark.filterwarnings("ignore")


def rodrigues(u, v, x):
    u = np.asarray(u, float) / np.linalg.norm(u)
    v = np.asarray(v, float) / np.linalg.norm(v)
    k = np.cross(u, v)
    s, c = np.linalg.norm(k), float(np.dot(u, v))
    k = k / s
    return x * c + np.cross(k, x) * s + k * np.dot(k, x) * (1 - c)


def test_rotation_between_two_normals(oracle_mod):
    rng = np.random.default_rng(4)
    for _ in range(200):
        u, v, x = rng.normal(size=(3, 3))
        got = oracle_mod.rotate_two_vectors(u, v, x)
        np.testing.assert_allclose(got, rodrigues(u, v, x), rtol=0, atol=1e-13 * np.linalg.norm(x))
        # u is turned onto the direction of v
        gu = oracle_mod.rotate_two_vectors(u, v, u)
        np.testing.assert_allclose(gu / np.linalg.norm(gu), v / np.linalg.norm(v), atol=1e-14)
    # opposite vectors: a half turn about an axis normal to u
    u = np.array([0.3, -0.2, 0.9])
    got = oracle_mod.rotate_two_vectors(u, -u, u)
    np.testing.assert_allclose(got, -u, atol=1e-15)


def _mass_inertia(r=1.0e-3, rho=workloads.DENSITY):
    m = rho * 4.0 * math.pi * r ** 3 / 3.0
    return m, 0.4 * m * r * r


def test_stick_oscillation_period(oracle_mod):
    """Tangential spring kt, no damping (en = 1), mu large: the contact point
    oscillates harmonically at sqrt(kt / m_eff)."""
    kt, kn = 1.0e4, 1.0e5
    prm = (kn, 1.0, -1.0, 0.9, kt, 0.0, 0.0, 0.0)
    dt = 1.0e-7
    wl = workloads.sphere_on_wall((MODEL_HOOKE_MEMORY, prm), v0=(2.0e-4, 0.0, 0.0), dt=dt)
    o = oracle_mod.Oracle(wl)
    o.prime()
    m, inertia = _mass_inertia()
    r = 1.0e-3
    a = r  # lever arm of the contact point (the overlap is 1e-9 m)
    m_eff = 1.0 / (1.0 / m + a * a / inertia)
    omega = math.sqrt(kt / m_eff)
    period = 2 * math.pi / omega
    nsteps = int(4.2 * period / dt)
    fx = np.empty(nsteps)
    for k in range(nsteps):
        o.step(1, dt)
        fx[k] = o.state()["F"][0, 0]
    # zero crossings of the tangential force, linearly interpolated
    t = (np.arange(nsteps) + 1) * dt
    idx = np.nonzero(np.sign(fx[1:]) != np.sign(fx[:-1]))[0]
    tz = t[idx] - fx[idx] * dt / (fx[idx + 1] - fx[idx])
    assert len(tz) >= 7
    measured = 2 * np.mean(np.diff(tz))
    assert measured == pytest.approx(period, rel=2e-3)
    # amplitude of the force: kt * v0 / omega (the contact point starts at the
    # spring's rest position with the full tangential velocity)
    assert np.abs(fx).max() == pytest.approx(kt * 2.0e-4 / omega, rel=1e-2)


def test_slide_to_roll_five_sevenths(oracle_mod):
    prm = (1.0e5, 0.5, -1.0, 0.3, 1.0e4, 0.0, 0.0, 0.0)
    dt = 2.0e-6
    v0 = 0.5
    wl = workloads.sphere_on_wall((MODEL_HOOKE_MEMORY, prm), v0=(v0, 0.0, 0.0), dt=dt)
    o = oracle_mod.Oracle(wl)
    o.prime()
    o.step(40000, dt)  # 0.08 s; rolling sets in after 2 v0 / (7 mu g) = 0.049 s
    st = o.state()
    m, inertia = _mass_inertia()
    a = 1.0e-3
    want = v0 / (1.0 + inertia / (m * a * a))
    assert st["v"][0, 0] == pytest.approx(want, rel=1e-4)
    # rolling without slipping: v = a w_y
    assert st["w"][0, 1] * a == pytest.approx(st["v"][0, 0], rel=1e-4)


def test_rolling_friction_deceleration(oracle_mod):
    mur = 0.05
    prm = (1.0e5, 0.5, -1.0, 0.5, 1.0e4, mur, 0.0, 0.0)
    dt = 2.0e-6
    r = 1.0e-3
    v0 = 0.05
    # start rolling: w_y = v0 / r
    wl = workloads.sphere_on_wall((MODEL_HOOKE_MEMORY, prm), v0=(v0, 0.0, 0.0),
                                  w0=(0.0, v0 / r, 0.0), dt=dt)
    o = oracle_mod.Oracle(wl)
    o.prime()
    o.step(5000, dt)  # the spring torque saturates
    v1 = o.state()["v"][0, 0]
    o.step(10000, dt)
    v2 = o.state()["v"][0, 0]
    decel = (v1 - v2) / (10000 * dt)
    m, inertia = _mass_inertia()
    half = (0.1, 0.02, 1.0e-3)
    vol = 8.0 * half[0] * half[1] * half[2]
    r_wall = (0.75 / math.pi * vol) ** (1.0 / 3.0)
    req_sphere = r
    rstar = 1.0 / (1.0 / req_sphere + 1.0 / r_wall)
    want = mur * rstar * 9.81 * r / (r * r + inertia / m)
    assert decel == pytest.approx(want, rel=2e-3)


def test_maps_are_mirrored_and_pruned(oracle_mod):
    wl = workloads.memory_box(300, seed=3)
    wl["particles"]["radius"] = wl["particles"]["radius"] * 1.08
    wl["cell_size"] = 2.0 * wl["particles"]["radius"].max()
    o = oracle_mod.Oracle(wl, log_capacity=5000)
    o.prime()
    for _ in range(4):
        o.step(25, wl["dt"])
        c = o.contacts()
        partners = {}
        for a, b in zip(c["id0"].tolist(), c["id1"].tolist()):
            partners.setdefault(a, set()).add(b)
            if b >= 0:
                partners.setdefault(b, set()).add(a)
        assert len(c["id0"]) > 5
        for p in range(o.n):
            mem = o.memory(p)
            # exactly the contacts of the last force evaluation are remembered
            assert set(mem) == partners.get(p, set()), p
            for q, (kd, nrm, spr) in mem.items():
                if q < 0:
                    continue
                kq, nq, sq = o.memory(q)[p]
                np.testing.assert_array_equal(kq, -kd)
                np.testing.assert_array_equal(nq, -nrm)
                np.testing.assert_array_equal(sq, -spr)
                assert np.linalg.norm(nrm) == pytest.approx(1.0, abs=1e-12)
