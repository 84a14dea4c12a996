"""Pins of the CPU oracle against closed forms, brute force and invariants.

The reference's C++ cannot be built here (every geometry header includes
WriterXML.hh, which needs Xerces-C, absent from this image), and its own tests carry no
numeric fixture for the Hooke/Hertz/leap-frog path, so each oracle function is
checked against mathematics that a plausible slip (swapped radii, wrong sign,
dropped term, wrong index) would break.
"""
import math

import numpy as np
import pytest

from pacific_amd import workloads

HOOKE = 1
HERTZ = 2
RHO = 2500.0
CRUST = 5e-5


def sphere_mass(r):
    return RHO * 4.0 * math.pi * r ** 3 / 3.0


def two_sphere_wl(xa, xb, va, vb, ra=1e-3, rb=1e-3, kind=HOOKE, params=(1e5, 0.9, -1.0, 0.5, 0.0),
                  size=0.02, gravity=(0.0, 0.0, 0.0), wa=(0, 0, 0), wb=(0, 0, 0)):
    p = {"x": np.array([xa, xb], float), "v": np.array([va, vb], float),
         "w": np.array([wa, wb], float), "q": np.array([[0, 0, 0, 1.0]] * 2),
         "radius": np.array([ra, rb]), "density": np.full(2, RHO), "crust": np.full(2, CRUST),
         "material": np.zeros(2, np.int32)}
    return {"origin": (0, 0, 0), "size": (size, size, size), "periodic": (0, 0, 0),
            "cell_size": 2.0 * max(ra, rb), "gravity": gravity, "models": [(0, 0, kind, params)],
            "walls": [], "particles": p}


# --------------------------------------------------------------------- geometry
def test_sphere_sphere_matches_geometry(oracle_mod):
    rng = np.random.default_rng(7)
    for _ in range(200):
        ra, rb = rng.uniform(0.8e-3, 1.2e-3, 2)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        a = rng.uniform(0.01, 0.02, 3)
        gap = rng.uniform(-0.9, 0.5) * 2 * CRUST
        b = a + (ra + rb + gap) * n
        rc, pt, ov, d = oracle_mod.sphere_sphere(a, ra, CRUST, b, rb, CRUST)
        true_gap = np.linalg.norm(b - a) - (ra + rb)
        assert d == pytest.approx(true_gap, abs=1e-18)
        if gap > 1e-12:
            assert rc == 0
            continue
        assert rc == 1
        # contact point: midpoint of the two surface points along the centre line
        sa = a + ra * n
        sb = b - rb * n
        np.testing.assert_allclose(pt, 0.5 * (sa + sb), rtol=0, atol=1e-17)
        # overlap vector: |d| long, pointing from B to A when overlapping
        np.testing.assert_allclose(ov, d * n, rtol=0, atol=1e-19)


def test_sphere_sphere_error_threshold(oracle_mod):
    a = np.array([0.01, 0.01, 0.01])
    n = np.array([1.0, 0.0, 0.0])
    # overlap just below / at / above the sum of crusts (RigidBodyWithCrust.cpp:616-624)
    rc, *_ = oracle_mod.sphere_sphere(a, 1e-3, CRUST, a + (2e-3 - 0.99 * 2 * CRUST) * n, 1e-3, CRUST)
    assert rc == 1
    rc, *_ = oracle_mod.sphere_sphere(a, 1e-3, CRUST, a + (2e-3 - 1.01 * 2 * CRUST) * n, 1e-3, CRUST)
    assert rc == -1


def _box_closest(c, e):
    """Closest point of the axis-aligned box [-e, e] to c, and the region kind."""
    q = np.clip(c, -e, e)
    outside = int(np.sum(np.abs(c) > e))
    return q, outside


@pytest.mark.parametrize("rotated", [False, True])
def test_sphere_box_matches_distance_to_box(oracle_mod, rotated):
    rng = np.random.default_rng(11)
    e = np.array([0.004, 0.003, 0.0005])
    r = 1e-3
    for _ in range(400):
        # local centre within r of the box, outside it
        q = rng.uniform(-e, e)
        dirn = rng.normal(size=3)
        dirn /= np.linalg.norm(dirn)
        cl = q + dirn * rng.uniform(0, 1.2 * r)
        if np.all(np.abs(cl) <= e):
            continue
        qc, kind = _box_closest(cl, e)
        dist = np.linalg.norm(cl - qc)
        if rotated:
            Rm, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            if np.linalg.det(Rm) < 0:
                Rm[:, 0] *= -1
        else:
            Rm = np.eye(3)
        origin = rng.uniform(0.0, 0.02, 3)
        cw = Rm @ cl + origin
        rc, pt, ov, d = oracle_mod.sphere_box(cw, r, CRUST, Rm.reshape(-1), origin, e, CRUST)
        if dist >= r:
            assert rc == 0
            continue
        if r - dist >= 2 * CRUST:
            assert rc == -1
            continue
        assert rc == 1
        assert d == pytest.approx(dist - r, abs=1e-15)
        nrm = (cl - qc) / dist
        # Box.cpp: a face contact point sits half the overlap inside the face;
        # an edge or corner contact point sits half the overlap outside it
        # (ContactEdgeSPHERE / ContactCornerSPHERE use the opposite sign).
        sgn = 1.0 if kind == 1 else -1.0
        expect_local = qc + sgn * 0.5 * d * nrm
        np.testing.assert_allclose(pt, Rm @ expect_local + origin, rtol=0, atol=2e-17 if not rotated else 5e-15)
        # overlap vector along (contact point - centre), length |d|
        u = (pt - cw) / np.linalg.norm(pt - cw)
        np.testing.assert_allclose(ov, d * u, rtol=0, atol=1e-18)


# --------------------------------------------------------------------- contact laws
def _binary_collision(oracle_mod, kind, params, v0, dt, ra=1e-3, rb=1e-3):
    size = 0.02
    c = size / 2
    gap = 2e-7
    xa = [c - ra - gap / 2, c, c]
    xb = [c + rb + gap / 2, c, c]
    wl = two_sphere_wl(xa, xb, [v0 / 2, 0, 0], [-v0 / 2, 0, 0], ra, rb, kind, params, size)
    o = oracle_mod.Oracle(wl)
    o.prime()
    t_contact = 0
    dmax = 0.0
    in_contact = False
    for k in range(200000):
        o.step(1, dt)
        s = o.state()
        d = np.linalg.norm(s["x"][1] - s["x"][0]) - (ra + rb)
        if d < 0:
            in_contact = True
            t_contact += 1
            dmax = max(dmax, -d)
        elif in_contact:
            vrel = s["v"][1][0] - s["v"][0][0]
            return vrel / v0, t_contact * dt, dmax
    raise AssertionError("collision did not end")


def test_hooke_binary_collision_restitution_and_duration(oracle_mod):
    kn, en = 1e5, 0.9
    v0 = 0.1
    dt = 2e-8
    e, tc, dmax = _binary_collision(oracle_mod, HOOKE, (kn, en, -1.0, 0.5, 0.0), v0, dt)
    m = sphere_mass(1e-3)
    mstar = m / 2
    beta = math.log(en) / math.sqrt(math.pi ** 2 + math.log(en) ** 2)
    omega0 = math.sqrt(kn / mstar)
    eta = -beta * omega0
    theta = math.sqrt(omega0 ** 2 - eta ** 2)
    # HookeContactForceModel::computeAndWriteEstimates (HookeContactForceModel.cpp:208-230)
    assert tc == pytest.approx(math.pi / theta, abs=3 * dt)
    assert e == pytest.approx(en, rel=2e-3)
    # maximum overlap of the damped oscillator, at t0 = atan(theta/eta)/theta
    t0 = math.atan2(theta, eta) / theta
    dm = (v0 / theta) * math.exp(-eta * t0) * math.sin(theta * t0)
    assert dmax == pytest.approx(dm, rel=2e-3)


def test_hertz_elastic_collision_time(oracle_mod):
    Es = 1e7
    v0 = 0.1
    e, tc, dmax = _binary_collision(oracle_mod, HERTZ, (Es, 0.99999999, Es / 4, 0.4, 0.0), v0, 2e-8)
    m = sphere_mass(1e-3)
    mstar = m / 2
    Rstar = 0.5e-3
    # elastic Hertz impact: t_c = 2.87 (m*^2 / (R* E*^2 v0))^(1/5),
    # delta_max = (15 m* v0^2 / (16 E* sqrt(R*)))^(2/5)
    assert tc == pytest.approx(2.868 * (mstar ** 2 / (Rstar * Es ** 2 * v0)) ** 0.2, rel=5e-3)
    assert dmax == pytest.approx((15 * mstar * v0 ** 2 / (16 * Es * math.sqrt(Rstar))) ** 0.4, rel=2e-3)
    assert e == pytest.approx(1.0, abs=2e-3)


def test_hertz_force_scales_as_overlap_to_three_halves(oracle_mod):
    p = (1e7, 0.85, 2.5e6, 0.4, 0.0)
    x0 = np.array([0.0, 0.0, 0.0])
    z = np.zeros(3)
    m = sphere_mass(1e-3)
    req = 1e-3
    f = []
    for delta in (1e-6, 4e-6):
        x1 = np.array([2e-3 - delta, 0, 0])
        ov = np.array([-delta, 0, 0])
        pt = np.array([1e-3 - delta / 2, 0, 0])
        f.append(oracle_mod.contact_force(HERTZ, p, pt, ov, -delta, x0, z, z, m, req, x1, z, z, m, req))
    assert f[0][0] < 0  # pushes body 0 away from body 1
    assert f[1][0] / f[0][0] == pytest.approx(8.0, rel=1e-12)
    # magnitude 4/3 E* sqrt(R* delta) delta with R* = R/2
    assert -f[0][0] == pytest.approx(4.0 / 3.0 * 1e7 * math.sqrt(0.5e-3 * 1e-6) * 1e-6, rel=1e-12)


def test_hooke_coulomb_cap(oracle_mod):
    kn, en, mu = 1e5, 0.9, 0.3
    m = sphere_mass(1e-3)
    x0 = np.zeros(3)
    x1 = np.array([2e-3 - 1e-6, 0, 0])
    pt = np.array([1e-3 - 0.5e-6, 0, 0])
    ov = np.array([-1e-6, 0, 0])
    z = np.zeros(3)
    # large sliding velocity along y: the tangential force saturates at mu |Fn|
    f = oracle_mod.contact_force(HOOKE, (kn, en, -1.0, mu, 0.0), pt, ov, -1e-6, x0,
                                 np.array([0, 1.0, 0]), z, m, 1e-3, x1, z, z, m, 1e-3)
    fn = abs(f[0])
    assert fn == pytest.approx(kn * 1e-6, rel=1e-12)
    assert f[1] == pytest.approx(-mu * fn, rel=1e-12)
    # slow sliding: viscous tangential force -gamma_t v_t below the cap
    vt = 1e-6
    f = oracle_mod.contact_force(HOOKE, (kn, en, -1.0, mu, 0.0), pt, ov, -1e-6, x0,
                                 np.array([0, vt, 0]), z, m, 1e-3, x1, z, z, m, 1e-3)
    beta = math.log(en) / math.sqrt(math.pi ** 2 + math.log(en) ** 2)
    mstar = m / 2
    gammat = 2 * (-beta * math.sqrt(kn / mstar)) * mstar
    assert f[1] == pytest.approx(-gammat * vt, rel=1e-12)


def test_rolling_contact_uses_surface_velocity(oracle_mod):
    # body 0 spins about z: surface velocity at the contact point is w x (c - x0)
    kn, en, mu = 1e5, 0.9, 10.0
    m = sphere_mass(1e-3)
    x0 = np.zeros(3)
    x1 = np.array([2e-3 - 1e-6, 0, 0])
    pt = np.array([1e-3 - 0.5e-6, 0, 0])
    ov = np.array([-1e-6, 0, 0])
    z = np.zeros(3)
    w = np.array([0, 0, 1e-3])
    f = oracle_mod.contact_force(HOOKE, (kn, en, -1.0, mu, 0.0), pt, ov, -1e-6, x0, z, w, m, 1e-3,
                                 x1, z, z, m, 1e-3)
    beta = math.log(en) / math.sqrt(math.pi ** 2 + math.log(en) ** 2)
    gammat = 2 * (-beta * math.sqrt(kn / (m / 2))) * (m / 2)
    vsurf = np.cross(w, pt - x0)
    assert f[1] == pytest.approx(-gammat * vsurf[1], rel=1e-12)


# --------------------------------------------------------------------- broad phase
def _brute_pairs(x, r, L=None):
    n = len(r)
    pairs = set()
    for i in range(n):
        d = x[i + 1:] - x[i]
        if L is not None:
            d -= L * np.round(d / L)
        dist = np.sqrt((d * d).sum(1)) - (r[i] + r[i + 1:])
        for j in np.nonzero(dist < 0)[0]:
            pairs.add((i, i + 1 + int(j)))
    return pairs


def _oracle_pairs(c):
    return {(min(a, b), max(a, b)) for a, b in zip(c["id0"], c["id1"]) if b >= 0}


@pytest.mark.parametrize("periodic", [False, True])
def test_linked_cell_finds_exactly_the_brute_force_pairs(oracle_mod, periodic):
    if periodic:
        wl = workloads.periodic_hertz(3000, seed=21, steps=1)
        L = wl["size"][0]
    else:
        wl = workloads.plain_box(1000, seed=3, steps=1)
        L = None
    p = wl["particles"]
    # grow the radii so the non-overlapping start has contacts
    p["radius"] = p["radius"] * 1.025 if periodic else p["radius"] * 1.08
    if not periodic:
        wl["cell_size"] = 2 * p["radius"].max()
    o = oracle_mod.Oracle(wl, log_capacity=100000)
    o.prime()
    got = _oracle_pairs(o.contacts())
    want = _brute_pairs(p["x"], p["radius"], L)
    assert len(want) > 50
    assert got == want


def test_grid_dimensions_and_cells(oracle_mod):
    wl = workloads.plain_box(10, steps=1)
    o = oracle_mod.Oracle(wl)
    nc, e, _ = o.grid()
    # LinkedCell::set: n = int((L + 1e-10) / h), edge = L / n
    assert list(nc) == [10, 10, 20]
    np.testing.assert_allclose(e, [0.002, 0.002, 0.002], rtol=1e-15)
    wl = workloads.periodic_hertz(2000, seed=5, steps=1)
    o = oracle_mod.Oracle(wl)
    nc, e, _ = o.grid()
    L = wl["size"][0]
    n = int((L + 1e-10) / wl["cell_size"])
    assert list(nc) == [n, n, n]
    assert e[0] == L / n


# --------------------------------------------------------------------- integrator
def test_free_flight_under_gravity_is_exact_parabola(oracle_mod):
    g = (0.0, 0.0, -9.81)
    wl = two_sphere_wl([0.005, 0.005, 0.015], [0.015, 0.015, 0.015], [0.01, 0.02, 0.05],
                       [-0.03, 0.0, 0.0], gravity=g)
    o = oracle_mod.Oracle(wl)
    o.prime()
    dt, n = 1e-5, 1000
    o.step(n, dt)
    s = o.state()
    t = n * dt
    x0 = wl["particles"]["x"]
    v0 = wl["particles"]["v"]
    np.testing.assert_allclose(s["x"], x0 + v0 * t + 0.5 * np.array(g) * t * t, rtol=0, atol=1e-15)
    np.testing.assert_allclose(s["v"], v0 + np.array(g) * t, rtol=0, atol=1e-14)
    np.testing.assert_allclose(s["F"], np.tile(np.array(g) * sphere_mass(1e-3), (2, 1)), rtol=1e-15)


def test_torque_free_sphere_rotates_at_constant_rate(oracle_mod):
    w0 = np.array([30.0, -10.0, 20.0])
    wl = two_sphere_wl([0.005, 0.005, 0.005], [0.015, 0.015, 0.015], [0, 0, 0], [0, 0, 0], wa=w0)
    o = oracle_mod.Oracle(wl)
    o.prime()
    dt, n = 1e-5, 2000
    o.step(n, dt)
    s = o.state()
    np.testing.assert_allclose(s["w"][0], w0, rtol=1e-12)
    q = s["q"][0]
    assert abs(np.linalg.norm(q) - 1.0) < 1e-14 * n  # unit to round-off per step
    ang = np.linalg.norm(w0) * n * dt
    axis = w0 / np.linalg.norm(w0)
    np.testing.assert_allclose(q, np.r_[math.sin(ang / 2) * axis, math.cos(ang / 2)], atol=1e-12)


# --------------------------------------------------------------------- invariants
def test_periodic_step_conserves_momentum(oracle_mod):
    wl = workloads.periodic_hertz(3000, seed=22, steps=1)
    wl["particles"]["radius"] = wl["particles"]["radius"] * 1.02
    o = oracle_mod.Oracle(wl, log_capacity=100000)
    o.prime()
    m = RHO * 4.0 * math.pi * wl["particles"]["radius"] ** 3 / 3.0
    s0 = o.state()
    fabs = np.abs(o.contacts()["force"]).sum()
    assert np.abs(s0["F"].sum(0)).max() <= 1e-12 * fabs
    p0 = (m[:, None] * s0["v"]).sum(0)
    o.step(50, wl["dt"])
    s1 = o.state()
    p1 = (m[:, None] * s1["v"]).sum(0)
    scale = (m[:, None] * np.abs(s0["v"])).sum()
    assert np.abs(p1 - p0).max() <= 1e-13 * scale


def _energy(s, m, inertia):
    return 0.5 * (m[:, None] * s["v"] ** 2).sum() + 0.5 * (inertia[:, None] * s["w"] ** 2).sum()


def test_elastic_frictionless_gas_conserves_energy_dissipative_loses_it(oracle_mod):
    kn = 1e5
    for en, mu, check in ((1.0, 0.0, "conserve"), (0.7, 0.5, "decrease")):
        wl = workloads.periodic_hertz(2000, seed=23, steps=1)
        wl["models"] = [(0, 0, HOOKE, (kn, en, -1.0, mu, 0.0))]
        r = wl["particles"]["radius"]
        m = RHO * 4.0 * math.pi * r ** 3 / 3.0
        inertia = RHO * 8.0 * math.pi * r ** 5 / 15.0
        o = oracle_mod.Oracle(wl, log_capacity=100000)
        o.prime()
        # kinetic energy plus the Hooke spring energy 1/2 kn delta^2 of open contacts
        e0 = _energy(o.state(), m, inertia) + 0.5 * kn * (o.contacts()["overlap"] ** 2).sum()
        o.step(400, 2e-6)
        e1 = _energy(o.state(), m, inertia) + 0.5 * kn * (o.contacts()["overlap"] ** 2).sum()
        if check == "conserve":
            assert abs(e1 - e0) / e0 < 1e-4
        else:
            assert e1 < e0 * 0.999


# --------------------------------------------------------------------- errors
def test_error_paths(oracle_mod):
    # ContactError: the overlap reaches ten crust sums, where even the retry's
    # shrunk spheres touch (an overlap of 1.2 crust sums is retried instead)
    wl = two_sphere_wl([0.01, 0.01, 0.01], [0.01 + 2e-3 - 1.0e-3, 0.01, 0.01], [0] * 3, [0] * 3)
    o = oracle_mod.Oracle(wl)
    with pytest.raises(oracle_mod.OracleError) as ei:
        o.prime()
    assert ei.value.code == oracle_mod.ERR_CONTACT
    # MotionError: |v| dt > crust
    wl = two_sphere_wl([0.005] * 3, [0.015] * 3, [10.0, 0, 0], [0] * 3)
    o = oracle_mod.Oracle(wl)
    o.prime()
    with pytest.raises(oracle_mod.OracleError) as ei:
        o.step(1, 1e-5)
    assert ei.value.code == oracle_mod.ERR_MOTION and ei.value.id0 == 0
    # SimulationError: leaves the (non-periodic) grid
    wl = two_sphere_wl([0.005] * 3, [0.019999, 0.015, 0.015], [0] * 3, [1.0, 0, 0])
    o = oracle_mod.Oracle(wl)
    o.prime()
    with pytest.raises(oracle_mod.OracleError) as ei:
        o.step(1, 2e-6)
    assert ei.value.code == oracle_mod.ERR_SIMULATION and ei.value.id0 == 1
