"""Parity of the HIP convex path (boxes + tetrahedra, GJK narrow phase, Johnson
or signed-volume: the "_sv" cases) with the CPU oracle, through the C ABI.

Bar as for spheres: pair sets bit-exact; each contact's point, overlap
vector, overlap and force bit-identical to the oracle's per-pair evaluation
with the same body order; per-particle F and T within 1e-12 of the
per-particle scale; positions, velocities, angular velocities, quaternions and
bases within 1e-9 relative after 100 steps.
"""
import numpy as np
import pytest

from gdem_testutil import force_scales, normwise_ok, pair_set, rel_err
from pacific_amd import gdem, workloads

pytestmark = pytest.mark.gpu


def _extents(wl):
    """Lowest and highest world coordinate of every particle's vertices."""
    p = wl["particles"]
    cube = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)],
                    float) * np.array(workloads.CUBE_HALF)
    verts = {0: cube, 1: workloads.tetrahedron_vertices()}
    n = len(p["shape"])
    lo, hi = np.empty((n, 3)), np.empty((n, 3))
    for i in range(n):
        w = verts[int(p["shape"][i])] @ p["basis"][i].reshape(3, 3).T + p["x"][i]
        lo[i], hi[i] = w.min(0), w.max(0)
    return lo, hi


def push_walls_in(wl, depth=5e-5):
    """Move each wall inwards so that the particle reaching furthest towards
    it overlaps it by depth (wall contacts from the first step)."""
    lo, hi = _extents(wl)
    walls = []
    for k, (basis, centre, half, crust, mat) in enumerate(wl["walls"]):
        d, side = k // 2, (-1 if k % 2 == 0 else 1)
        c = list(centre)
        if side < 0:
            c[d] += lo[:, d].min() + depth - wl["origin"][d]
        else:
            c[d] -= wl["origin"][d] + wl["size"][d] - (hi[:, d].max() - depth)
        walls.append((basis, tuple(c), half, crust, mat))
    wl["walls"] = walls
    return wl


def walls_to_rectangles(wl):
    """Each (pushed-in) box wall replaced by a Rectangle on its inner face,
    of the same extent; the particles meet them through the reference's
    ClosestPointRECTANGLE (a particle's deepest support point, projected)."""
    rects = []
    for k, (basis, centre, half, crust, mat) in enumerate(wl["walls"]):
        d, side = k // 2, (-1 if k % 2 == 0 else 1)
        a, b = (d + 1) % 3, (d + 2) % 3
        c = list(centre)
        c[d] += half[d] if side < 0 else -half[d]
        rects.append((workloads.FACE_BASIS[d], tuple(c), 2.0 * half[a], 2.0 * half[b], crust, mat))
    wl["walls"] = []
    wl["rectangles"] = rects
    return wl


def case(name):
    if name.endswith("_sv"):
        # the signed-volume narrow phase (CollisionDetection Type="GJK_SV")
        wl = case(name[:-3])
        wl["narrow_phase"] = "GJK_SV"
        return wl
    if name == "convex_rectangles":
        return walls_to_rectangles(case("convex_walls"))
    if name == "convex_walls":
        return push_walls_in(workloads.convex_box(600, seed=5, phi=0.15, height=0.015,
                                                  overlap=4e-5, sigma_v=0.2))
    if name == "convex_dense":
        return push_walls_in(workloads.convex_box(2000, seed=9, phi=0.18, height=0.02,
                                                  overlap=4e-5, sigma_v=0.3))
    raise KeyError(name)


CASES = ["convex_walls", "convex_dense", "convex_rectangles", "convex_walls_sv",
         "convex_dense_sv"]


def _masses(orc, wl):
    p = wl["particles"]
    vols = [orc.shape_info(k)[0] for k in range(len(wl["shapes"]))]
    return p["density"] * np.array([vols[s] for s in p["shape"]])


@pytest.mark.parametrize("name", CASES)
def test_convex_initial_contacts_and_forces(oracle_mod, name):
    print(f"[case] {name} prime", flush=True)
    wl = case(name)
    n = len(wl["particles"]["shape"])
    orc = oracle_mod.ConvexOracle(wl, log_capacity=20 * n)
    orc.prime()
    oc = orc.contacts()
    os_ = orc.state()
    sysg = gdem.build_system(wl)
    sysg.set_contact_recording(32)
    sysg.prime()
    gs = sysg.get_state()
    gc = sysg.get_contacts()
    assert len(oc["id0"]) > 20 and (oc["id1"] < 0).any()
    assert pair_set(gc) == pair_set(oc)
    for k in range(len(gc["id0"])):
        a, b = int(gc["id0"][k]), int(gc["id1"][k])
        if b >= 0:
            rc, pt, ov, dist, f = orc.pair(a, b)
        else:
            rc, pt, ov, dist, f = orc.wall_pair(a, -1 - b)
        assert rc == 1, (a, b)
        assert np.array_equal(gc["point"][k], pt), (a, b)
        assert np.array_equal(gc["ovec"][k], ov), (a, b)
        assert gc["overlap"][k] == dist, (a, b)
        assert np.array_equal(gc["force"][k], f), (a, b, gc["force"][k], f)
    x = wl["particles"]["x"]
    weight = _masses(orc, wl)[:, None] * np.array(wl["gravity"])[None, :]
    sf, st = force_scales(n, oc, x, weight)
    ok, err, bad = normwise_ok(gs["F"], os_["F"], sf, 1e-12)
    assert ok, (np.nonzero(bad)[0][:5], err[bad][:5])
    ok, err, bad = normwise_ok(gs["T"], os_["T"], st, 1e-12)
    assert ok, (np.nonzero(bad)[0][:5], err[bad][:5])
    sysg.close()


@pytest.mark.parametrize("name", CASES)
def test_convex_trajectory_100_steps(oracle_mod, name):
    print(f"[case] {name} 100 steps", flush=True)
    wl = case(name)
    n = len(wl["particles"]["shape"])
    orc = oracle_mod.ConvexOracle(wl, log_capacity=20 * n)
    orc.prime()
    orc.step(100, wl["dt"])
    o = orc.state()
    sysg = gdem.build_system(wl)
    sysg.set_contact_recording(32)
    sysg.step(100, wl["dt"])
    g = sysg.get_state()
    gb = sysg.get_basis()
    x0 = wl["particles"]["x"]
    assert rel_err(g["x"] - x0, o["x"] - x0) <= 1e-9
    assert rel_err(g["v"], o["v"]) <= 1e-9
    assert rel_err(g["w"], o["w"]) <= 1e-9
    assert rel_err(g["q"], o["q"]) <= 1e-9
    assert rel_err(gb, o["basis"]) <= 1e-9
    assert pair_set(sysg.get_contacts()) == pair_set(orc.contacts())
    sysg.close()


def test_signed_volume_narrow_phase_is_the_one_run():
    """The GJK_SV switch reaches the kernel: its contacts differ from the
    Johnson GJK's in the last bits (same pairs, different rounding), and the
    context reports the choice."""
    print("[case] convex narrow phase switch", flush=True)
    out = {}
    for name in ("convex_dense", "convex_dense_sv"):
        wl = case(name)
        s = gdem.build_system(wl)
        assert s.narrow_phase() == ("GJK_SV" if name.endswith("_sv") else "GJK")
        s.set_contact_recording(32)
        s.prime()
        out[name] = s.get_contacts()
        s.close()
    a, b = out["convex_dense"], out["convex_dense_sv"]
    assert pair_set(a) == pair_set(b)
    ka = {(int(i), int(j)): k for k, (i, j) in enumerate(zip(a["id0"], a["id1"]))}
    differ = 0
    for k, (i, j) in enumerate(zip(b["id0"], b["id1"])):
        m = ka[(int(i), int(j))]
        assert b["overlap"][k] == pytest.approx(a["overlap"][m], rel=1e-5, abs=1e-13)
        differ += not np.array_equal(b["point"][k], a["point"][m])
    assert differ > 0


def test_convex_phase_calls_equal_fused_step():
    print("[case] convex phases vs fused", flush=True)
    wl = case("convex_walls")
    dt = wl["dt"]
    a = gdem.build_system(wl)
    a.step(10, dt)
    sa, ba = a.get_state(), a.get_basis()
    b = gdem.build_system(wl)
    b.prime()
    for _ in range(10):
        b.move(0.5 * dt, dt)
        b.link_update()
        b.compute_forces()
        b.advance_velocity(0.5 * dt)
    sb, bb = b.get_state(), b.get_basis()
    for k in ("x", "v", "w", "q", "F", "T"):
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(ba, bb)
    a.close()
    b.close()


def test_convex_contact_error_names_pair(oracle_mod):
    """A cube and a tetrahedron whose shrunk bodies overlap raise ContactError
    naming both."""
    print("[case] convex contact error", flush=True)
    wl = case("convex_walls")
    p = wl["particles"]
    # put particle 1 (a tetrahedron) on particle 0's centre
    p["x"][1] = p["x"][0] + np.array([1e-5, 0.0, 0.0])
    orc = oracle_mod.ConvexOracle(wl)
    with pytest.raises(oracle_mod.OracleError) as eo:
        orc.prime()
    sysg = gdem.build_system(wl)
    with pytest.raises(gdem.ContactError) as eg:
        sysg.prime()
        sysg.synchronize()
    assert {eg.value.id0, eg.value.id1} == {0, 1}
    assert {eo.value.id0, eo.value.id1} == {0, 1}
    sysg.close()
