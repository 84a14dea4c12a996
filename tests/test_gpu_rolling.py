"""Rolling resistance (kr != 0) of the Hooke and Hertz models on the GPU
(canonical force kernel; the packed step and the convex, differentiated
paths refuse it) against the oracle, whose moment is pinned in
tests/test_oracle_rolling_pins.py.

* at prime, every particle's force and torque within 1e-12 of its scale
  (contact forces times lever arms, plus |weight| and |torque|);
* trajectories within 1e-9 relative after 100 steps (or 3 times the
  oracle's own one-ulp divergence where larger: the walled box), with skin
  lists on (gdem_step falls back from the packed step) and with a wall box;
* the models refused where the moment is not carried."""
import numpy as np
import pytest

from gdem_testutil import force_scales, normwise_ok, rel_err
from pacific_amd import gdem, workloads

pytestmark = pytest.mark.gpu

RHO = 2500.0


def _spun(wl, seed, kr):
    p = wl["particles"]
    rng = np.random.default_rng(seed)
    p["w"] = rng.normal(size=p["x"].shape) * 40.0
    models = []
    for (a, b, kind, prm) in wl["models"]:
        prm = list(prm)
        prm[4] = kr
        models.append((a, b, kind, tuple(prm)))
    wl["models"] = models
    return wl


def case(name):
    if name == "box_hooke_walls":
        wl = workloads.plain_box(1000, seed=1)
        wl["particles"]["radius"] = wl["particles"]["radius"] * 1.08
        wl["cell_size"] = 2.0 * wl["particles"]["radius"].max()
        # the moment acts like a rotational damper of rate kr R* FN a / I:
        # kr keeps that rate times dt below one at the overlaps of this box
        return _spun(wl, 5, 0.02)
    if name == "periodic_hertz":
        wl = workloads.periodic_hertz(20000, seed=2)
        wl["particles"]["radius"] = wl["particles"]["radius"] * 1.02
        wl["cell_size"] = 2.4 * wl["particles"]["radius"].max()
        return _spun(wl, 6, 3.0)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["box_hooke_walls", "periodic_hertz"])
def test_prime_forces_and_torques(oracle_mod, name):
    print(f"[case] rolling resistance prime, {name}", flush=True)
    wl = case(name)
    p = wl["particles"]
    n = len(p["radius"])
    orc = oracle_mod.Oracle(wl, log_capacity=10 * n)
    orc.prime()
    oc, os_ = orc.contacts(), orc.state()
    s = gdem.build_system(wl)
    s.prime()
    gs = s.get_state()
    assert len(oc["id0"]) > 50
    m = RHO * (4.0 * np.pi * p["radius"] ** 3 / 3.0)
    weight = m[:, None] * np.array(wl["gravity"])[None, :]
    sf, st = force_scales(n, oc, p["x"], weight)
    st = st + np.linalg.norm(os_["T"], axis=1)
    # the moments matter: the torques differ from those without rolling resistance
    wl0 = case(name)
    wl0["models"] = [(a, b, k, tuple(list(prm[:4]) + [0.0])) for (a, b, k, prm) in wl0["models"]]
    o0 = oracle_mod.Oracle(wl0)
    o0.prime()
    assert rel_err(o0.state()["T"], os_["T"]) > 1e-3
    ok, err, bad = normwise_ok(gs["F"], os_["F"], sf, 1e-12)
    assert ok, (np.nonzero(bad)[0][:5], err[bad][:5])
    ok, err, bad = normwise_ok(gs["T"], os_["T"], st, 1e-12)
    assert ok, (np.nonzero(bad)[0][:5], err[bad][:5])
    s.close()


def _disp(x, x0, wl):
    d = x - x0
    if any(wl["periodic"]):
        L = np.array(wl["size"])
        d -= L * np.round(d / L)
    return d


def _oracle_floor(oracle_mod, name, steps, ref):
    """The oracle's own divergence after `steps` steps when every initial
    velocity moves by one ulp (two draws of the directions): the rounding
    floor of the case."""
    out = {}
    for seed in (7, 8):
        wl = case(name)
        v = wl["particles"]["v"]
        up = np.random.default_rng(seed).random(v.shape) < 0.5
        wl["particles"]["v"] = np.nextafter(v, np.where(up, np.inf, -np.inf))
        o = oracle_mod.Oracle(wl)
        o.prime()
        o.step(steps, wl["dt"])
        st = o.state()
        x0 = wl["particles"]["x"]
        errs = {"x": rel_err(_disp(st["x"], x0, wl), _disp(ref["x"], x0, wl)),
                "v": rel_err(st["v"], ref["v"]), "w": rel_err(st["w"], ref["w"])}
        for k, e in errs.items():
            out[k] = max(out.get(k, 0.0), e)
    return out


@pytest.mark.parametrize("name", ["box_hooke_walls", "periodic_hertz"])
def test_trajectory_100_steps(oracle_mod, name):
    """Bar: 1e-9 relative, or 3 times the oracle's own one-ulp divergence
    where that is larger (the walled box amplifies rounding)."""
    print(f"[case] rolling resistance 100 steps, {name}", flush=True)
    wl = case(name)
    steps = 100
    orc = oracle_mod.Oracle(wl)
    orc.prime()
    orc.step(steps, wl["dt"])
    o = orc.state()
    floor = _oracle_floor(oracle_mod, name, steps, o)
    bar = {k: max(1e-9, 3.0 * e) for k, e in floor.items()}
    s = gdem.build_system(wl)
    s.prime()
    s.step(steps, wl["dt"])
    g = s.get_state()
    x0 = wl["particles"]["x"]
    errs = {"x": rel_err(_disp(g["x"], x0, wl), _disp(o["x"], x0, wl)),
            "v": rel_err(g["v"], o["v"]), "w": rel_err(g["w"], o["w"])}
    print(f"[case] errors {errs} oracle one-ulp divergence {floor}", flush=True)
    for k in errs:
        assert errs[k] <= bar[k], (k, errs[k], bar[k])
    s.close()


def test_refused_where_the_moment_is_not_carried():
    print("[case] rolling resistance refusals", flush=True)
    wl = workloads.convex_box(64, seed=3)
    s = gdem.GdemSystem(wl["origin"], wl["size"], wl["periodic"], wl["cell_size"], wl["gravity"])
    for sh in wl["shapes"]:
        if sh[0] == "box":
            s.add_box_shape(sh[1])
        else:
            s.add_polyhedron_shape(sh[1], sh[2])
    p = wl["particles"]
    s.set_convex_particles(p["x"], p["v"], p["w"], p["q"], p["basis"], p["shape"], p["density"],
                           p["crust"], p["material"])
    with pytest.raises(gdem.GdemError):
        s.set_contact_model(0, 0, gdem.MODEL_HOOKE, (2e5, 0.8, -1.0, 0.5, 1.0))
    s.close()
