"""Wall loads on the GPU path (gdem_set_wall_loads / gdem_get_wall_loads)
against the oracle's wall torsors, through the C ABI.

* At the priming evaluation and after 50 steps (states equal to rounding)
  each wall's force and torque match the oracle's to 1e-11 of the summed
  contact-force magnitudes on the walls (the sums run in another order);
* the balance of linear and angular momentum (particle contact forces and
  torques against wall loads) holds on the GPU's own state, with graph
  replay, for Hooke and HookeMemory with rolling friction;
* the loads are reproducible bit for bit from run to run (fixed-order
  reduction, no atomics)."""
import numpy as np
import pytest

from pacific_amd import gdem, workloads
from test_oracle_wall_load_pins import _balances

pytestmark = pytest.mark.gpu


def case(name):
    if name == "shaken_box":
        wl = workloads.shaken_box(1000, seed=1)
    elif name == "memory_box":
        wl = workloads.memory_box(1000, seed=1)   # walls touched from the start
        wl["particles"]["radius"] = wl["particles"]["radius"] * 1.08
        wl["cell_size"] = 2.0 * wl["particles"]["radius"].max()
        return wl
    else:
        return workloads.rotating_drum(memory=True)
    wl["particles"]["radius"] = wl["particles"]["radius"] * 1.04
    wl["cell_size"] = 2.0 * wl["particles"]["radius"].max()
    return wl


def _wall_scale(contacts):
    wall = contacts["id1"] < 0
    return float(np.linalg.norm(contacts["force"][wall], axis=1).sum())


@pytest.mark.parametrize("name", ["shaken_box", "memory_box", "rotating_drum"])
def test_loads_match_oracle(oracle_mod, name):
    print(f"[case] wall loads {name}", flush=True)
    wl = case(name)
    n = len(wl["particles"]["radius"])
    orc = oracle_mod.Oracle(wl, log_capacity=10 * n)
    orc.prime()
    s = gdem.build_system(wl)
    s.set_wall_loads(True)
    s.prime()
    touched = False
    for steps in (0, 50):
        if steps:
            orc.step(steps, wl["dt"])
            s.step(steps, wl["dt"])
        fo, to = orc.wall_loads()
        fg, tg = s.get_wall_loads()
        scale = _wall_scale(orc.contacts())
        touched = touched or scale > 0.0
        print(f"[case] steps {steps}: max force diff {np.abs(fg - fo).max():.3e}, "
              f"torque diff {np.abs(tg - to).max():.3e}, scale {scale:.3e}", flush=True)
        np.testing.assert_allclose(fg, fo, rtol=0, atol=1e-11 * scale)
        np.testing.assert_allclose(tg, to, rtol=0, atol=1e-11 * scale * 0.05)
    assert touched
    s.close()


@pytest.mark.parametrize("memory", [False, True])
def test_momentum_balance_on_gpu(memory):
    print(f"[case] wall-load balance, memory {memory}", flush=True)
    wl = workloads.memory_box(1000, seed=1) if memory else workloads.plain_box(1000, seed=1)
    wl["particles"]["radius"] = wl["particles"]["radius"] * 1.08
    wl["cell_size"] = 2.0 * wl["particles"]["radius"].max()
    s = gdem.build_system(wl)
    s.set_wall_loads(True)
    for _ in range(3):
        s.step(20, wl["dt"])        # graph replay for all but the last step
        f, t = s.get_wall_loads()
        assert np.abs(f).sum() > 0.0
        lin, ang, sf, stq = _balances(wl, s.get_state(), f, t)
        assert np.abs(lin).max() <= 1e-13 * sf
        assert np.abs(ang).max() <= 1e-13 * stq
    s.close()


def test_loads_reproducible():
    print("[case] wall loads reproducible", flush=True)
    wl = case("shaken_box")
    out = []
    for _ in range(2):
        s = gdem.build_system(wl)
        s.set_wall_loads(True)
        s.step(80, wl["dt"])
        out.append(s.get_wall_loads())
        s.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_loads_switched_on_after_priming():
    """Switched on after gdem_prime, the loads are those of the priming
    evaluation (re-run on the unchanged state), equal to a run that switched
    them on first; switching on twice keeps them; after a step, or with a
    memory model, a late switch is refused."""
    print("[case] wall loads switched on after prime", flush=True)
    wl = case("memory_box")  # walls touched from the start
    hooke = (1.0e5, 0.9, -1.0, 0.5, 0.0)
    wl["models"] = [(0, 0, workloads.MODEL_HOOKE, hooke), (0, 1, workloads.MODEL_HOOKE, hooke)]
    a = gdem.build_system(wl)
    a.set_wall_loads(True)
    a.prime()
    b = gdem.build_system(wl)
    b.prime()
    b.set_wall_loads(True)
    fa, ta = a.get_wall_loads()
    fb, tb = b.get_wall_loads()
    assert np.abs(fa).max() > 0
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(ta, tb)
    b.set_wall_loads(True)  # already on: kept
    np.testing.assert_array_equal(b.get_wall_loads()[0], fa)
    a.step(3, wl["dt"])
    b.step(3, wl["dt"])
    np.testing.assert_array_equal(a.get_state()["x"], b.get_state()["x"])
    c = gdem.build_system(wl)
    c.prime()
    c.step(1, wl["dt"])
    with pytest.raises(gdem.GdemError):
        c.set_wall_loads(True)
    for s in (a, b, c):
        s.close()
