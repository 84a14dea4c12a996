"""Adjoint gradient of calibration trajectories (gdem_diff_*): checked against
central finite differences of the GPU loss, against finite differences of the
CPU oracle's loss, and through the torch.autograd harness.

Bar: relative error <= 1e-5 against central differences with h = 1e-6
relative (the contact set does not change within +-h at these seeds)."""
import numpy as np
import pytest

from gdem_testutil import rel_err
from pacific_amd import workloads

pytestmark = pytest.mark.gpu

NSTEPS, K, OBS, DT = 120, 25, (40, 80, 120), 1.0e-5
GUESS = np.array([1.2, 0.9, 1.1])


def setup_case(nsys=3, n=300, seed=21):
    from pacific_amd import gdem
    systems, hidden = workloads.calibration_systems(nsys, n=n, seed=seed)
    eng = gdem.GdemDiff(systems, NSTEPS, DT, K, OBS)
    v0 = np.concatenate([w["particles"]["v"] for w in systems])
    eng.set_parameters(hidden, v0)
    target = eng.record()
    eng.set_targets(target)
    theta = hidden * GUESS[None, :]
    eng.set_parameters(theta, v0)
    return eng, systems, hidden, theta, v0, target


def test_forward_follows_oracle(oracle_mod):
    print("[case] diff forward vs oracle", flush=True)
    eng, systems, hidden, theta, v0, target = setup_case()
    eng.set_parameters(hidden, v0)
    rec = eng.record()
    off = eng.offsets
    for s, wl in enumerate(systems):
        o = oracle_mod.Oracle(wl)
        o.prime()
        last = 0
        x0 = wl["particles"]["x"]
        for k, step in enumerate(OBS):
            o.step(step - last, DT)
            last = step
            xo = o.state()["x"]
            xg = rec[k, off[s]:off[s + 1]]
            assert rel_err(xg - x0, xo - x0) <= 1e-9, (s, step)
    eng.close()


def _fd_theta(eng, theta, s, c, h):
    tp, tm = theta.copy(), theta.copy()
    tp[s, c] += h
    tm[s, c] -= h
    eng.set_parameters(tp)
    lp = eng.loss()[s]
    eng.set_parameters(tm)
    lm = eng.loss()[s]
    eng.set_parameters(theta)
    return (lp - lm) / (2 * h)


def test_theta_gradient_matches_central_differences():
    print("[case] adjoint d/dtheta vs central differences (GPU loss)", flush=True)
    eng, systems, hidden, theta, v0, target = setup_case()
    loss, gth, gv0 = eng.gradient()
    assert np.all(loss > 0)
    for s in range(len(systems)):
        for c in range(3):
            h = 1e-6 * theta[s, c]
            fd = _fd_theta(eng, theta, s, c, h)
            tol = 1e-5 * abs(fd) + 1e-9 * loss[s] / theta[s, c]
            assert abs(gth[s, c] - fd) <= tol, (s, c, gth[s, c], fd)
    eng.close()


def test_v0_gradient_matches_central_differences():
    print("[case] adjoint d/dv0 vs central differences", flush=True)
    eng, systems, hidden, theta, v0, target = setup_case(nsys=2)
    loss, gth, gv0 = eng.gradient()
    off = eng.offsets
    for s in range(len(systems)):
        block = np.abs(gv0[off[s]:off[s + 1]])
        picks = np.argsort(block.reshape(-1))[-4:]
        for flat in picks:
            i, c = off[s] + flat // 3, flat % 3
            h = 1e-6
            vp, vm = v0.copy(), v0.copy()
            vp[i, c] += h
            vm[i, c] -= h
            eng.set_parameters(None, vp)
            lp = eng.loss()[s]
            eng.set_parameters(None, vm)
            lm = eng.loss()[s]
            eng.set_parameters(None, v0)
            fd = (lp - lm) / (2 * h)
            assert abs(gv0[i, c] - fd) <= 1e-5 * abs(fd) + 1e-9 * loss[s], (s, i, c, gv0[i, c], fd)
    eng.close()


def test_theta_gradient_matches_oracle_differences(oracle_mod):
    """Central differences of the loss computed by the CPU oracle."""
    print("[case] adjoint d/dtheta vs central differences of the oracle", flush=True)
    eng, systems, hidden, theta, v0, target = setup_case(nsys=1)
    loss, gth, gv0 = eng.gradient()
    wl = systems[0]
    n = len(wl["particles"]["radius"])

    def oracle_loss(th):
        w = dict(wl)
        hooke = (th[0], th[1], -1.0, th[2], 0.0)
        w["models"] = [(0, 0, 1, hooke), (0, 1, 1, hooke)]
        o = oracle_mod.Oracle(w)
        o.prime()
        last, acc = 0, 0.0
        for k, step in enumerate(OBS):
            o.step(step - last, DT)
            last = step
            d = o.state()["x"] - target[k, :n]
            acc += float((d * d).sum())
        return acc / (len(OBS) * 3 * n)

    assert oracle_loss(theta[0]) == pytest.approx(loss[0], rel=1e-8)
    for c in range(3):
        h = 1e-6 * theta[0, c]
        tp, tm = theta[0].copy(), theta[0].copy()
        tp[c] += h
        tm[c] -= h
        fd = (oracle_loss(tp) - oracle_loss(tm)) / (2 * h)
        assert abs(gth[0, c] - fd) <= 1e-5 * abs(fd) + 1e-9 * loss[0] / theta[0, c], (c, gth[0, c], fd)
    eng.close()


def test_autograd_harness_and_determinism():
    print("[case] torch.autograd harness", flush=True)
    import torch
    from pacific_amd.calibration import trajectory_loss
    eng, systems, hidden, theta, v0, target = setup_case(nsys=2)
    _, gth, gv0 = eng.gradient()
    _, gth2, gv02 = eng.gradient()
    assert np.array_equal(gth, gth2) and np.array_equal(gv0, gv02)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    vv = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    loss = trajectory_loss(th, vv, eng)
    w = torch.tensor([2.0, 0.5], dtype=torch.float64)
    (loss * w).sum().backward()
    np.testing.assert_array_equal(th.grad.numpy(), gth * w.numpy()[:, None])
    off = eng.offsets
    np.testing.assert_array_equal(vv.grad.numpy()[off[1]:off[2]], gv0[off[1]:off[2]] * 0.5)
    with torch.no_grad():
        l2 = trajectory_loss(th, vv, eng)
    np.testing.assert_array_equal(l2.numpy(), loss.detach().numpy())
    eng.close()
