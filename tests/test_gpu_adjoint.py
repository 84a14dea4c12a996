"""Adjoint gradient of calibration trajectories (gdem_diff_*): checked against
central finite differences of the GPU loss, against finite differences of the
CPU oracle's loss, and through the torch.autograd harness.

Bar: 1e-5 relative against central differences, normwise over the gradient
scaled by the parameters (theta_i dL/dtheta_i) and componentwise for every
component the differences resolve.  The loss is piecewise smooth: a contact
starting or ending, or switching between sticking and Coulomb sliding, puts a
kink in it, and rounding noise grows as the step shrinks, so the step is
chosen per component among 1e-4 .. 1e-8 relative as the one whose next
smaller step agrees best (calibration.central_difference); a component is
resolved when those two agree to 1e-6.  Independently of differences, the
reverse sweep is checked against a forward-mode (dual tangent) derivative
along random directions, to rounding, also over horizons where differences
no longer resolve the derivative."""
import numpy as np
import pytest

from gdem_testutil import rel_err
from pacific_amd import calibration, workloads

pytestmark = pytest.mark.gpu

# A short horizon: over longer ones the colliding packing amplifies rounding
# and the loss becomes too rough for central differences to resolve 1e-5
# (the adjoint itself is exact there too).
NSTEPS, K, OBS, DT = 48, 10, (16, 32, 48), 1.0e-5
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
        sub = wl["substeps"]
        for k, step in enumerate(OBS):
            o.step((step - last) * sub, DT / sub)
            last = step
            xo = o.state()["x"]
            xg = rec[k, off[s]:off[s + 1]]
            assert rel_err(xg - x0, xo - x0) <= 1e-9, (s, step)
    eng.close()


@pytest.mark.parametrize("nsteps", [48, 200])
def test_adjoint_equals_forward_mode_directional_derivatives(nsteps):
    """<gradient, direction> from the reverse sweep equals the forward-mode
    (dual tangent) derivative along random directions: both are exact
    derivatives of the same discrete map, so they agree to rounding, also over
    horizons where central differences no longer resolve the derivative."""
    print(f"[case] adjoint vs forward mode, {nsteps} steps", flush=True)
    from pacific_amd import gdem
    systems, hidden = workloads.calibration_systems(3, n=300, seed=21)
    obs = tuple(int(v) for v in np.linspace(nsteps / 3, nsteps, 3).round())
    eng = gdem.GdemDiff(systems, nsteps, DT, 10, obs)
    v0 = np.concatenate([w["particles"]["v"] for w in systems])
    eng.set_parameters(hidden, v0)
    eng.set_targets(eng.record())
    theta = hidden * GUESS[None, :]
    eng.set_parameters(theta, v0)
    loss, gth, gv0 = eng.gradient()
    rng = np.random.default_rng(3)
    off = eng.offsets
    for trial in range(3):
        dth = rng.normal(size=theta.shape) * theta
        dv = rng.normal(size=v0.shape) * (0.0 if trial == 0 else 1.0)
        if trial == 2:
            dth[:] = 0.0
        loss2, dl = eng.directional(dth, dv)
        np.testing.assert_array_equal(loss2, loss)
        for s in range(len(systems)):
            want = float((gth[s] * dth[s]).sum() + (gv0[off[s]:off[s + 1]] * dv[off[s]:off[s + 1]]).sum())
            scale = float(np.abs(gth[s] * dth[s]).sum() + np.abs(gv0[off[s]:off[s + 1]] * dv[off[s]:off[s + 1]]).sum())
            assert abs(dl[s] - want) <= 1e-10 * scale, (trial, s, dl[s], want)
    eng.close()


def _check(grad, fd, agree, scale, label):
    """Normwise over the parameter-scaled vector and componentwise where
    differences resolve the derivative: both within 1e-5 relative."""
    normwise, comp, nres = calibration.gradient_error(grad, fd, agree, scale)
    print(f"[case] {label}: normwise {normwise:.2e}, resolved components {nres} "
          f"(worst {comp:.2e}), agreement {np.array(agree)}", flush=True)
    assert normwise <= 1e-5, (label, normwise, grad, fd)
    assert comp <= 1e-5, (label, comp, grad, fd, agree)
    assert nres >= 1, (label, agree)


def test_theta_gradient_matches_central_differences():
    print("[case] adjoint d/dtheta vs central differences (GPU loss)", flush=True)
    eng, systems, hidden, theta, v0, target = setup_case()
    loss, gth, gv0 = eng.gradient()
    assert np.all(loss > 0)
    for s in range(len(systems)):
        fd, agree = np.zeros(3), np.zeros(3)
        for c in range(3):
            def f(h):
                t = theta.copy()
                t[s, c] += h
                eng.set_parameters(t)
                return eng.loss()[s]
            fd[c], agree[c] = calibration.central_difference(f, theta[s, c])
        eng.set_parameters(theta)
        _check(gth[s], fd, agree, theta[s], f"system {s}")
    eng.close()


def test_v0_gradient_matches_central_differences():
    print("[case] adjoint d/dv0 vs central differences", flush=True)
    eng, systems, hidden, theta, v0, target = setup_case(nsys=2)
    loss, gth, gv0 = eng.gradient()
    off = eng.offsets
    for s in range(len(systems)):
        block = np.abs(gv0[off[s]:off[s + 1]])
        picks = np.argsort(block.reshape(-1))[-4:]
        g, fd, agree = np.zeros(4), np.zeros(4), np.zeros(4)
        for j, flat in enumerate(picks):
            i, c = off[s] + flat // 3, flat % 3

            def f(h):
                vv = v0.copy()
                vv[i, c] += h
                eng.set_parameters(None, vv)
                return eng.loss()[s]
            g[j] = gv0[i, c]
            fd[j], agree[j] = calibration.central_difference(f, 0.2)
            eng.set_parameters(None, v0)
        _check(g, fd, agree, np.full(4, 0.2), f"system {s} v0")
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
        sub = wl["substeps"]
        for k, step in enumerate(OBS):
            o.step((step - last) * sub, DT / sub)
            last = step
            d = o.state()["x"] - target[k, :n]
            acc += float((d * d).sum())
        return acc / (len(OBS) * 3 * n)

    assert oracle_loss(theta[0]) == pytest.approx(loss[0], rel=1e-8)
    fd, agree = np.zeros(3), np.zeros(3)
    for c in range(3):
        def f(h):
            t = theta[0].copy()
            t[c] += h
            return oracle_loss(t)
        fd[c], agree[c] = calibration.central_difference(f, theta[0, c])
    _check(gth[0], fd, agree, theta[0], "oracle differences")
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
