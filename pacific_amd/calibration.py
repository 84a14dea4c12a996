"""Trajectory calibration with the differentiated step: a torch.autograd.Function
whose forward is the loss of every system (mean square position error at the
observation steps) and whose backward is the adjoint gradient of
libgdem.so (gdem_diff_gradient) with respect to (kn, en, mu) of each system
and to the initial velocities.

torch is only the autograd plumbing here: parameters and gradients cross the
C ABI as host arrays, the computation is the HIP kernels'.
"""
from __future__ import annotations

import numpy as np
import torch

from . import gdem


class TrajectoryLoss(torch.autograd.Function):
    """loss[s] = mean over observation steps, particles and axes of
    (x - x_target)^2 for system s.

    forward(theta (nsys, 3), v0 (total particles, 3), engine) -> loss (nsys,)
    The gradient is computed by the same launch as the loss when any input
    needs it (checkpointed recomputation inside the launch), so backward only
    scales the stored gradients by the incoming loss gradient."""

    @staticmethod
    def forward(ctx, theta, v0, engine):
        engine.set_parameters(theta.detach().cpu().double().numpy(),
                              v0.detach().cpu().double().numpy())
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            loss, gth, gv0 = engine.gradient()
            ctx.save_for_backward(torch.from_numpy(gth), torch.from_numpy(gv0))
        else:
            loss = engine.loss()
        ctx.offsets = engine.offsets
        return torch.from_numpy(loss).to(theta.device)

    @staticmethod
    def backward(ctx, gloss):
        gth, gv0 = ctx.saved_tensors
        g = gloss.detach().cpu().double()
        gtheta = gth * g[:, None]
        reps = torch.from_numpy(np.diff(ctx.offsets))
        gv = gv0 * torch.repeat_interleave(g, reps)[:, None]
        return gtheta, gv, None


def trajectory_loss(theta, v0, engine):
    return TrajectoryLoss.apply(theta, v0, engine)


def make_engine(systems, nsteps=400, dt=1.0e-5, checkpoint_every=50, obs_steps=(100, 200, 300, 400),
                block_size=256):
    return gdem.GdemDiff(systems, nsteps, dt, checkpoint_every, obs_steps, block_size=block_size)


def initial_velocities(systems):
    return np.concatenate([np.asarray(w["particles"]["v"], float) for w in systems])


def central_difference(f, scale, rel_steps=(1e-4, 1e-5, 1e-6, 1e-7, 1e-8)):
    """Central differences (f(h) - f(-h)) / 2h of a scalar function at h = 0,
    for h = r * scale with each r of rel_steps (largest first).  The loss is
    piecewise smooth (a contact starting or ending, or switching between
    sticking and Coulomb sliding, puts a kink in it) and carries rounding
    noise that a small step amplifies, so no single step suits every
    parameter: the estimate taken is the one whose next smaller step agrees
    with it best.  Returns (estimate, agreement), agreement being the
    relative difference of those two estimates: how well differences resolve
    the derivative there."""
    ds = []
    for r in rel_steps:
        h = r * scale
        ds.append((f(h) - f(-h)) / (2.0 * h))
    best, agree = 0, float("inf")
    for k in range(len(ds) - 1):
        den = max(abs(ds[k]), abs(ds[k + 1]))
        a = abs(ds[k] - ds[k + 1]) / den if den > 0 else 0.0
        if a < agree:
            best, agree = k, a
    return ds[best], agree


def gradient_error(grad, fd, agree, scale, resolved=1e-6):
    """Errors of a gradient against central differences: the normwise
    relative error of the scaled vectors (scale_i * grad_i against
    scale_i * fd_i, i.e. of the changes of the loss over relative parameter
    changes), and the largest componentwise relative error among the
    components the differences resolve (agreement <= resolved).  Returns
    (normwise, componentwise, number of components resolved)."""
    g = np.asarray(grad, float) * np.asarray(scale, float)
    d = np.asarray(fd, float) * np.asarray(scale, float)
    den = np.linalg.norm(d)
    normwise = float(np.linalg.norm(g - d) / den) if den > 0 else float(np.linalg.norm(g))
    ok = np.asarray(agree) <= resolved
    comp = 0.0
    for gi, di, oki in zip(np.asarray(grad, float), np.asarray(fd, float), ok):
        if oki and di != 0.0:
            comp = max(comp, abs(gi - di) / abs(di))
    return normwise, comp, int(ok.sum())
