"""Shared helpers of the GPU parity tests: comparisons against the oracle."""
import numpy as np


def pair_set(c):
    """Unordered contact set; a wall partner (negative id) keeps its side."""
    out = set()
    for a, b in zip(c["id0"].tolist(), c["id1"].tolist()):
        out.add((a, b) if b < 0 else (min(a, b), max(a, b)))
    return out


def force_scales(n, contacts, x, weight):
    """Per-particle scale of the force sum: sum of |contact force| + |weight|
    (the normwise bound used for the 1e-12 check), and of the torque sum:
    sum of |lever| |f|."""
    sf = np.linalg.norm(weight, axis=1).copy()
    st = np.zeros(n)
    fa = np.linalg.norm(contacts["force"], axis=1)
    for k, (a, b) in enumerate(zip(contacts["id0"].tolist(), contacts["id1"].tolist())):
        pt = contacts["point"][k]
        sf[a] += fa[k]
        st[a] += fa[k] * np.linalg.norm(pt - x[a]) * 1.01
        if b >= 0:
            sf[b] += fa[k]
            st[b] += fa[k] * np.linalg.norm(pt - x[b]) * 1.01 + 1e-30
    return sf, st


def normwise_ok(got, want, scale, tol):
    err = np.linalg.norm(got - want, axis=1)
    bad = err > tol * np.maximum(scale, 1e-300)
    return not bad.any(), err, bad


def rel_err(got, want):
    den = np.abs(want).max()
    if den == 0.0:
        return float(np.abs(got).max())
    return float(np.abs(got - want).max() / den)
