"""The detection sweep's single-precision pre-filter never rejects a pair the
FP64 contact test accepts.

The kernel rejects a candidate when, in float arithmetic,
(xb + s*L - xa)^2 + ... > (ra + rb + margin)^2 with
margin = 2^-19 (max |coordinate| + max L).  This replays that arithmetic with
numpy float32 on pairs placed at and just inside contact, with coordinates at
the worst magnitudes of a box, and checks that none is rejected.
"""
import numpy as np


def float_reject(xa, xb, ra, rb, shift, L, margin):
    f = np.float32
    d = (xb.astype(f) + (shift * L).astype(f)) - xa.astype(f)
    rs = f(ra) + f(rb) + f(margin)
    return (d * d).sum(axis=-1, dtype=f) > rs * rs


def test_prefilter_keeps_every_exact_contact():
    rng = np.random.default_rng(5)
    for lo, L in ((np.zeros(3), np.full(3, 0.244)), (np.full(3, -3.0), np.full(3, 7.5))):
        mag = np.max(np.maximum(np.abs(lo), np.abs(lo + L)))
        margin = 2.0 ** -19 * (mag + L.max())
        n = 200000
        ra = rng.uniform(0.8e-3, 1.2e-3, n)
        rb = rng.uniform(0.8e-3, 1.2e-3, n)
        u = rng.normal(size=(n, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        # exact separation at or just inside contact
        sep = (ra + rb) * (1.0 - rng.uniform(0.0, 1e-9, n))
        xa = lo + rng.uniform(0.0, 1.0, (n, 3)) * L
        xb = xa + u * sep[:, None]
        shift = np.zeros((n, 3))
        # periodic images: move xb to the other side and shift it back
        wrap = rng.random(n) < 0.5
        side = np.where(xb > lo + L / 2, -1.0, 1.0)
        shift[wrap] = -side[wrap]
        xb[wrap] = xb[wrap] + side[wrap] * L
        exact = np.linalg.norm((xb + shift * L) - xa, axis=1) - (ra + rb)
        keep = exact <= 0.0
        assert keep.sum() > n // 2
        rej = float_reject(xa[keep], xb[keep], ra[keep], rb[keep], shift[keep], L, margin)
        assert not rej.any()
