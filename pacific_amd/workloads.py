"""Seeded synthetic workloads of the DEM step.

Each builder returns a plain dict (domain, grid request, contact models, walls,
particles, dt, steps) consumed by ``gdem.build_system`` and by the CPU oracle's
binding, so both sides start from bit-identical inputs.  Parameters the
workload descriptions leave open are fixed here:

* crust thickness 5e-5 m (mandatory in the reference, Grains.cpp:877-883);
* Hertz E* = 1e7 Pa and G* = E*/4 (the reference's default,
  HertzContactForceModel.cpp:193-196) for the periodic 1M-sphere workload;
* tangential damping etat = -1 (gamma_t = gamma_n) and no rolling resistance.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .gdem import MODEL_HERTZ, MODEL_HOOKE

_HERE = os.path.dirname(os.path.abspath(__file__))
_WL_PATH = os.path.join(_HERE, "lib", "libgdem_workloads.so")
_wl = None

CRUST = 5.0e-5
DENSITY = 2500.0
IDENTITY = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


def _lib():
    global _wl
    if _wl is None:
        if not os.path.exists(_WL_PATH):
            raise OSError(f"{_WL_PATH} is missing: build it with __graft_entry__.build()")
        lib = C.CDLL(_WL_PATH)
        vp, d, i32, u64 = C.c_void_p, C.c_double, C.c_int, C.c_uint64
        lib.gdw_lattice.restype = i32
        lib.gdw_lattice.argtypes = [i32, d, vp, vp, d, d, u64, vp, vp]
        lib.gdw_radii.restype = d
        lib.gdw_radii.argtypes = [i32, d, d, d, u64, vp]
        lib.gdw_rsa_periodic.restype = i32
        lib.gdw_rsa_periodic.argtypes = [i32, vp, d, d, u64, C.c_long, vp, vp]
        lib.gdw_random_quaternions.restype = None
        lib.gdw_random_quaternions.argtypes = [i32, u64, vp]
        lib.gdw_uniform.restype = None
        lib.gdw_uniform.argtypes = [i32, d, d, i32, u64, vp]
        _wl = lib
    return _wl


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def box_walls(lo, hi, material, thickness=5.0e-4, margin=5.0e-4, crust=CRUST):
    """Six thin box walls just outside [lo, hi], as the reference's settling
    test sets them (src/Grains3D/Tests/data/spheres-settling-new-grains/Grains/
    Init/insert.xml:47-98: 0.5 mm thick, 0.5 mm wider than the domain)."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    c = 0.5 * (lo + hi)
    h = 0.5 * (hi - lo) + margin
    t = 0.5 * thickness
    walls = []
    for d in range(3):
        for side in (-1, 1):
            centre = c.copy()
            half = h.copy()
            centre[d] = (lo[d] - t) if side < 0 else (hi[d] + t)
            half[d] = t
            walls.append((IDENTITY, tuple(centre), tuple(half), crust, material))
    return walls


def _particles(x, v, radius, seed_q=None):
    n = len(radius)
    q = np.zeros((n, 4))
    q[:, 3] = 1.0
    if seed_q is not None:
        _lib().gdw_random_quaternions(n, seed_q, _p(q))
    return {"x": x, "v": v, "w": np.zeros((n, 3)), "q": q, "radius": radius,
            "density": np.full(n, DENSITY), "crust": np.full(n, CRUST),
            "material": np.zeros(n, np.int32)}


def plain_box(n=1000, seed=1, steps=2000):
    """1,000 monodisperse spheres (r = 1 mm) in a closed 2x2x4 cm box with six
    walls; jittered cubic lattice (2.2 mm, +-0.05 mm), v ~ N(0, 0.05^2), gravity
    -9.81 z; Hooke kn 1e5, en 0.9, mu 0.5; dt 1e-5."""
    lo, hi = (0.0, 0.0, 0.0), (0.02, 0.02, 0.04)
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    rc = _lib().gdw_lattice(n, 2.2e-3, _p(np.array(lo)), _p(np.array(hi)), 5.0e-5, 0.05, seed,
                            _p(x), _p(v))
    if rc < 0:
        raise ValueError("lattice too small")
    r = np.full(n, 1.0e-3)
    hooke = (1.0e5, 0.9, -1.0, 0.5, 0.0)
    return {"name": f"plain_box_{n}", "origin": lo, "size": tuple(np.subtract(hi, lo)),
            "periodic": (0, 0, 0), "cell_size": 2.0 * 1.0 * r.max(),
            "gravity": (0.0, 0.0, -9.81), "dt": 1.0e-5, "steps": steps,
            "models": [(0, 0, MODEL_HOOKE, hooke), (0, 1, MODEL_HOOKE, hooke)],
            "walls": box_walls(lo, hi, 1), "particles": _particles(x, v, r)}


def periodic_hertz(n=1_000_000, seed=2, steps=500, phi=0.30, cell_factor=1.2):
    """n polydisperse spheres, r ~ U[0.8, 1.2] mm, triply periodic cube at solid
    fraction 0.30, random sequential addition, v ~ N(0, 0.1^2), no gravity;
    Hertz en 0.85, mu 0.4; cell = 2 * 1.2 * r_max; dt 5e-6."""
    r = np.empty(n)
    L = _lib().gdw_radii(n, 0.8e-3, 1.2e-3, phi, seed, _p(r))
    x = np.empty((n, 3))
    v = np.empty((n, 3))
    rc = _lib().gdw_rsa_periodic(n, _p(r), L, 0.1, seed, 400 * n + 100000, _p(x), _p(v))
    if rc != 0:
        raise RuntimeError("random sequential addition did not finish")
    hertz = (1.0e7, 0.85, 1.0e7 / 4.0, 0.4, 0.0)
    return {"name": f"periodic_hertz_{n}", "origin": (0.0, 0.0, 0.0), "size": (L, L, L),
            "periodic": (1, 1, 1), "cell_size": 2.0 * cell_factor * r.max(),
            "gravity": (0.0, 0.0, 0.0), "dt": 5.0e-6, "steps": steps,
            "models": [(0, 0, MODEL_HERTZ, hertz)], "walls": [],
            "particles": _particles(x, v, r)}


def periodic_hooke(n=2000, seed=12, phi=0.30):
    """Small periodic Hooke variant (parity of the Hooke branch with the minimum image)."""
    wl = periodic_hertz(n, seed, steps=100, phi=phi)
    wl["name"] = f"periodic_hooke_{n}"
    wl["models"] = [(0, 0, MODEL_HOOKE, (1.0e5, 0.9, -1.0, 0.5, 0.0))]
    wl["dt"] = 1.0e-5
    return wl
