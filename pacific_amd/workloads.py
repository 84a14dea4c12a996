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

from .gdem import MODEL_HERTZ, MODEL_HERTZ_MEMORY, MODEL_HOOKE, MODEL_HOOKE_MEMORY

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


# row-major bases whose local z (the Rectangle normal) is world axis d, the
# local x and y the next two axes in turn
FACE_BASIS = ((0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0),
              (0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0),
              IDENTITY)


def box_rectangles(lo, hi, material, margin=5.0e-4, crust=CRUST):
    """Six Rectangle obstacles on the faces of [lo, hi] (no thickness), each
    margin wider than the face on every side: (basis, centre, lx, ly, crust,
    material)."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    c = 0.5 * (lo + hi)
    rects = []
    for d in range(3):
        a, b = (d + 1) % 3, (d + 2) % 3
        for side in (-1, 1):
            centre = c.copy()
            centre[d] = lo[d] if side < 0 else hi[d]
            rects.append((FACE_BASIS[d], tuple(centre), hi[a] - lo[a] + 2 * margin,
                          hi[b] - lo[b] + 2 * margin, crust, material))
    return rects


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


# ------------------------------------------------------------------ memory models
def sphere_on_wall(model, v0=(0.0, 0.0, 0.0), w0=(0.0, 0.0, 0.0), r=1.0e-3, dt=1.0e-6,
                   length=0.1):
    """One sphere (density 2500) resting on the top face (z = 0) of a box
    wall, at the static overlap m g / kn of a Hooke model (kn = model[1][0]),
    with velocity v0 and angular velocity w0; the domain runs length along x.
    model = (kind, params) for the sphere-wall pair."""
    kind, prm = model
    m = DENSITY * 4.0 * np.pi * r ** 3 / 3.0
    delta = m * 9.81 / prm[0] if kind in (MODEL_HOOKE, MODEL_HOOKE_MEMORY) else 0.0
    x = np.array([[0.1 * length, 0.01, r - delta]])
    lo, size = (0.0, 0.0, -2.0e-3), (length, 0.02, 0.01)
    wall = (IDENTITY, (0.5 * length, 0.01, -1.0e-3), (length, 0.02, 1.0e-3), CRUST, 1)
    return {"name": "sphere_on_wall", "origin": lo, "size": size, "periodic": (0, 0, 0),
            "cell_size": 2.0 * r, "gravity": (0.0, 0.0, -9.81), "dt": dt, "steps": 1000,
            "models": [(0, 0, kind, prm), (0, 1, kind, prm)], "walls": [wall],
            "particles": _particles(x, np.array([v0], dtype=float), np.array([r]))
            | {"w": np.array([w0], dtype=float)}}


def sphere_on_plate(model, half=(0.05, 0.05), x0=(0.0, 0.0, None), r=1.0e-3, dt=5.0e-6,
                    gravity=-9.81, motions=()):
    """One sphere (density 2500) on a square plate (top face z = 0, centred
    on the z axis, half extents half x 1 mm), at the static overlap of a
    Hooke model when x0[2] is None; motions = imposed plate motions
    (wall 0) as (kind, t_start, t_end, params)."""
    kind, prm = model
    m = DENSITY * 4.0 * np.pi * r ** 3 / 3.0
    delta = m * -gravity / prm[0] if kind in (MODEL_HOOKE, MODEL_HOOKE_MEMORY) else 0.0
    z = r - delta if x0[2] is None else x0[2]
    x = np.array([[x0[0], x0[1], z]])
    hx, hy = half
    wall = (IDENTITY, (0.0, 0.0, -1.0e-3), (hx, hy, 1.0e-3), CRUST, 1)
    return {"name": "sphere_on_plate", "origin": (-hx, -hy, -2.0e-3), "size": (2 * hx, 2 * hy, 0.01),
            "periodic": (0, 0, 0), "cell_size": 2.0 * r, "gravity": (0.0, 0.0, gravity), "dt": dt,
            "models": [(0, 0, kind, prm), (0, 1, kind, prm)], "walls": [wall], "time": 0.0,
            "motions": [(0,) + tuple(mo) for mo in motions],
            "particles": _particles(x, np.zeros((1, 3)), np.array([r]))}


def rectangle_box(n=1000, seed=1, memory=False):
    """plain_box (memory_box when memory) closed by six Rectangle obstacles on
    the faces of the domain instead of box walls."""
    wl = memory_box(n, seed) if memory else plain_box(n, seed)
    lo = np.asarray(wl["origin"], float)
    wl["name"] = f"rectangle_box_{n}" + ("_memory" if memory else "")
    wl["rectangles"] = box_rectangles(lo, lo + np.asarray(wl["size"], float), 1)
    wl["walls"] = []
    return wl


def shaken_box(n=1000, seed=1, memory=False):
    """plain_box (HookeMemory between particles and walls when memory) with an
    imposed motion on five walls (wall k of box_walls: 2 d + (side > 0)):
    floor sine in z (0.1 m/s, 5 ms), x-min crenel along +x (0.1 m/s, 2 ms),
    x-max multi-sine inwards (0.2 m/s, 6 ms, phase 180 degrees: the
    direction enters squared, so its sign alone would not turn it), y-min constant 0.05 m/s
    along +y over [0.5, 2.5] ms, y-max rotating at 10 rad/s about x through
    its centre.  The walls reach the lattice (50 um from the floor and x-min,
    250 um from the others) within about a millisecond."""
    wl = memory_box(n, seed) if memory else plain_box(n, seed)
    wl["name"] = f"shaken_box_{n}" + ("_memory" if memory else "")
    wl["time"] = 0.0
    wl["motions"] = [
        (4, 2, 0.0, 1.0, (0.0, 0.0, 1.0, 0.1, 5.0e-3, 0.0)),
        (0, 3, 0.0, 1.0, (1.0, 0.0, 0.0, 0.1, 2.0e-3)),
        (1, 4, 0.0, 1.0, (-1.0, 0.0, 0.0, 0.2, 0.0, 0.0, 6.0e-3, 1.0, 1.0, 180.0, 0.0, 0.0)),
        (2, 1, 5.0e-4, 2.5e-3, (0.0, 0.05, 0.0)),
        (3, 5, 0.0, 1.0, (10.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)),
    ]
    return wl


def piston_box(n=1000, seed=1, memory=False, force=0.05, vmax=0.2, mass=2.5e-5):
    """plain_box (HookeMemory when memory) whose lid (wall 5) starts 50 um
    above the highest sphere and is driven by an imposed force of `force` N
    (ObstacleImposedForce, ConstantTranslation, vmax m/s, controller mass
    `mass` kg): it closes in at vmax, then presses the packing until its load
    is `force`."""
    wl = memory_box(n, seed) if memory else plain_box(n, seed)
    wl["name"] = f"piston_box_{n}" + ("_memory" if memory else "")
    x, r = wl["particles"]["x"], wl["particles"]["radius"]
    top = float((x[:, 2] + r).max())
    basis, centre, half, crust, mat = wl["walls"][5]
    wl["walls"][5] = (basis, (centre[0], centre[1], top + 5.0e-5 + half[2]), half, crust, mat)
    wl["time"] = 0.0
    wl["wall_forces"] = [(5, 0.0, 1.0, (0.0, 0.0, force, mass, vmax))]
    return wl


# axis of the drums along y: the shell's local z is the world's y
DRUM_BASIS = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, -1.0, 0.0)


def rotating_drum(seed=6, radius=0.01, length=0.04, nbox=24, omega=5.0, fill=0.55, memory=False,
                  r=1.0e-3, gap=2.0e-5, elements="box"):
    """Spheres (r 1 mm, density 2500) in a horizontal drum: a box-element
    CylindricalShell (inner radius 1 cm, wall 1 mm, 24 boxes, axis along y)
    and two end caps in one composite (the drum: caps, then the shell's own
    composite) rotating at omega about the drum axis (y); in each 2.2 mm slice along the axis, a ring of sites gap from the
    shell's faces (2.2 mm apart) and a jittered 2.2 mm lattice inside it, in
    the lower part of the drum (z below (2 fill - 1) radius), v ~ N(0,
    0.05^2); Hooke kn 1e5, en 0.9, mu 0.5 (HookeMemory when memory); dt 1e-5.
    elements="rectangle" builds the shell from its default Rectangle
    elements (inner faces at the same radius)."""
    rng = np.random.default_rng(seed)
    spacing = 2.2e-3
    ring = radius - r - gap        # a ring of sites gap from the shell's faces
    rin = ring - spacing           # lattice sites inside it
    zmax = (2.0 * fill - 1.0) * radius
    dth = 2.0 * np.pi / np.floor(2.0 * np.pi * ring / spacing)   # chord >= spacing - tiny
    pts = []
    for y in np.arange(-length / 2 + r + gap + 1e-4, length / 2 - r - gap, spacing):
        for th in np.arange(0.0, 2.0 * np.pi - dth / 2, dth) + 0.5 * dth:
            if ring * np.sin(th) <= zmax:
                pts.append((ring * np.cos(th), y, ring * np.sin(th)))
        for xx in np.arange(-rin, rin + 1e-12, spacing):
            for zz in np.arange(-rin, rin + 1e-12, spacing):
                if xx * xx + zz * zz <= rin * rin and zz <= zmax:
                    pts.append((xx, y, zz))
    pts = np.array(pts)
    jit = rng.uniform(-1.0e-5, 1.0e-5, size=pts.shape)
    # ring sites move only along the shell (no jitter towards it)
    rad = np.hypot(pts[:, 0], pts[:, 2])
    on_ring = rad > rin + 1e-4
    jit[on_ring, 0] = 0.0
    jit[on_ring, 2] = 0.0
    x = pts + jit
    n = len(x)
    v = rng.normal(0.0, 0.05, size=(n, 3))
    width = 1.0e-3
    cap = radius + width + 1.0e-3
    walls = [(IDENTITY, (0.0, -length / 2 - 2.5e-4, 0.0), (cap, 2.5e-4, cap), CRUST, 1),
             (IDENTITY, (0.0, length / 2 + 2.5e-4, 0.0), (cap, 2.5e-4, cap), CRUST, 1)]
    hooke = (1.0e5, 0.9, -1.0, 0.5, 0.0)
    models = ([(0, 0, MODEL_HOOKE_MEMORY, HOOKE_MEMORY), (0, 1, MODEL_HOOKE_MEMORY, HOOKE_MEMORY)]
              if memory else [(0, 0, MODEL_HOOKE, hooke), (0, 1, MODEL_HOOKE, hooke)])
    half = (radius + width + 2 * r, length / 2 + r, radius + width + 2 * r)
    return {"name": f"rotating_drum_{n}" + ("_memory" if memory else ""),
            "origin": tuple(-h for h in half), "size": tuple(2 * h for h in half),
            "periodic": (0, 0, 0), "cell_size": 2.0 * r, "gravity": (0.0, 0.0, -9.81), "dt": 1.0e-5,
            "models": models, "walls": walls, "time": 0.0,
            # composite 0 (the drum) holds the caps and the shell (composite 1)
            "composites": [(-1, [0, 1])],
            ("rect_shells" if elements == "rectangle" else "shells"):
                [(DRUM_BASIS, (0.0, 0.0, 0.0), length, radius, width, nbox, CRUST, 1, 0)],
            "composite_motions": [(0, 5, 0.0, 10.0, (0.0, omega, 0.0, 0.0, 0.0, 0.0, 1.0))],
            "particles": _particles(x, v, np.full(n, r))}


# HookeMemory {kn, en, etat, muc, kt, mur, etarpf, epst} with rolling friction
HOOKE_MEMORY = (1.0e5, 0.9, -1.0, 0.5, 2.0e4, 0.02, 0.5, 0.0)
# HertzMemory {Es, en, Gs, muc, mur, etarpf, epst}
HERTZ_MEMORY = (1.0e7, 0.85, 1.0e7 / 4.0, 0.4, 0.02, 0.5, 0.0)


def memory_box(n=1000, seed=1):
    """plain_box with the HookeMemory model (tangential spring kt 2e4,
    rolling friction mur 0.02, etarpf 0.5) between particles and walls."""
    wl = plain_box(n, seed)
    wl["name"] = f"memory_box_{n}"
    wl["models"] = [(0, 0, MODEL_HOOKE_MEMORY, HOOKE_MEMORY), (0, 1, MODEL_HOOKE_MEMORY, HOOKE_MEMORY)]
    return wl


def memory_periodic(n=20000, seed=2):
    """periodic_hertz with the HertzMemory model (rolling friction mur 0.02)."""
    wl = periodic_hertz(n, seed, steps=100)
    wl["name"] = f"memory_periodic_{n}"
    wl["models"] = [(0, 0, MODEL_HERTZ_MEMORY, HERTZ_MEMORY)]
    return wl


def periodic_hooke(n=2000, seed=12, phi=0.30):
    """Small periodic Hooke variant (parity of the Hooke branch with the minimum image)."""
    wl = periodic_hertz(n, seed, steps=100, phi=phi)
    wl["name"] = f"periodic_hooke_{n}"
    wl["models"] = [(0, 0, MODEL_HOOKE, (1.0e5, 0.9, -1.0, 0.5, 0.0))]
    wl["dt"] = 1.0e-5
    return wl


# ------------------------------------------------------------------ convex
CUBE_HALF = (1.0e-3, 1.0e-3, 1.0e-3)   # Box(2 mm, 2 mm, 2 mm)
TET_EDGE = 2.5e-3
# faces of the regular tetrahedron as Grains3D's shape tool writes them
# (src/Grains3D/Grains/Tools/PrePost/ShapeFile/src/shapefile.cpp, case 4)
TET_FACES = ((0, 1, 2), (0, 3, 1), (2, 1, 3), (2, 3, 0))


def tetrahedron_vertices(edge=TET_EDGE):
    """Regular tetrahedron centred on its centroid, vertices as the shape tool
    places them: three at height -h on a circle of radius l, one at rcirc."""
    import math
    theta = 2.0 * math.atan(math.sqrt(2.0)) - math.pi / 2.0
    rcirc = math.sqrt(3.0 / 8.0) * edge
    h = rcirc * math.sin(theta)
    l = rcirc * math.cos(theta)
    return np.array([[l, 0.0, -h],
                     [l * math.cos(2.0 * math.pi / 3.0), l * math.sin(2.0 * math.pi / 3.0), -h],
                     [l * math.cos(4.0 * math.pi / 3.0), l * math.sin(4.0 * math.pi / 3.0), -h],
                     [0.0, 0.0, rcirc]])


def convex_box(n=200_000, seed=3, steps=300, phi=0.25, height=0.20, overlap=0.0, cell_factor=1.0,
               sigma_v=0.0, attempts_per_particle=40000):
    """n convex particles, alternating cubes (edge 2 mm) and regular tetrahedra
    (edge 2.5 mm), uniform random orientations (Shoemake), random sequential
    addition (exact separating-axis overlap test) at solid fraction phi in a
    square box of the given height with six walls, the cross-section sized
    for n particles; gravity -9.81 z; Hooke kn 2e5, en 0.8, mu 0.5; dt 2e-6.
    Near 0.25 the addition approaches its saturation slowly: 200,000 particles
    take about 2e9 attempts (minutes on a few dozen host threads).  overlap > 0
    lets neighbours overlap by up to that much at the start (contacts from
    step 0)."""
    import math
    tv = tetrahedron_vertices()
    v_cube = 8.0 * CUBE_HALF[0] * CUBE_HALF[1] * CUBE_HALF[2]
    v_tet = TET_EDGE ** 3 / (6.0 * math.sqrt(2.0))
    types = (np.arange(n) % 2).astype(np.int32)   # 0 cube, 1 tetrahedron
    vs = (n - types.sum()) * v_cube + types.sum() * v_tet
    side = math.sqrt(vs / phi / height)
    lo = np.zeros(3)
    hi = np.array([side, side, height])
    x = np.empty((n, 3))
    v = np.empty((n, 3))
    q = np.empty((n, 4))
    basis = np.empty((n, 9))
    faces = np.array(TET_FACES, np.int32).reshape(-1)
    rc = _lib_convex().gdw_rsa_convex(n, _p(types), _p(np.array(CUBE_HALF)),
                                      _p(np.ascontiguousarray(tv)), _p(faces), _p(lo), _p(hi),
                                      -overlap, sigma_v, seed, attempts_per_particle * n + 100000, _p(x), _p(v), _p(q),
                                      _p(basis))
    if rc != 0:
        raise RuntimeError(f"random sequential addition stopped at particle {-1 - rc} of {n}")
    circ = max(math.sqrt(3.0) * CUBE_HALF[0], math.sqrt(3.0 / 8.0) * TET_EDGE)
    hooke = (2.0e5, 0.8, -1.0, 0.5, 0.0)
    parts = {"x": x, "v": v, "w": np.zeros((n, 3)), "q": q, "basis": basis,
             "shape": types, "density": np.full(n, DENSITY), "crust": np.full(n, CRUST),
             "material": np.zeros(n, np.int32)}
    return {"name": f"convex_box_{n}", "origin": tuple(lo), "size": tuple(hi - lo),
            "periodic": (0, 0, 0), "cell_size": 2.0 * cell_factor * circ,
            "gravity": (0.0, 0.0, -9.81), "dt": 2.0e-6, "steps": steps,
            "models": [(0, 0, MODEL_HOOKE, hooke), (0, 1, MODEL_HOOKE, hooke)],
            "walls": box_walls(lo, hi, 1),
            "shapes": [("box", CUBE_HALF), ("polyhedron", tv, TET_FACES)],
            "particles": parts}


def _lib_convex():
    lib = _lib()
    if not hasattr(lib, "_convex_ready"):
        vp, d, i32, u64 = C.c_void_p, C.c_double, C.c_int, C.c_uint64
        lib.gdw_rsa_convex.restype = i32
        lib.gdw_rsa_convex.argtypes = [i32, vp, vp, vp, vp, vp, vp, d, d, u64, C.c_long, vp, vp, vp, vp]
        lib._convex_ready = True
    return lib


# ------------------------------------------------------------------ batched
def stable_substeps(kn, r_min, dt, density=DENSITY, limit=1.0):
    """Substeps of dt that keep a Hooke contact of stiffness kn between the two
    lightest spheres (radius r_min) at omega dt / substeps <= limit, omega =
    sqrt(kn / m*) with m* = m / 2: half the leap-frog's stability bound
    omega dt < 2.  The policy of the batched and calibration workloads for
    stiff systems; the oracle and gdem_step run such a system as
    nsteps * substeps steps of dt / substeps."""
    import math
    m = density * 4.0 * math.pi * r_min ** 3 / 3.0
    omega = math.sqrt(kn / (0.5 * m))
    return max(1, math.ceil(omega * dt / limit))


def batch_system(n, seed, aspect, kn, en, mu, cell_factor=1.0, spacing=2.3e-3, jitter=4.0e-5,
                 sigma_v=0.05):
    """One small settling system: n spheres, r ~ U[0.9, 1.1] mm, on a jittered
    cubic lattice (spacing 2.3 mm) in a closed box whose height is about
    `aspect` times its width, six walls, gravity -9.81 z, Hooke (kn, en, mu)
    for particle-particle and particle-wall contacts, dt 1e-5 s, advanced in
    stable_substeps(kn, r_min, dt) substeps per step."""
    import math
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.9e-3, 1.1e-3, n)
    nx = max(1, math.ceil((n / aspect) ** (1.0 / 3.0)))
    nz = max(1, math.ceil(n / (nx * nx)))
    lo = np.zeros(3)
    hi = np.array([nx + 0.5, nx + 0.5, nz + 0.5]) * spacing
    x = np.empty((n, 3))
    v = np.empty((n, 3))
    rc = _lib().gdw_lattice(n, spacing, _p(lo), _p(hi), jitter, sigma_v, int(seed), _p(x), _p(v))
    if rc < 0:
        raise ValueError("lattice too small")
    hooke = (float(kn), float(en), -1.0, float(mu), 0.0)
    parts = _particles(x, v, r)
    return {"name": f"batch_system_{n}_s{seed}", "origin": tuple(lo), "size": tuple(hi - lo),
            "periodic": (0, 0, 0), "cell_size": 2.0 * cell_factor * r.max(),
            "gravity": (0.0, 0.0, -9.81), "dt": 1.0e-5, "steps": 1000,
            "substeps": stable_substeps(kn, float(r.min()), 1.0e-5),
            "models": [(0, 0, MODEL_HOOKE, hooke), (0, 1, MODEL_HOOKE, hooke)],
            "walls": box_walls(lo, hi, 1), "particles": parts}


def batch_systems(nsys=4096, seed=4, nmin=64, nmax=2048, kn_range=(1.0e4, 1.0e6), cell_factor=1.0):
    """Independent settling systems of a calibration sweep: particle count
    U{nmin..nmax}, aspect ratio U[1, 4], kn log-uniform in kn_range, en
    U[0.5, 0.95], mu U[0.1, 0.8] per system.  For two 0.9 mm spheres (m* =
    3.8e-6 kg) the leap-frog step of 1e-5 s is unstable above kn = 4 m* /
    dt^2 = 1.5e5 N/m (omega dt > 2); each system therefore runs
    stable_substeps() substeps per step (up to 6 at kn = 1e6)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, nsys)
    aspect = rng.uniform(1.0, 4.0, nsys)
    kn = np.exp(rng.uniform(np.log(kn_range[0]), np.log(kn_range[1]), nsys))
    en = rng.uniform(0.5, 0.95, nsys)
    mu = rng.uniform(0.1, 0.8, nsys)
    seeds = rng.integers(1, 2 ** 62, nsys)
    return [batch_system(int(ns[s]), int(seeds[s]), aspect[s], kn[s], en[s], mu[s], cell_factor)
            for s in range(nsys)]


# ------------------------------------------------------------------ calibration
def calibration_system(n=1000, seed=1, spacing=2.1e-3, jitter=4.0e-5, sigma_v=0.2,
                       theta=(5.0e4, 0.8, 0.4)):
    """One calibration system: n spheres of radius 1 mm on a jittered cubic
    lattice (spacing 2.1 mm, so neighbours collide within the first
    milliseconds), v ~ N(0, sigma_v^2), in a closed box of aspect ratio about
    2, gravity -9.81 z, Hooke (kn, en, mu) = theta for particle-particle and
    particle-wall contacts, dt 1e-5 s in stable_substeps(1.5 kn) substeps."""
    import math
    nx = max(1, math.ceil((n / 2.0) ** (1.0 / 3.0)))
    nz = max(1, math.ceil(n / (nx * nx)))
    lo = np.zeros(3)
    hi = np.array([nx + 0.5, nx + 0.5, nz + 0.5]) * spacing
    x = np.empty((n, 3))
    v = np.empty((n, 3))
    rc = _lib().gdw_lattice(n, spacing, _p(lo), _p(hi), jitter, sigma_v, int(seed), _p(x), _p(v))
    if rc < 0:
        raise ValueError("lattice too small")
    r = np.full(n, 1.0e-3)
    hooke = (float(theta[0]), float(theta[1]), -1.0, float(theta[2]), 0.0)
    return {"name": f"calibration_{n}_s{seed}", "origin": tuple(lo), "size": tuple(hi - lo),
            "periodic": (0, 0, 0), "cell_size": 2.0 * r.max(), "gravity": (0.0, 0.0, -9.81),
            "dt": 1.0e-5, "steps": 400,
            # room for a calibration to move kn up to 1.5 times its value
            "substeps": stable_substeps(1.5 * float(theta[0]), float(r.min()), 1.0e-5),
            "models": [(0, 0, MODEL_HOOKE, hooke), (0, 1, MODEL_HOOKE, hooke)],
            "walls": box_walls(lo, hi, 1), "particles": _particles(x, v, r)}


def calibration_systems(nsys=512, n=1000, seed=5, kn_range=(1.0e4, 1.0e6)):
    """Systems of a calibration run and their hidden parameters (kn, en, mu),
    drawn as for batch_systems: kn log-uniform in kn_range, en U[0.5, 0.95],
    mu U[0.1, 0.8]."""
    rng = np.random.default_rng(seed)
    kn = np.exp(rng.uniform(np.log(kn_range[0]), np.log(kn_range[1]), nsys))
    en = rng.uniform(0.5, 0.95, nsys)
    mu = rng.uniform(0.1, 0.8, nsys)
    seeds = rng.integers(1, 2 ** 62, nsys)
    theta = np.stack([kn, en, mu], 1)
    return [calibration_system(n, int(seeds[s]), theta=theta[s]) for s in range(nsys)], theta
