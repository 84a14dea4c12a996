"""ctypes binding of the CPU oracle (oracle/build/libgdem_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libgdem_oracle.so")

OK, ERR_CONTACT, ERR_MOTION, ERR_SIMULATION, _, ERR_ARG = range(6)
MODEL_HOOKE, MODEL_HERTZ = 1, 2


class _Config(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("size", C.c_double * 3),
                ("periodic", C.c_int * 3), ("cell_size", C.c_double),
                ("gravity", C.c_double * 3)]


class OracleError(RuntimeError):
    def __init__(self, code, what, id0, id1):
        super().__init__(f"{what}: oracle error {code} (ids {id0}, {id1})")
        self.code, self.id0, self.id1 = code, id0, id1


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        vp, d, i32 = C.c_void_p, C.c_double, C.c_int
        P = C.POINTER
        sig = {
            "orc_create": (i32, [P(_Config), P(vp)]),
            "orc_destroy": (None, [vp]),
            "orc_set_particles": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp]),
            "orc_add_wall": (i32, [vp, vp, vp, vp, d, i32]),
            "orc_set_contact_model": (i32, [vp, i32, i32, i32, vp]),
            "orc_set_time_step": (i32, [vp, d]),
            "orc_add_wall_motion": (i32, [vp, i32, i32, d, d, vp]),
            "orc_get_wall_loads": (i32, [vp, vp, vp]),
            "orc_add_wall_force": (i32, [vp, i32, d, d, vp]),
            "orc_add_composite_force": (i32, [vp, i32, d, d, vp]),
            "orc_set_motion_restriction": (i32, [vp, i32, i32, i32]),
            "orc_add_composite": (i32, [vp, i32, P(i32)]),
            "orc_attach_wall": (i32, [vp, i32, i32]),
            "orc_add_composite_motion": (i32, [vp, i32, i32, d, d, vp]),
            "orc_get_composite": (i32, [vp, i32, vp, vp, vp]),
            "orc_cylindrical_shell_boxes": (None, [vp, vp, d, d, d, i32, vp, vp, vp]),
            "orc_add_cylindrical_shell": (i32, [vp, vp, vp, d, d, d, i32, d, i32, i32, P(i32)]),
            "orc_add_rectangle_wall": (i32, [vp, vp, vp, d, d, d, i32]),
            "orc_cylindrical_shell_rectangles": (None, [vp, vp, d, d, d, i32, vp, vp, vp]),
            "orc_add_cylindrical_shell_rectangles": (i32, [vp, vp, vp, d, d, d, i32, d, i32, i32,
                                                           P(i32)]),
            "orc_set_time": (i32, [vp, d]),
            "orc_get_time": (d, [vp]),
            "orc_move_walls": (i32, [vp, d, d]),
            "orc_get_wall": (i32, [vp, i32, vp, vp, vp, vp]),
            "orc_get_memory": (i32, [vp, i32, i32, vp, vp, vp, vp]),
            "orc_rotate_two_vectors": (None, [vp, vp, vp, vp]),
            "orc_link": (i32, [vp]),
            "orc_link_update": (i32, [vp]),
            "orc_compute_forces": (i32, [vp]),
            "orc_move": (i32, [vp, d, d]),
            "orc_advance_velocity": (i32, [vp, d]),
            "orc_prime": (i32, [vp]),
            "orc_step": (i32, [vp, i32, d]),
            "orc_step_parallel": (i32, [vp, i32, d, i32]),
            "orc_get_state": (i32, [vp, vp, vp, vp, vp, vp, vp]),
            "orc_set_state": (i32, [vp, vp, vp, vp, vp, vp, vp]),
            "orc_get_grid": (i32, [vp, vp, vp, vp]),
            "orc_enable_contact_log": (None, [vp, i32]),
            "orc_get_contacts": (i32, [vp, i32, P(i32), vp, vp, vp, vp, vp, vp]),
            "orc_last_error": (i32, [vp, P(i32), P(C.c_int32), P(C.c_int32), P(d), P(d)]),
            "orc_candidates": (C.c_int64, [vp]),
            "orc_get_contact_stats": (None, [vp, vp]),
            "orc_sphere_sphere_retry": (i32, [vp, d, d, vp, d, d, vp, vp, P(d), P(i32)]),
            "orc_sphere_sphere": (i32, [vp, d, d, vp, d, d, vp, vp, P(d)]),
            "orc_sphere_box": (i32, [vp, d, d, vp, vp, vp, d, vp, vp, P(d)]),
            "orc_sphere_rectangle": (i32, [vp, d, vp, vp, vp, d, d, vp, vp, P(d)]),
            "orc_contact_force": (None, [i32, vp, vp, vp, d, vp, vp, vp, d, d, vp, vp, vp, d, d, vp,
                                         vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape))


class Oracle:
    """One oracle system built from a workload dict (pacific_amd.workloads)."""

    def __init__(self, wl, log_capacity=0):
        self.lib = load()
        cfg = _Config()
        for d in range(3):
            cfg.origin[d] = wl["origin"][d]
            cfg.size[d] = wl["size"][d]
            cfg.periodic[d] = int(wl["periodic"][d])
            cfg.gravity[d] = wl["gravity"][d]
        cfg.cell_size = wl["cell_size"]
        h = C.c_void_p()
        rc = self.lib.orc_create(C.byref(cfg), C.byref(h))
        if rc:
            raise OracleError(rc, "orc_create", -1, -1)
        self.h = h
        for (a, b, kind, p) in wl["models"]:
            npar = {3: 8, 4: 7}.get(int(kind), 5)
            self._chk(self.lib.orc_set_contact_model(h, a, b, kind, _p(_f(p, (npar,)))), "model")
        if "dt" in wl:
            self._chk(self.lib.orc_set_time_step(h, float(wl["dt"])), "time step")
        self.nwalls = (len(wl["walls"]) + len(wl.get("rectangles", []))
                       + sum(int(sh[5]) for sh in wl.get("shells", []) + wl.get("rect_shells", [])))
        for (basis, centre, half, crust, mat) in wl["walls"]:
            self._chk(self.lib.orc_add_wall(h, _p(_f(basis, (9,))), _p(_f(centre, (3,))),
                                            _p(_f(half, (3,))), crust, mat), "wall")
        for (basis, centre, lx, ly, crust, mat) in wl.get("rectangles", []):
            self._chk(self.lib.orc_add_rectangle_wall(h, _p(_f(basis, (9,))), _p(_f(centre, (3,))),
                                                      float(lx), float(ly), float(crust), int(mat)),
                      "rectangle")
        p = wl["particles"]
        n = len(p["radius"])
        self.n = n
        keep = [_f(p["x"], (n, 3)), _f(p["v"], (n, 3)), _f(p["w"], (n, 3)), _f(p["q"], (n, 4)),
                _f(p["radius"], (n,)), _f(p["density"], (n,)), _f(p["crust"], (n,))]
        mat = np.ascontiguousarray(p["material"], dtype=np.int32)
        self._chk(self.lib.orc_set_particles(h, n, *[_p(a) for a in keep], _p(mat)), "particles")
        if "time" in wl:
            self.lib.orc_set_time(h, float(wl["time"]))
        for (parent, members) in wl.get("composites", []):
            k = C.c_int()
            self._chk(self.lib.orc_add_composite(h, int(parent), C.byref(k)), "composite")
            for w in members:
                self._chk(self.lib.orc_attach_wall(h, int(w), k.value), "attach")
        for (basis, centre, height, radius, width, n, crust, mat, parent) in wl.get("shells", []):
            k = C.c_int()
            self._chk(self.lib.orc_add_cylindrical_shell(
                h, _p(_f(basis, (9,))), _p(_f(centre, (3,))), float(height), float(radius),
                float(width), int(n), float(crust), int(mat), int(parent), C.byref(k)), "shell")
        for (basis, centre, height, radius, width, n, crust, mat, parent) in wl.get("rect_shells",
                                                                                    []):
            k = C.c_int()
            self._chk(self.lib.orc_add_cylindrical_shell_rectangles(
                h, _p(_f(basis, (9,))), _p(_f(centre, (3,))), float(height), float(radius),
                float(width), int(n), float(crust), int(mat), int(parent), C.byref(k)), "shell")
        for (wall, comp) in wl.get("attach", []):
            self._chk(self.lib.orc_attach_wall(h, int(wall), int(comp)), "attach")
        for (wall, comp, mask) in wl.get("restrictions", []):
            self._chk(self.lib.orc_set_motion_restriction(h, int(wall), int(comp), int(mask)),
                      "restriction")
        for (wall, kind, t0, t1, prm) in wl.get("motions", []):
            self.add_wall_motion(wall, kind, t0, t1, prm)
        for (comp, kind, t0, t1, prm) in wl.get("composite_motions", []):
            npar = {1: 3, 2: 6, 3: 5, 4: 12, 5: 7}[int(kind)]
            self._chk(self.lib.orc_add_composite_motion(h, int(comp), int(kind), float(t0), float(t1),
                                                        _p(_f(prm, (npar,)))), "composite motion")
        for (wall, t0, t1, prm) in wl.get("wall_forces", []):
            self._chk(self.lib.orc_add_wall_force(h, int(wall), float(t0), float(t1),
                                                  _p(_f(prm, (5,)))), "wall force")
        for (comp, t0, t1, prm) in wl.get("composite_forces", []):
            self._chk(self.lib.orc_add_composite_force(h, int(comp), float(t0), float(t1),
                                                       _p(_f(prm, (5,)))), "composite force")
        if log_capacity:
            self.lib.orc_enable_contact_log(h, int(log_capacity))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def _chk(self, rc, what):
        if rc:
            code, a, b = C.c_int(), C.c_int32(), C.c_int32()
            v, l = C.c_double(), C.c_double()
            self.lib.orc_last_error(self.h, C.byref(code), C.byref(a), C.byref(b), C.byref(v), C.byref(l))
            raise OracleError(rc, what, a.value, b.value)

    def prime(self):
        self._chk(self.lib.orc_prime(self.h), "prime")

    def step(self, n, dt):
        self._chk(self.lib.orc_step(self.h, int(n), float(dt)), "step")

    def step_parallel(self, n, dt, threads):
        """orc_step on `threads` OpenMP threads (memoryless models, static walls)."""
        self._chk(self.lib.orc_step_parallel(self.h, int(n), float(dt), int(threads)),
                  "step_parallel")

    def link(self):
        self._chk(self.lib.orc_link(self.h), "link")

    def link_update(self):
        self._chk(self.lib.orc_link_update(self.h), "link_update")

    def compute_forces(self):
        self._chk(self.lib.orc_compute_forces(self.h), "compute_forces")

    def move(self, dtv, dtd):
        self._chk(self.lib.orc_move(self.h, dtv, dtd), "move")

    def advance_velocity(self, dtv):
        self._chk(self.lib.orc_advance_velocity(self.h, dtv), "advance_velocity")

    def add_wall_motion(self, wall, kind, t0, t1, prm):
        npar = {1: 3, 2: 6, 3: 5, 4: 12, 5: 7}[int(kind)]
        self._chk(self.lib.orc_add_wall_motion(self.h, int(wall), int(kind), float(t0), float(t1),
                                               _p(_f(prm, (npar,)))), "wall motion")

    def move_walls(self, time, dt):
        self._chk(self.lib.orc_move_walls(self.h, float(time), float(dt)), "move walls")

    def wall_loads(self):
        nw = self.nwalls
        f, t = np.zeros((max(nw, 1), 3)), np.zeros((max(nw, 1), 3))
        self.lib.orc_get_wall_loads(self.h, _p(f), _p(t))
        return f[:nw], t[:nw]

    def composite(self, k):
        c, v, w = np.empty(3), np.empty(3), np.empty(3)
        self._chk(self.lib.orc_get_composite(self.h, int(k), _p(c), _p(v), _p(w)), "composite")
        return c, v, w

    def time(self):
        return self.lib.orc_get_time(self.h)

    def wall(self, k):
        c, b, v, w = np.empty(3), np.empty(9), np.empty(3), np.empty(3)
        self._chk(self.lib.orc_get_wall(self.h, int(k), _p(c), _p(b), _p(v), _p(w)), "wall")
        return c, b, v, w

    def set_time_step(self, dt):
        self._chk(self.lib.orc_set_time_step(self.h, float(dt)), "time step")

    def memory(self, p, cap=64):
        """Contact map of particle p: {key: (kdelta, normal, spring torque)}."""
        key = np.empty(cap, np.int32)
        kd, nrm, spr = (np.empty((cap, 3)) for _ in range(3))
        n = self.lib.orc_get_memory(self.h, int(p), cap, _p(key), _p(kd), _p(nrm), _p(spr))
        return {int(key[k]): (kd[k].copy(), nrm[k].copy(), spr[k].copy()) for k in range(min(n, cap))}

    def state(self):
        n = self.n
        out = {k: np.empty((n, 3)) for k in ("x", "v", "w", "F", "T")}
        out["q"] = np.empty((n, 4))
        self.lib.orc_get_state(self.h, _p(out["x"]), _p(out["v"]), _p(out["w"]), _p(out["q"]),
                               _p(out["F"]), _p(out["T"]))
        return out

    def set_state(self, x=None, v=None, w=None, q=None, F=None, T=None):
        n = self.n
        arrs = [None if a is None else _f(a, (n, k)) for a, k in
                ((x, 3), (v, 3), (w, 3), (q, 4), (F, 3), (T, 3))]
        self.lib.orc_set_state(self.h, *[_p(a) for a in arrs])

    def grid(self):
        nc = np.zeros(3, np.int32)
        e = np.zeros(3)
        o = np.zeros(3)
        self.lib.orc_get_grid(self.h, _p(nc), _p(e), _p(o))
        return nc, e, o

    def contacts(self, cap=10_000_000):
        n = C.c_int()
        self.lib.orc_get_contacts(self.h, 0, C.byref(n), None, None, None, None, None, None)
        m = n.value
        res = {"id0": np.empty(m, np.int32), "id1": np.empty(m, np.int32),
               "point": np.empty((m, 3)), "ovec": np.empty((m, 3)), "overlap": np.empty(m),
               "force": np.empty((m, 3))}
        self.lib.orc_get_contacts(self.h, m, C.byref(n), _p(res["id0"]), _p(res["id1"]),
                                  _p(res["point"]), _p(res["ovec"]), _p(res["overlap"]),
                                  _p(res["force"]))
        return res

    def contact_stats(self):
        """AppCollision's contact features of the run: overlap max, mean, time of
        the max, mean GJK iterations, and the numbers of overlaps and GJK
        reports counted."""
        out = np.zeros(6)
        self.lib.orc_get_contact_stats(self.h, _p(out))
        return {"overlap_max": out[0], "overlap_mean": out[1], "time_overlap_max": out[2],
                "gjk_mean": out[3], "overlaps": int(out[4]), "gjk_reports": int(out[5])}

    def candidates(self):
        return self.lib.orc_candidates(self.h)


def cylindrical_shell_boxes(basis, centre, height, radius, width, n):
    lib = load()
    b, c, h = np.empty((n, 9)), np.empty((n, 3)), np.empty((n, 3))
    lib.orc_cylindrical_shell_boxes(_p(_f(basis, (9,))), _p(_f(centre, (3,))), float(height),
                                    float(radius), float(width), int(n), _p(b), _p(c), _p(h))
    return b, c, h


def cylindrical_shell_rectangles(basis, centre, height, radius, width, n):
    lib = load()
    b, c, h = np.empty((n, 9)), np.empty((n, 3)), np.empty((n, 3))
    lib.orc_cylindrical_shell_rectangles(_p(_f(basis, (9,))), _p(_f(centre, (3,))),
                                         float(height), float(radius), float(width), int(n),
                                         _p(b), _p(c), _p(h))
    return b, c, h


def rotate_two_vectors(u, v, x):
    """Quaternion::setRotFromTwoVectors(u, v) then rotateVector(x)."""
    out = np.empty(3)
    load().orc_rotate_two_vectors(_p(_f(u, (3,))), _p(_f(v, (3,))), _p(_f(x, (3,))), _p(out))
    return out


def sphere_sphere_retry(a, ra, cra, b, rb, crb):
    """ClosestPoint_ErreurHandling of two spheres: (rc, point, overlap vector,
    overlap, GJK iterations)."""
    lib = load()
    pt, ov = np.zeros(3), np.zeros(3)
    dist, it = C.c_double(), C.c_int()
    rc = lib.orc_sphere_sphere_retry(_p(_f(a, (3,))), ra, cra, _p(_f(b, (3,))), rb, crb, _p(pt),
                                     _p(ov), C.byref(dist), C.byref(it))
    return rc, pt, ov, dist.value, it.value


def sphere_sphere(a, ra, cra, b, rb, crb):
    lib = load()
    pt, ov = np.zeros(3), np.zeros(3)
    d = C.c_double()
    rc = lib.orc_sphere_sphere(_p(_f(a, (3,))), ra, cra, _p(_f(b, (3,))), rb, crb, _p(pt), _p(ov),
                               C.byref(d))
    return rc, pt, ov, d.value


def sphere_box(centre, r, cr, basis, origin, half, crb):
    lib = load()
    pt, ov = np.zeros(3), np.zeros(3)
    d = C.c_double()
    rc = lib.orc_sphere_box(_p(_f(centre, (3,))), r, cr, _p(_f(basis, (9,))), _p(_f(origin, (3,))),
                            _p(_f(half, (3,))), crb, _p(pt), _p(ov), C.byref(d))
    return rc, pt, ov, d.value


def sphere_rectangle(centre, r, sphere_basis, basis, origin, lx, ly):
    lib = load()
    pt, ov = np.zeros(3), np.zeros(3)
    d = C.c_double()
    rc = lib.orc_sphere_rectangle(_p(_f(centre, (3,))), r, _p(_f(sphere_basis, (9,))),
                                  _p(_f(basis, (9,))), _p(_f(origin, (3,))), float(lx), float(ly),
                                  _p(pt), _p(ov), C.byref(d))
    return rc, pt, ov, d.value


def contact_force(kind, params, point, ovec, dist, x0, v0, w0, m0, req0, x1, v1, w1, m1, req1,
                  with_moment=False):
    """Force on body 0; with_moment: (force, rolling-resistance moment on body 0)."""
    lib = load()
    f = np.zeros(3)
    mo = np.zeros(3)
    lib.orc_contact_force(kind, _p(_f(params, (5,))), _p(_f(point, (3,))), _p(_f(ovec, (3,))), dist,
                          _p(_f(x0, (3,))), _p(_f(v0, (3,))), _p(_f(w0, (3,))), m0, req0,
                          _p(_f(x1, (3,))), _p(_f(v1, (3,))), _p(_f(w1, (3,))), m1, req1, _p(f),
                          _p(mo))
    return (f, mo) if with_moment else f


# --------------------------------------------------------------------------- convex
SHAPE_BOX, SHAPE_POLYHEDRON = 1, 2


def _load_convex(lib):
    if getattr(lib, "_convex_sig", False):
        return lib
    vp, d, i32 = C.c_void_p, C.c_double, C.c_int
    P = C.POINTER
    sig = {
        "orc_c_create": (i32, [P(_Config), P(vp)]),
        "orc_c_destroy": (None, [vp]),
        "orc_c_add_box_shape": (i32, [vp, vp]),
        "orc_c_add_sphere_shape": (i32, [vp, d]),
        "orc_c_add_polyhedron_shape": (i32, [vp, i32, vp, i32, vp, vp]),
        "orc_c_shape_info": (i32, [vp, i32, vp, vp, vp, vp]),
        "orc_c_set_particles": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "orc_c_add_wall": (i32, [vp, vp, vp, vp, d, i32]),
        "orc_c_add_rectangle_wall": (i32, [vp, vp, vp, d, d, d, i32]),
        "orc_c_set_contact_model": (i32, [vp, i32, i32, i32, vp]),
        "orc_c_prime": (i32, [vp]),
        "orc_c_step": (i32, [vp, i32, d]),
        "orc_c_get_state": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "orc_c_enable_contact_log": (None, [vp, i32]),
        "orc_c_get_contacts": (i32, [vp, i32, P(i32), vp, vp, vp, vp, vp, vp]),
        "orc_c_last_error": (i32, [vp, P(i32), P(C.c_int32), P(C.c_int32)]),
        "orc_c_gjk_calls": (C.c_int64, [vp]),
        "orc_c_get_contact_stats": (None, [vp, vp]),
        "orc_c_closest_points": (i32, [vp, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, P(i32), P(d)]),
        "orc_c_set_narrow_phase": (i32, [vp, i32]),
        "orc_c_set_support_history": (i32, [vp, i32]),
        "orc_c_support": (i32, [vp, i32, vp, i32, vp, P(i32)]),
        "orc_c_pair": (i32, [vp, i32, i32, vp, vp, vp, P(d), vp]),
        "orc_c_wall_pair": (i32, [vp, i32, i32, vp, vp, P(d), vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib._convex_sig = True
    return lib


class ConvexOracle:
    """Oracle of the convex (GJK) step built from a workload dict with
    "shapes" (pacific_amd.workloads.convex_box)."""

    def __init__(self, wl, log_capacity=0, support_history=False):
        """support_history: keep each polyhedron's Polyhedron::m_curr_vertex
        between support calls, as the reference does (default: every call
        starts from the first vertex with neighbours, as the GPU path)."""
        self.lib = _load_convex(load())
        cfg = _Config()
        for d in range(3):
            cfg.origin[d] = wl["origin"][d]
            cfg.size[d] = wl["size"][d]
            cfg.periodic[d] = int(wl["periodic"][d])
            cfg.gravity[d] = wl["gravity"][d]
        cfg.cell_size = wl["cell_size"]
        h = C.c_void_p()
        rc = self.lib.orc_c_create(C.byref(cfg), C.byref(h))
        if rc:
            raise OracleError(rc, "orc_c_create", -1, -1)
        self.h = h
        # CollisionDetection Type (Base/src/Grains.cpp:660-666): "GJK" or "GJK_SV"
        self.set_narrow_phase(wl.get("narrow_phase", "GJK"))
        for sh in wl["shapes"]:
            self.add_shape(sh)
        for (a, b, kind, p) in wl["models"]:
            self._chk(self.lib.orc_c_set_contact_model(h, a, b, kind, _p(_f(p, (5,)))), "model")
        for (basis, centre, half, crust, mat) in wl["walls"]:
            self._chk(self.lib.orc_c_add_wall(h, _p(_f(basis, (9,))), _p(_f(centre, (3,))),
                                              _p(_f(half, (3,))), crust, mat), "wall")
        for (basis, centre, lx, ly, crust, mat) in wl.get("rectangles", []):
            self._chk(self.lib.orc_c_add_rectangle_wall(h, _p(_f(basis, (9,))), _p(_f(centre, (3,))),
                                                        float(lx), float(ly), float(crust), int(mat)),
                      "rectangle")
        p = wl["particles"]
        n = len(p["shape"])
        self.n = n
        keep = [_f(p["x"], (n, 3)), _f(p["v"], (n, 3)), _f(p["w"], (n, 3)), _f(p["q"], (n, 4)),
                _f(p["basis"], (n, 9))]
        shape = np.ascontiguousarray(p["shape"], dtype=np.int32)
        dens, crust = _f(p["density"], (n,)), _f(p["crust"], (n,))
        mat = np.ascontiguousarray(p["material"], dtype=np.int32)
        self._chk(self.lib.orc_c_set_particles(h, n, *[_p(a) for a in keep], _p(shape), _p(dens),
                                               _p(crust), _p(mat)), "particles")
        if support_history:
            self._chk(self.lib.orc_c_set_support_history(h, 1), "support history")
        if log_capacity:
            self.lib.orc_c_enable_contact_log(h, int(log_capacity))

    def add_shape(self, sh):
        if sh[0] == "box":
            return self.lib.orc_c_add_box_shape(self.h, _p(_f(sh[1], (3,))))
        if sh[0] == "sphere":
            return self.lib.orc_c_add_sphere_shape(self.h, float(sh[1]))
        verts = _f(sh[1], (-1, 3))
        faces = [list(f) for f in sh[2]]
        sizes = np.array([len(f) for f in faces], np.int32)
        idx = np.array([i for f in faces for i in f], np.int32)
        return self.lib.orc_c_add_polyhedron_shape(self.h, len(verts), _p(verts), len(faces),
                                                   _p(sizes), _p(idx))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_c_destroy(self.h)
            self.h = None

    def _chk(self, rc, what):
        if rc:
            code, a, b = C.c_int(), C.c_int32(), C.c_int32()
            self.lib.orc_c_last_error(self.h, C.byref(code), C.byref(a), C.byref(b))
            raise OracleError(rc, what, a.value, b.value)

    def prime(self):
        self._chk(self.lib.orc_c_prime(self.h), "prime")

    def step(self, n, dt):
        self._chk(self.lib.orc_c_step(self.h, int(n), float(dt)), "step")

    def state(self):
        n = self.n
        out = {k: np.empty((n, 3)) for k in ("x", "v", "w", "F", "T")}
        out["q"] = np.empty((n, 4))
        out["basis"] = np.empty((n, 9))
        self.lib.orc_c_get_state(self.h, _p(out["x"]), _p(out["v"]), _p(out["w"]), _p(out["q"]),
                                 _p(out["basis"]), _p(out["F"]), _p(out["T"]))
        return out

    def contacts(self):
        n = C.c_int()
        self.lib.orc_c_get_contacts(self.h, 0, C.byref(n), None, None, None, None, None, None)
        m = n.value
        res = {"id0": np.empty(m, np.int32), "id1": np.empty(m, np.int32),
               "point": np.empty((m, 3)), "ovec": np.empty((m, 3)), "overlap": np.empty(m),
               "force": np.empty((m, 3))}
        self.lib.orc_c_get_contacts(self.h, m, C.byref(n), _p(res["id0"]), _p(res["id1"]),
                                    _p(res["point"]), _p(res["ovec"]), _p(res["overlap"]),
                                    _p(res["force"]))
        return res

    def shape_info(self, k):
        vol, circ = C.c_double(), C.c_double()
        inertia, bb = np.zeros(6), np.zeros(3)
        self.lib.orc_c_shape_info(self.h, k, C.byref(vol), _p(inertia), C.byref(circ), _p(bb))
        return vol.value, inertia, circ.value, bb

    def set_narrow_phase(self, name):
        if name not in ("GJK", "GJK_SV"):
            raise ValueError(f"unknown narrow phase {name!r}")
        self._chk(self.lib.orc_c_set_narrow_phase(self.h, int(name == "GJK_SV")), "narrow phase")

    def support(self, shape, direction, start=-1):
        """Support point of a shape along direction, climbing from vertex
        start (-1: the first vertex with neighbours); returns (point, vertex)."""
        pt, vx = np.zeros(3), C.c_int()
        self._chk(self.lib.orc_c_support(self.h, int(shape), _p(_f(direction, (3,))), int(start),
                                         _p(pt), C.byref(vx)), "support")
        return pt, vx.value

    def closest_points(self, sa, basis_a, origin_a, scale_a, sb, basis_b, origin_b, scale_b):
        pa, pb = np.zeros(3), np.zeros(3)
        it, d = C.c_int(), C.c_double()
        self.lib.orc_c_closest_points(self.h, sa, _p(_f(basis_a, (9,))), _p(_f(origin_a, (3,))),
                                      _p(_f(scale_a, (3,))), sb, _p(_f(basis_b, (9,))),
                                      _p(_f(origin_b, (3,))), _p(_f(scale_b, (3,))), _p(pa),
                                      _p(pb), C.byref(it), C.byref(d))
        return d.value, pa, pb, it.value

    def gjk_calls(self):
        return self.lib.orc_c_gjk_calls(self.h)

    def contact_stats(self):
        """AppCollision's contact features of the run, as Oracle.contact_stats."""
        out = np.zeros(6)
        self.lib.orc_c_get_contact_stats(self.h, _p(out))
        return {"overlap_max": out[0], "overlap_mean": out[1], "time_overlap_max": out[2],
                "gjk_mean": out[3], "overlaps": int(out[4]), "gjk_reports": int(out[5])}

    def pair(self, p0, p1, shift=(0.0, 0.0, 0.0)):
        """(rc, point, overlap vector, overlap, force on p0) of one pair with
        p0 as the first body, on the current state."""
        pt, ov, f = np.zeros(3), np.zeros(3), np.zeros(3)
        dist = C.c_double()
        rc = self.lib.orc_c_pair(self.h, int(p0), int(p1), _p(_f(shift, (3,))), _p(pt), _p(ov),
                                 C.byref(dist), _p(f))
        return rc, pt, ov, dist.value, f

    def wall_pair(self, p, wall):
        pt, ov, f = np.zeros(3), np.zeros(3), np.zeros(3)
        dist = C.c_double()
        rc = self.lib.orc_c_wall_pair(self.h, int(p), int(wall), _p(pt), _p(ov), C.byref(dist), _p(f))
        return rc, pt, ov, dist.value, f
