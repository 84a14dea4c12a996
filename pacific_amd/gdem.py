"""ctypes binding of libgdem.so (include/gdem.h).

Python mirror of the batched plugin seams of Grains3D: one ``GdemSystem`` is
one ``Grains`` application restricted to its step (LinkedCell binning,
ContactForceModel forces, SecondOrderLeapFrog integration).  Errors the device
raises come back as the reference's exception types (ContactError,
MotionError, SimulationError; src/Grains3D/Grains/Base/include/Error.hh:19,75,129).

There is no CPU fallback: if the HIP library is missing the import of this
module fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgdem.so")

OK, ERR_CONTACT, ERR_MOTION, ERR_SIMULATION, ERR_HIP, ERR_ARG, ERR_CAPACITY = range(7)
MODEL_HOOKE, MODEL_HERTZ = 1, 2
NARROW_PHASES = {"GJK": 0, "GJK_SV": 1}  # GDEM_GJK, GDEM_GJK_SV
MODEL_HOOKE_MEMORY, MODEL_HERTZ_MEMORY = 3, 4
# parameters per model kind (include/gdem.h)
MODEL_NPARAMS = {MODEL_HOOKE: 5, MODEL_HERTZ: 5, MODEL_HOOKE_MEMORY: 8, MODEL_HERTZ_MEMORY: 7}
# imposed wall motions (ObstacleImposedVelocity types) and their parameter counts
MOTION_CONSTANT_TRANSLATION, MOTION_SIN_TRANSLATION, MOTION_CRENEL_TRANSLATION = 1, 2, 3
MOTION_MULTISIN_TRANSLATION, MOTION_CONSTANT_ROTATION = 4, 5
MOTION_NPARAMS = {1: 3, 2: 6, 3: 5, 4: 12, 5: 7}


class GdemError(RuntimeError):
    def __init__(self, code, msg, id0=-1, id1=-1):
        super().__init__(f"{msg} (code {code}, ids {id0}, {id1})")
        self.code, self.id0, self.id1 = code, id0, id1


class ContactError(GdemError):
    """Overlap reached the sum of the crust thicknesses (RigidBodyWithCrust.cpp:624-630)."""


class MotionError(GdemError):
    """Displacement over one step exceeded the crust (Particle.cpp:447-457)."""


class SimulationError(GdemError):
    """Particle left the linked-cell grid (LinkedCell.cpp:1506-1531)."""


class CapacityError(GdemError):
    """A contact buffer was too small."""


_ERRORS = {ERR_CONTACT: ContactError, ERR_MOTION: MotionError,
           ERR_SIMULATION: SimulationError, ERR_CAPACITY: CapacityError}


class _Config(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("size", C.c_double * 3),
                ("periodic", C.c_int32 * 3), ("cell_size", C.c_double),
                ("gravity", C.c_double * 3), ("device", C.c_int32),
                ("reserved", C.c_int32)]


class _BatchSystem(C.Structure):
    _fields_ = [("cfg", _Config), ("n", C.c_int64),
                ("x", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p), ("q", C.c_void_p),
                ("radius", C.c_void_p), ("density", C.c_void_p), ("crust", C.c_void_p),
                ("kind", C.c_int32), ("nwalls", C.c_int32),
                ("pp", C.c_double * 5), ("pw", C.c_double * 5),
                ("wall_basis", C.c_void_p), ("wall_centre", C.c_void_p),
                ("wall_half", C.c_void_p), ("wall_crust", C.c_double),
                ("substeps", C.c_int32), ("reserved", C.c_int32)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libgdem.so once; raises OSError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.POINTER
    d, i32, i64 = C.c_double, C.c_int32, C.c_int64
    vp = C.c_void_p
    sig = {
        "gdem_create": (C.c_int, [P(_Config), P(vp)]),
        "gdem_destroy": (C.c_int, [vp]),
        "gdem_set_particles": (C.c_int, [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp]),
        "gdem_add_wall": (C.c_int, [vp, vp, vp, vp, d, i32]),
        "gdem_set_contact_model": (C.c_int, [vp, i32, i32, i32, vp]),
        "gdem_set_time_step": (C.c_int, [vp, d]),
        "gdem_link_update": (C.c_int, [vp]),
        "gdem_compute_forces": (C.c_int, [vp]),
        "gdem_move": (C.c_int, [vp, d, d]),
        "gdem_advance_velocity": (C.c_int, [vp, d]),
        "gdem_prime": (C.c_int, [vp]),
        "gdem_step": (C.c_int, [vp, i32, d]),
        "gdem_set_graph": (C.c_int, [vp, i32]),
        "gdem_set_skin": (C.c_int, [vp, d]),
        "gdem_get_skin": (C.c_int, [vp, P(d), P(i64)]),
        "gdem_synchronize": (C.c_int, [vp]),
        "gdem_get_state": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
        "gdem_set_state": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
        "gdem_get_memory": (C.c_int, [vp, i64, P(i64), vp, vp, vp, vp, vp]),
        "gdem_set_memory": (C.c_int, [vp, i64, vp, vp, vp, vp, vp]),
        "gdem_info": (C.c_int, [vp, P(i64), P(i32), P(i32), P(i32)]),
        "gdem_add_wall_motion": (C.c_int, [vp, i32, i32, d, d, vp]),
        "gdem_set_wall_loads": (C.c_int, [vp, i32]),
        "gdem_set_motion_restriction": (C.c_int, [vp, i32, i32, i32]),
        "gdem_add_wall_force": (C.c_int, [vp, i32, d, d, vp]),
        "gdem_add_composite_force": (C.c_int, [vp, i32, d, d, vp]),
        "gdem_get_wall_loads": (C.c_int, [vp, vp, vp]),
        "gdem_add_composite": (C.c_int, [vp, i32, P(i32)]),
        "gdem_attach_wall": (C.c_int, [vp, i32, i32]),
        "gdem_add_composite_motion": (C.c_int, [vp, i32, i32, d, d, vp]),
        "gdem_get_composite": (C.c_int, [vp, i32, vp, vp, vp]),
        "gdem_add_cylindrical_shell": (C.c_int, [vp, vp, vp, d, d, d, i32, d, i32, i32, P(i32)]),
        "gdem_cylindrical_shell_boxes": (C.c_int, [vp, vp, d, d, d, i32, vp, vp, vp]),
        "gdem_add_rectangle_wall": (C.c_int, [vp, vp, vp, d, d, d, i32]),
        "gdem_add_cylindrical_shell_rectangles": (C.c_int, [vp, vp, vp, d, d, d, i32, d, i32, i32,
                                                            P(i32)]),
        "gdem_cylindrical_shell_rectangles": (C.c_int, [vp, vp, d, d, d, i32, vp, vp, vp]),
        "gdem_set_time": (C.c_int, [vp, d]),
        "gdem_get_time": (C.c_int, [vp, P(d)]),
        "gdem_move_walls": (C.c_int, [vp, d, d]),
        "gdem_get_wall": (C.c_int, [vp, i32, vp, vp, vp, vp]),
        "gdem_write_particles_2014": (C.c_int, [vp, C.c_char_p, i32, i32, i32, vp]),
        "gdem_read_particles_2014": (C.c_int, [vp, C.c_char_p, i32, i32, i32, vp]),
        "gdem_set_contact_recording": (C.c_int, [vp, i32]),
        "gdem_get_contacts": (C.c_int, [vp, i64, P(i64), vp, vp, vp, vp, vp, vp]),
        "gdem_last_error": (C.c_int, [vp, P(i32), P(i32), P(i32)]),
        "gdem_set_kernel_timing": (C.c_int, [vp, i32]),
        "gdem_get_counters": (C.c_int, [vp, P(i64), P(i64), P(d), P(i64)]),
        "gdem_reset_counters": (C.c_int, [vp]),
        "gdem_get_kernel_times": (C.c_int, [vp, P(d), P(i64)]),
        "gdem_set_contact_stats": (C.c_int, [vp, i32]),
        "gdem_get_contact_stats": (C.c_int, [vp, P(d)]),
        "gdem_set_narrow_phase": (C.c_int, [vp, i32]),
        "gdem_set_skin_adapt": (C.c_int, [vp, i32]),
        "gdem_get_list_occupancy": (C.c_int, [vp, P(i32), P(i32), P(i32)]),
        "gdem_get_narrow_phase": (C.c_int, [vp, P(i32)]),
        "gdem_get_grid": (C.c_int, [vp, P(i32), P(d)]),
        "gdem_get_particles": (C.c_int, [vp, vp, vp, vp, vp]),
        "gdem_get_contact_model": (C.c_int, [vp, i32, i32, P(i32), vp]),
        "gdem_get_wall_shape": (C.c_int, [vp, i32, vp, vp, vp, P(d), P(i32), P(i32)]),
        "gdem_set_material_name": (C.c_int, [vp, i32, C.c_char_p]),
        "gdem_get_material_name": (C.c_int, [vp, i32, C.c_char_p, i32]),
        "gdem_set_wall_name": (C.c_int, [vp, i32, C.c_char_p]),
        "gdem_get_wall_name": (C.c_int, [vp, i32, C.c_char_p, i32]),
        "gdem_write_reload": (C.c_int, [vp, C.c_char_p, C.c_char_p, d, i32, i32, i32]),
        "gdem_read_reload": (C.c_int, [P(_Config), C.c_char_p, i32, i32, P(vp), P(d)]),
        "gdem_version": (C.c_char_p, []),
        "gdem_sphere_forces_noalloc": (C.c_int, [P(_Config), i32, vp, i64, vp, vp, vp, d, vp, vp,
                                                 vp, P(i32)]),
        "gdem_add_box_shape": (C.c_int, [vp, vp, P(i32)]),
        "gdem_add_polyhedron_shape": (C.c_int, [vp, i32, vp, i32, vp, vp, P(i32)]),
        "gdem_set_convex_particles": (C.c_int, [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "gdem_get_basis": (C.c_int, [vp, vp]),
        "gdem_batch_create": (C.c_int, [i32, P(vp)]),
        "gdem_batch_destroy": (C.c_int, [vp]),
        "gdem_batch_add_system": (C.c_int, [vp, P(_BatchSystem), P(i32)]),
        "gdem_batch_commit": (C.c_int, [vp]),
        "gdem_batch_set_launch": (C.c_int, [vp, i32, i32]),
        "gdem_batch_prime": (C.c_int, [vp]),
        "gdem_batch_step": (C.c_int, [vp, i32, d]),
        "gdem_batch_synchronize": (C.c_int, [vp]),
        "gdem_batch_get_state": (C.c_int, [vp, i32, vp, vp, vp, vp, vp, vp]),
        "gdem_batch_last_error": (C.c_int, [vp, P(i32), P(i32), P(i32), P(i32)]),
        "gdem_batch_set_kernel_timing": (C.c_int, [vp, i32]),
        "gdem_batch_get_counters": (C.c_int, [vp, P(i64), P(i64), P(d), P(i64)]),
        "gdem_batch_info": (C.c_int, [vp, P(i32), P(i64)]),
        "gdem_diff_create": (C.c_int, [i32, P(vp)]),
        "gdem_diff_destroy": (C.c_int, [vp]),
        "gdem_diff_add_system": (C.c_int, [vp, P(_BatchSystem), P(i32)]),
        "gdem_diff_commit": (C.c_int, [vp, i32, d, i32, i32, vp]),
        "gdem_diff_set_launch": (C.c_int, [vp, i32, i32]),
        "gdem_diff_set_parameters": (C.c_int, [vp, vp, vp]),
        "gdem_diff_set_targets": (C.c_int, [vp, vp]),
        "gdem_diff_record": (C.c_int, [vp, vp]),
        "gdem_diff_loss": (C.c_int, [vp, vp]),
        "gdem_diff_gradient": (C.c_int, [vp, vp, vp, vp]),
        "gdem_diff_directional": (C.c_int, [vp, vp, vp, vp, vp]),
        "gdem_diff_last_error": (C.c_int, [vp, P(i32), P(i32), P(i32), P(i32)]),
        "gdem_diff_info": (C.c_int, [vp, P(i32), P(i64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


EXPORTED = ["gdem_create", "gdem_destroy", "gdem_set_particles", "gdem_add_wall",
            "gdem_set_contact_model", "gdem_set_time_step", "gdem_link_update", "gdem_compute_forces",
            "gdem_move", "gdem_advance_velocity", "gdem_prime", "gdem_step", "gdem_set_graph",
            "gdem_set_skin", "gdem_get_skin",
            "gdem_synchronize", "gdem_get_state", "gdem_set_state", "gdem_get_memory",
            "gdem_set_memory", "gdem_info", "gdem_add_wall_motion", "gdem_set_wall_loads",
            "gdem_get_wall_loads", "gdem_add_wall_force", "gdem_set_motion_restriction",
            "gdem_add_composite",
            "gdem_attach_wall", "gdem_add_composite_motion", "gdem_get_composite",
            "gdem_add_cylindrical_shell", "gdem_cylindrical_shell_boxes", "gdem_set_time",
            "gdem_add_rectangle_wall", "gdem_add_cylindrical_shell_rectangles",
            "gdem_add_composite_force",
            "gdem_cylindrical_shell_rectangles",
            "gdem_get_time", "gdem_move_walls", "gdem_get_wall", "gdem_write_particles_2014",
            "gdem_read_particles_2014", "gdem_set_contact_recording",
            "gdem_get_contacts", "gdem_last_error", "gdem_set_kernel_timing",
            "gdem_get_counters", "gdem_reset_counters", "gdem_get_kernel_times",
            "gdem_set_contact_stats", "gdem_get_contact_stats", "gdem_get_grid",
            "gdem_set_narrow_phase", "gdem_get_narrow_phase", "gdem_set_skin_adapt",
            "gdem_get_list_occupancy",
            "gdem_get_particles", "gdem_get_contact_model", "gdem_get_wall_shape",
            "gdem_set_material_name", "gdem_get_material_name", "gdem_set_wall_name",
            "gdem_get_wall_name", "gdem_write_reload", "gdem_read_reload",
            "gdem_version", "gdem_sphere_forces_noalloc",
            "gdem_add_box_shape", "gdem_add_polyhedron_shape", "gdem_set_convex_particles",
            "gdem_get_basis", "gdem_batch_create", "gdem_batch_destroy", "gdem_batch_add_system",
            "gdem_batch_commit", "gdem_batch_set_launch", "gdem_batch_prime", "gdem_batch_step",
            "gdem_batch_synchronize", "gdem_batch_get_state", "gdem_batch_last_error",
            "gdem_batch_set_kernel_timing", "gdem_batch_get_counters", "gdem_batch_info",
            "gdem_diff_create", "gdem_diff_destroy", "gdem_diff_add_system", "gdem_diff_commit",
            "gdem_diff_set_launch", "gdem_diff_set_parameters", "gdem_diff_set_targets",
            "gdem_diff_record", "gdem_diff_loss", "gdem_diff_gradient", "gdem_diff_directional",
            "gdem_diff_last_error",
            "gdem_diff_info"]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape))


class GdemSystem:
    """One DEM system on one GPU (see include/gdem.h for the phase mapping)."""

    def __init__(self, origin, size, periodic=(0, 0, 0), cell_size=0.0,
                 gravity=(0.0, 0.0, 0.0), device=0):
        self.lib = load_library()
        cfg = _Config()
        for d in range(3):
            cfg.origin[d] = float(origin[d])
            cfg.size[d] = float(size[d])
            cfg.periodic[d] = int(bool(periodic[d]))
            cfg.gravity[d] = float(gravity[d])
        cfg.cell_size = float(cell_size)
        cfg.device = int(device)
        h = C.c_void_p()
        self._check(self.lib.gdem_create(C.byref(cfg), C.byref(h)), "gdem_create")
        self.h = h
        self.n = 0

    # -- lifecycle
    def close(self):
        if getattr(self, "h", None):
            self.lib.gdem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc == OK:
            return
        id0 = id1 = -1
        if getattr(self, "h", None):
            code, a, b = C.c_int32(), C.c_int32(), C.c_int32()
            self.lib.gdem_last_error(self.h, C.byref(code), C.byref(a), C.byref(b))
            id0, id1 = a.value, b.value
        raise _ERRORS.get(rc, GdemError)(rc, what, id0, id1)

    # -- setup
    def set_particles(self, x, v, w, q, radius, density, crust, material=None):
        n = len(radius)
        self._keep = [_f64(x, (n, 3)), _f64(v, (n, 3)), _f64(w, (n, 3)), _f64(q, (n, 4)),
                      _f64(radius, (n,)), _f64(density, (n,)), _f64(crust, (n,))]
        mat = None if material is None else np.ascontiguousarray(material, dtype=np.int32)
        self._check(self.lib.gdem_set_particles(self.h, n, *[_ptr(a) for a in self._keep],
                                                _ptr(mat)), "gdem_set_particles")
        self.n = n

    def add_box_shape(self, half_extent):
        """Box with the given half extents; returns its shape id."""
        sid = C.c_int32()
        self._check(self.lib.gdem_add_box_shape(self.h, _ptr(_f64(half_extent, (3,))), C.byref(sid)),
                    "gdem_add_box_shape")
        return sid.value

    def add_polyhedron_shape(self, vertices, faces):
        """Polyhedron from vertices (about the centre of mass) and faces (lists
        of vertex indices, each loop oriented outwards); returns its shape id."""
        v = _f64(vertices, (-1, 3))
        size = np.array([len(f) for f in faces], np.int32)
        idx = np.array([i for f in faces for i in f], np.int32)
        sid = C.c_int32()
        self._check(self.lib.gdem_add_polyhedron_shape(self.h, len(v), _ptr(v), len(size), _ptr(size),
                                                       _ptr(idx), C.byref(sid)),
                    "gdem_add_polyhedron_shape")
        return sid.value

    def set_convex_particles(self, x, v, w, q, basis, shape, density, crust, material=None):
        n = len(shape)
        self._keep = [_f64(x, (n, 3)), _f64(v, (n, 3)), _f64(w, (n, 3)), _f64(q, (n, 4)),
                      _f64(basis, (n, 9))]
        sh = np.ascontiguousarray(shape, dtype=np.int32)
        rest = [_f64(density, (n,)), _f64(crust, (n,))]
        self._keep += [sh] + rest
        mat = None if material is None else np.ascontiguousarray(material, dtype=np.int32)
        self._check(self.lib.gdem_set_convex_particles(
            self.h, n, *[_ptr(a) for a in self._keep[:5]], _ptr(sh), *[_ptr(a) for a in rest],
            _ptr(mat)), "gdem_set_convex_particles")
        self.n = n

    def get_basis(self):
        """Row-major rotation of every convex particle, (n, 9)."""
        b = np.empty((self.n, 9))
        self._check(self.lib.gdem_get_basis(self.h, _ptr(b)), "gdem_get_basis")
        return b

    def add_wall(self, basis, centre, half_extent, crust, material):
        self._check(self.lib.gdem_add_wall(self.h, _ptr(_f64(basis, (9,))), _ptr(_f64(centre, (3,))),
                                           _ptr(_f64(half_extent, (3,))), float(crust), int(material)),
                    "gdem_add_wall")

    def add_wall_motion(self, wall, kind, t_start, t_end, params):
        """Imposed velocity of a wall over [t_start, t_end] (MOTION_* kinds,
        parameters as MOTION_NPARAMS counts them)."""
        p = _f64(params, (MOTION_NPARAMS[int(kind)],))
        self._check(self.lib.gdem_add_wall_motion(self.h, int(wall), int(kind), float(t_start),
                                                  float(t_end), _ptr(p)), "gdem_add_wall_motion")

    def set_wall_loads(self, on=True):
        """Track the load on every wall in each force evaluation."""
        self._check(self.lib.gdem_set_wall_loads(self.h, int(bool(on))), "gdem_set_wall_loads")

    def get_wall_loads(self):
        """(force (nwalls, 3), torque about each wall's centre (nwalls, 3)) of
        the last force evaluation."""
        nw = C.c_int32()
        self._check(self.lib.gdem_info(self.h, None, C.byref(nw), None, None), "gdem_info")
        f, t = np.zeros((max(nw.value, 1), 3)), np.zeros((max(nw.value, 1), 3))
        self._check(self.lib.gdem_get_wall_loads(self.h, _ptr(f), _ptr(t)), "gdem_get_wall_loads")
        return f[:nw.value], t[:nw.value]

    def add_wall_force(self, wall, t_start, t_end, params):
        """Imposed force on a wall: params = (ax, ay, az, mass (<= 0: automatic), vmax)."""
        self._check(self.lib.gdem_add_wall_force(self.h, int(wall), float(t_start), float(t_end),
                                                 _ptr(_f64(params, (5,)))), "gdem_add_wall_force")

    def add_composite_force(self, composite, t_start, t_end, params):
        """Imposed force on a composite's load: params = (ax, ay, az, mass > 0, vmax)."""
        self._check(self.lib.gdem_add_composite_force(self.h, int(composite), float(t_start),
                                                      float(t_end), _ptr(_f64(params, (5,)))),
                    "gdem_add_composite_force")

    def set_motion_restriction(self, mask, wall=-1, composite=-1):
        """No geometric translation along the directions of mask (bit 0 x, 1 y,
        2 z) for a wall or a composite; its velocity is kept."""
        self._check(self.lib.gdem_set_motion_restriction(self.h, int(wall), int(composite), int(mask)),
                    "gdem_set_motion_restriction")

    def add_composite(self, parent=-1):
        """A composite obstacle under parent (-1: the root); returns its index."""
        k = C.c_int32()
        self._check(self.lib.gdem_add_composite(self.h, int(parent), C.byref(k)), "gdem_add_composite")
        return k.value

    def attach_wall(self, wall, composite):
        self._check(self.lib.gdem_attach_wall(self.h, int(wall), int(composite)), "gdem_attach_wall")

    def add_composite_motion(self, composite, kind, t_start, t_end, params):
        p = _f64(params, (MOTION_NPARAMS[int(kind)],))
        self._check(self.lib.gdem_add_composite_motion(self.h, int(composite), int(kind),
                                                       float(t_start), float(t_end), _ptr(p)),
                    "gdem_add_composite_motion")

    def get_composite(self, k):
        """(centre, velocity, angular velocity) of composite k."""
        c, v, w = np.empty(3), np.empty(3), np.empty(3)
        self._check(self.lib.gdem_get_composite(self.h, int(k), _ptr(c), _ptr(v), _ptr(w)),
                    "gdem_get_composite")
        return c, v, w

    def add_cylindrical_shell(self, basis, centre, height, radius, width, n, crust, material,
                              parent=-1):
        """Box-element CylindricalShell as a composite of n walls; returns its index."""
        k = C.c_int32()
        self._check(self.lib.gdem_add_cylindrical_shell(
            self.h, _ptr(_f64(basis, (9,))), _ptr(_f64(centre, (3,))), float(height), float(radius),
            float(width), int(n), float(crust), int(material), int(parent), C.byref(k)),
            "gdem_add_cylindrical_shell")
        return k.value

    def add_rectangle_wall(self, basis, centre, lx, ly, crust, material):
        """Rectangle obstacle of sides (lx, ly) in its local xy plane (normal
        local z); sphere and convex particles.  Its wall index is the next one."""
        self._check(self.lib.gdem_add_rectangle_wall(
            self.h, _ptr(_f64(basis, (9,))), _ptr(_f64(centre, (3,))), float(lx), float(ly),
            float(crust), int(material)), "gdem_add_rectangle_wall")

    def add_cylindrical_shell_rectangles(self, basis, centre, height, radius, width, n, crust,
                                         material, parent=-1):
        """CylindricalShell with its default Rectangle elements as a
        composite of n walls; returns its index."""
        k = C.c_int32()
        self._check(self.lib.gdem_add_cylindrical_shell_rectangles(
            self.h, _ptr(_f64(basis, (9,))), _ptr(_f64(centre, (3,))), float(height), float(radius),
            float(width), int(n), float(crust), int(material), int(parent), C.byref(k)),
            "gdem_add_cylindrical_shell_rectangles")
        return k.value

    def set_time(self, t):
        self._check(self.lib.gdem_set_time(self.h, float(t)), "gdem_set_time")

    def get_time(self):
        t = C.c_double()
        self._check(self.lib.gdem_get_time(self.h, C.byref(t)), "gdem_get_time")
        return t.value

    def move_walls(self, time, dt):
        """Moves the walls over [time - dt, time] (obstacle part of a step)."""
        self._check(self.lib.gdem_move_walls(self.h, float(time), float(dt)), "gdem_move_walls")

    def get_wall(self, k):
        """(centre, row-major basis (9,), velocity, angular velocity) of wall k."""
        c, b, v, w = np.empty(3), np.empty(9), np.empty(3), np.empty(3)
        self._check(self.lib.gdem_get_wall(self.h, int(k), _ptr(c), _ptr(b), _ptr(v), _ptr(w)),
                    "gdem_get_wall")
        return c, b, v, w

    def set_contact_model(self, mat_a, mat_b, kind, params):
        p = _f64(params, (MODEL_NPARAMS.get(int(kind), 5),))
        self._check(self.lib.gdem_set_contact_model(self.h, int(mat_a), int(mat_b), int(kind), _ptr(p)),
                    "gdem_set_contact_model")

    def set_time_step(self, dt):
        """dt of the force evaluations for the memory models (gdem_step and
        move set it too)."""
        self._check(self.lib.gdem_set_time_step(self.h, float(dt)), "gdem_set_time_step")

    # -- phases
    def link_update(self):
        self._check(self.lib.gdem_link_update(self.h), "gdem_link_update")

    def compute_forces(self):
        self._check(self.lib.gdem_compute_forces(self.h), "gdem_compute_forces")

    def move(self, dt_vel, dt_disp):
        self._check(self.lib.gdem_move(self.h, float(dt_vel), float(dt_disp)), "gdem_move")

    def advance_velocity(self, dt_vel):
        self._check(self.lib.gdem_advance_velocity(self.h, float(dt_vel)), "gdem_advance_velocity")

    def prime(self):
        self._check(self.lib.gdem_prime(self.h), "gdem_prime")

    def step(self, nsteps, dt):
        self._check(self.lib.gdem_step(self.h, int(nsteps), float(dt)), "gdem_step")

    def synchronize(self):
        self._check(self.lib.gdem_synchronize(self.h), "gdem_synchronize")

    def set_skin(self, skin):
        """Skin of the candidate lists (m); negative = automatic, 0 = off."""
        self._check(self.lib.gdem_set_skin(self.h, float(skin)), "gdem_set_skin")

    def get_skin(self):
        """(effective skin in m, list rebuilds since the particles or skin were set)."""
        s, k = C.c_double(), C.c_int64()
        self._check(self.lib.gdem_get_skin(self.h, C.byref(s), C.byref(k)), "gdem_get_skin")
        return s.value, k.value

    def set_graph(self, on):
        self._check(self.lib.gdem_set_graph(self.h, int(bool(on))), "gdem_set_graph")

    # -- state transfer and reload files
    def set_state(self, x=None, v=None, w=None, q=None, F=None, T=None):
        """Restore particle state in id order (None: kept); F and T together
        become the leap-frog's a0 (no force evaluation before the next step)."""
        n = self.n
        arrs = [None if a is None else _f64(a, (n, k)) for a, k in
                ((x, 3), (v, 3), (w, 3), (q, 4), (F, 3), (T, 3))]
        self._check(self.lib.gdem_set_state(self.h, *[_ptr(a) for a in arrs]), "gdem_set_state")

    def get_memory(self):
        """Contact memory: dict of arrays owner, key, kd, nrm, spr (by owner, key)."""
        cnt = C.c_int64()
        self._check(self.lib.gdem_get_memory(self.h, 0, C.byref(cnt), None, None, None, None, None),
                    "gdem_get_memory")
        m = cnt.value
        out = {"owner": np.empty(m, np.int32), "key": np.empty(m, np.int32),
               "kd": np.empty((m, 3)), "nrm": np.empty((m, 3)), "spr": np.empty((m, 3))}
        self._check(self.lib.gdem_get_memory(self.h, m, C.byref(cnt), _ptr(out["owner"]),
                                             _ptr(out["key"]), _ptr(out["kd"]), _ptr(out["nrm"]),
                                             _ptr(out["spr"])), "gdem_get_memory")
        return out

    def set_memory(self, mem):
        m = len(mem["owner"])
        own = np.ascontiguousarray(mem["owner"], dtype=np.int32)
        key = np.ascontiguousarray(mem["key"], dtype=np.int32)
        arrs = [_f64(mem[k], (m, 3)) for k in ("kd", "nrm", "spr")]
        self._check(self.lib.gdem_set_memory(self.h, m, _ptr(own), _ptr(key), *[_ptr(a) for a in arrs]),
                    "gdem_set_memory")

    def write_particles_2014(self, path, binary=True, id_offset=0, wall_id_offset=0, geom_type=None):
        gt = None if geom_type is None else np.ascontiguousarray(geom_type, dtype=np.int32)
        self._check(self.lib.gdem_write_particles_2014(self.h, str(path).encode(), int(bool(binary)),
                                                       int(id_offset), int(wall_id_offset), _ptr(gt)),
                    "gdem_write_particles_2014")

    def set_material_name(self, k, name):
        self._check(self.lib.gdem_set_material_name(self.h, int(k), str(name).encode()),
                    "gdem_set_material_name")

    def material_name(self, k):
        buf = C.create_string_buffer(256)
        self._check(self.lib.gdem_get_material_name(self.h, int(k), buf, 256), "gdem_get_material_name")
        return buf.value.decode()

    def set_wall_name(self, k, name):
        self._check(self.lib.gdem_set_wall_name(self.h, int(k), str(name).encode()),
                    "gdem_set_wall_name")

    def wall_name(self, k):
        buf = C.create_string_buffer(256)
        self._check(self.lib.gdem_get_wall_name(self.h, int(k), buf, 256), "gdem_get_wall_name")
        return buf.value.decode()

    def particles(self):
        """The particles' constants (radius, density, crust, material) in id order."""
        r, rho, cr = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        mat = np.empty(self.n, np.int32)
        self._check(self.lib.gdem_get_particles(self.h, _ptr(r), _ptr(rho), _ptr(cr), _ptr(mat)),
                    "gdem_get_particles")
        return {"radius": r, "density": rho, "crust": cr, "material": mat}

    def contact_model(self, a, b):
        kind = C.c_int32()
        p = np.zeros(8)
        self._check(self.lib.gdem_get_contact_model(self.h, int(a), int(b), C.byref(kind), _ptr(p)),
                    "gdem_get_contact_model")
        return kind.value, p

    def wall_shape(self, k):
        basis, centre, half = np.empty(9), np.empty(3), np.empty(3)
        crust, mat, kind = C.c_double(), C.c_int32(), C.c_int32()
        self._check(self.lib.gdem_get_wall_shape(self.h, int(k), _ptr(basis), _ptr(centre), _ptr(half),
                                                 C.byref(crust), C.byref(mat), C.byref(kind)),
                    "gdem_get_wall_shape")
        return {"basis": basis, "centre": centre, "half": half, "crust": crust.value,
                "material": mat.value, "kind": kind.value}

    def write_reload(self, root, suffix, time, binary=True, id_offset=1, wall_id_offset=0):
        """The reference's reload file set for this system (gdem_write_reload)."""
        self._check(self.lib.gdem_write_reload(self.h, str(root).encode(), str(suffix).encode(),
                                               float(time), int(bool(binary)), int(id_offset),
                                               int(wall_id_offset)), "gdem_write_reload")

    def read_particles_2014(self, path, binary=True, id_offset=0, wall_id_offset=0):
        """Restore state and memory from a particle reload file; returns the
        records' geometric types in id order."""
        gt = np.zeros(self.n, np.int32)
        self._check(self.lib.gdem_read_particles_2014(self.h, str(path).encode(), int(bool(binary)),
                                                      int(id_offset), int(wall_id_offset), _ptr(gt)),
                    "gdem_read_particles_2014")
        return gt

    # -- results
    def get_state(self):
        n = self.n
        out = {k: np.empty((n, 3)) for k in ("x", "v", "w", "F", "T")}
        out["q"] = np.empty((n, 4))
        self._check(self.lib.gdem_get_state(self.h, _ptr(out["x"]), _ptr(out["v"]), _ptr(out["w"]),
                                            _ptr(out["q"]), _ptr(out["F"]), _ptr(out["T"])),
                    "gdem_get_state")
        return out

    def set_contact_recording(self, cap):
        self._check(self.lib.gdem_set_contact_recording(self.h, int(cap)), "gdem_set_contact_recording")

    def get_contacts(self):
        nout = C.c_int64()
        rc = self.lib.gdem_get_contacts(self.h, 0, C.byref(nout), None, None, None, None, None, None)
        if rc not in (OK, ERR_CAPACITY):
            self._check(rc, "gdem_get_contacts")
        m = nout.value
        res = {"id0": np.empty(m, np.int32), "id1": np.empty(m, np.int32),
               "point": np.empty((m, 3)), "ovec": np.empty((m, 3)),
               "overlap": np.empty(m), "force": np.empty((m, 3))}
        self._check(self.lib.gdem_get_contacts(self.h, m, C.byref(nout), _ptr(res["id0"]),
                                               _ptr(res["id1"]), _ptr(res["point"]), _ptr(res["ovec"]),
                                               _ptr(res["overlap"]), _ptr(res["force"])),
                    "gdem_get_contacts")
        return res

    def set_kernel_timing(self, on):
        self._check(self.lib.gdem_set_kernel_timing(self.h, int(bool(on))), "gdem_set_kernel_timing")

    def counters(self):
        fe, ct, tl = C.c_int64(), C.c_int64(), C.c_int64()
        ms = C.c_double()
        self._check(self.lib.gdem_get_counters(self.h, C.byref(fe), C.byref(ct), C.byref(ms), C.byref(tl)),
                    "gdem_get_counters")
        return {"force_evals": fe.value, "contacts": ct.value, "force_kernel_ms": ms.value,
                "timed_launches": tl.value}

    def set_contact_stats(self, on=True):
        """Switch on (and reset) the contact statistics of the run."""
        self._check(self.lib.gdem_set_contact_stats(self.h, int(bool(on))),
                    "gdem_set_contact_stats")

    def set_skin_adapt(self, on=True):
        """Halve the skin when a list build nears the list capacity (default on)."""
        self._check(self.lib.gdem_set_skin_adapt(self.h, int(bool(on))), "gdem_set_skin_adapt")

    def list_occupancy(self):
        """Longest list of the last build, the list capacity and the skin
        halvings so far."""
        peak, cap, halv = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.gdem_get_list_occupancy(self.h, C.byref(peak), C.byref(cap),
                                                     C.byref(halv)), "gdem_get_list_occupancy")
        return {"peak": peak.value, "capacity": cap.value, "skin_halvings": halv.value}

    def set_narrow_phase(self, name):
        """Narrow phase of convex contacts: "GJK" (Johnson, the default) or
        "GJK_SV" (signed volumes), as CollisionDetection Type in the
        reference's input file."""
        if name not in NARROW_PHASES:
            raise ValueError(f"unknown narrow phase {name!r}")
        self._check(self.lib.gdem_set_narrow_phase(self.h, NARROW_PHASES[name]),
                    "gdem_set_narrow_phase")

    def narrow_phase(self):
        alg = C.c_int32()
        self._check(self.lib.gdem_get_narrow_phase(self.h, C.byref(alg)), "gdem_get_narrow_phase")
        return {v: k for k, v in NARROW_PHASES.items()}[alg.value]

    def contact_stats(self):
        """The run's contact features, as the oracle's contact_stats()."""
        out = (C.c_double * 6)()
        self._check(self.lib.gdem_get_contact_stats(self.h, out), "gdem_get_contact_stats")
        return {"overlap_max": out[0], "overlap_mean": out[1], "time_overlap_max": out[2],
                "gjk_mean": out[3], "overlaps": int(out[4]), "gjk_reports": int(out[5])}

    def kernel_times(self):
        """Event-timed totals (ms) and launch counts: [force / integrator kernel,
        pair kernel of the packed layout]."""
        ms = (C.c_double * 2)()
        n = (C.c_int64 * 2)()
        self._check(self.lib.gdem_get_kernel_times(self.h, ms, n), "gdem_get_kernel_times")
        return {"ms": list(ms), "launches": list(n)}

    def reset_counters(self):
        self._check(self.lib.gdem_reset_counters(self.h), "gdem_reset_counters")

    def grid(self):
        nc = (C.c_int32 * 3)()
        e = (C.c_double * 3)()
        self._check(self.lib.gdem_get_grid(self.h, nc, e), "gdem_get_grid")
        return list(nc), list(e)


def read_reload(main_path, origin, size, periodic=(0, 0, 0), cell_size=0.0,
                gravity=(0.0, 0.0, 0.0), id_offset=1, wall_id_offset=0, device=0):
    """A GdemSystem restarted from a reload file set in the reference's format
    (gdem_read_reload; the domain comes from the caller).  Returns (system,
    time)."""
    lib = load_library()
    cfg = _Config()
    for d in range(3):
        cfg.origin[d] = float(origin[d])
        cfg.size[d] = float(size[d])
        cfg.periodic[d] = int(bool(periodic[d]))
        cfg.gravity[d] = float(gravity[d])
    cfg.cell_size = float(cell_size)
    cfg.device = int(device)
    h = C.c_void_p()
    t = C.c_double()
    rc = lib.gdem_read_reload(C.byref(cfg), str(main_path).encode(), int(id_offset),
                              int(wall_id_offset), C.byref(h), C.byref(t))
    if rc != OK:
        raise _ERRORS.get(rc, GdemError)(rc, "gdem_read_reload", -1, -1)
    s = GdemSystem.__new__(GdemSystem)
    s.lib = lib
    s.h = h
    n = C.c_int64()
    nw, cv, mm = C.c_int32(), C.c_int32(), C.c_int32()
    s._check(lib.gdem_info(h, C.byref(n), C.byref(nw), C.byref(cv), C.byref(mm)), "gdem_info")
    s.n = n.value
    return s, t.value


def build_system(wl, device=0):
    """Create a GdemSystem from a workload dict (pacific_amd.workloads)."""
    s = GdemSystem(wl["origin"], wl["size"], wl["periodic"], wl["cell_size"], wl["gravity"], device)
    for (a, b, kind, p) in wl["models"]:
        s.set_contact_model(a, b, kind, p)
    for wdef in wl["walls"]:
        s.add_wall(*wdef)
    for rdef in wl.get("rectangles", []):
        s.add_rectangle_wall(*rdef)
    p = wl["particles"]
    if "narrow_phase" in wl:
        s.set_narrow_phase(wl["narrow_phase"])
    if "shapes" in wl:
        for sh in wl["shapes"]:
            if sh[0] == "box":
                s.add_box_shape(sh[1])
            elif sh[0] == "polyhedron":
                s.add_polyhedron_shape(sh[1], sh[2])
            else:
                raise ValueError(f"unknown shape {sh[0]}")
        s.set_convex_particles(p["x"], p["v"], p["w"], p["q"], p["basis"], p["shape"],
                               p["density"], p["crust"], p["material"])
    else:
        s.set_particles(p["x"], p["v"], p["w"], p["q"], p["radius"], p["density"], p["crust"],
                        p["material"])
    if "dt" in wl:
        s.set_time_step(wl["dt"])
    if "time" in wl:
        s.set_time(wl["time"])
    add_obstacle_tree(s, wl)
    return s


def add_obstacle_tree(s, wl):
    """Composites, shells and imposed motions of a workload dict, in the
    order both the library and the oracle create them: composites
    (parent, member walls), then shells (basis, centre, height, radius, width,
    n, crust, material, parent), box elements then rectangle elements
    ("rect_shells"), then further members (wall, composite), then wall and
    composite motions.  Rectangle walls ("rectangles": basis, centre, lx, ly,
    crust, material) follow the box walls."""
    for (parent, members) in wl.get("composites", []):
        k = s.add_composite(parent)
        for w in members:
            s.attach_wall(w, k)
    for sh in wl.get("shells", []):
        s.add_cylindrical_shell(*sh)
    for sh in wl.get("rect_shells", []):
        s.add_cylindrical_shell_rectangles(*sh)
    for (wall, comp) in wl.get("attach", []):
        s.attach_wall(wall, comp)
    for (wall, comp, mask) in wl.get("restrictions", []):
        s.set_motion_restriction(mask, wall, comp)
    for (wall, kind, t0, t1, prm) in wl.get("motions", []):
        s.add_wall_motion(wall, kind, t0, t1, prm)
    for (comp, kind, t0, t1, prm) in wl.get("composite_motions", []):
        s.add_composite_motion(comp, kind, t0, t1, prm)
    for (wall, t0, t1, prm) in wl.get("wall_forces", []):
        s.add_wall_force(wall, t0, t1, prm)
    for (comp, t0, t1, prm) in wl.get("composite_forces", []):
        s.add_composite_force(comp, t0, t1, prm)


def cylindrical_shell_boxes(basis, centre, height, radius, width, n):
    """The box walls of a box-element CylindricalShell (no context needed):
    (basis (n, 9), centre (n, 3), half edges (n, 3))."""
    lib = load_library()
    b, c, h = np.empty((n, 9)), np.empty((n, 3)), np.empty((n, 3))
    rc = lib.gdem_cylindrical_shell_boxes(_ptr(_f64(basis, (9,))), _ptr(_f64(centre, (3,))),
                                          float(height), float(radius), float(width), int(n),
                                          _ptr(b), _ptr(c), _ptr(h))
    if rc:
        raise GdemError(rc, "gdem_cylindrical_shell_boxes")
    return b, c, h


def cylindrical_shell_rectangles(basis, centre, height, radius, width, n):
    """The rectangle walls of a CylindricalShell with its default Rectangle
    elements: (basis (n, 9), centre (n, 3), half sides (n, 3): height/2,
    ly/2, 0)."""
    lib = load_library()
    b, c, h = np.empty((n, 9)), np.empty((n, 3)), np.empty((n, 3))
    rc = lib.gdem_cylindrical_shell_rectangles(_ptr(_f64(basis, (9,))), _ptr(_f64(centre, (3,))),
                                               float(height), float(radius), float(width), int(n),
                                               _ptr(b), _ptr(c), _ptr(h))
    if rc:
        raise GdemError(rc, "gdem_cylindrical_shell_rectangles")
    return b, c, h


def _batch_system(wl):
    """gdem_batch_system record of a workload dict (and the arrays it points to)."""
    p = wl["particles"]
    n = len(p["radius"])
    bs = _BatchSystem()
    for d in range(3):
        bs.cfg.origin[d] = float(wl["origin"][d])
        bs.cfg.size[d] = float(wl["size"][d])
        bs.cfg.periodic[d] = int(bool(wl["periodic"][d]))
        bs.cfg.gravity[d] = float(wl["gravity"][d])
    bs.cfg.cell_size = float(wl["cell_size"])
    keep = [_f64(p["x"], (n, 3)), _f64(p["v"], (n, 3)), _f64(p["w"], (n, 3)),
            _f64(p["q"], (n, 4)), _f64(p["radius"], (n,)), _f64(p["density"], (n,)),
            _f64(p["crust"], (n,))]
    bs.n = n
    bs.x, bs.v, bs.w, bs.q, bs.radius, bs.density, bs.crust = [a.ctypes.data for a in keep]
    models = {(a, b): (kind, prm) for (a, b, kind, prm) in wl["models"]}
    kind, pp = models[(0, 0)]
    kind_w, pw = models.get((0, 1), models[(0, 0)])
    if kind_w != kind:
        raise ValueError("particle-particle and particle-wall models must be of one kind")
    bs.kind = int(kind)
    for k in range(5):
        bs.pp[k] = float(pp[k])
        bs.pw[k] = float(pw[k])
    walls = wl["walls"]
    bs.nwalls = len(walls)
    if walls:
        wb = _f64([w[0] for w in walls], (len(walls), 9))
        wc = _f64([w[1] for w in walls], (len(walls), 3))
        wh = _f64([w[2] for w in walls], (len(walls), 3))
        keep += [wb, wc, wh]
        bs.wall_basis, bs.wall_centre, bs.wall_half = wb.ctypes.data, wc.ctypes.data, wh.ctypes.data
        bs.wall_crust = float(walls[0][3])
    bs.substeps = int(wl.get("substeps", 1))
    return bs, keep


class GdemBatch:
    """Many independent sphere systems advanced by one persistent launch per
    call (include/gdem.h, gdem_batch_*).  Each system is a workload dict with
    one particle material (0) and walls of material 1, as
    pacific_amd.workloads.batch_systems builds them."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = C.c_void_p()
        self._check(self.lib.gdem_batch_create(int(device), C.byref(h)), "gdem_batch_create")
        self.h = h
        self.sizes = []

    def close(self):
        if getattr(self, "h", None):
            self.lib.gdem_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc == OK:
            return
        sysid = id0 = id1 = -1
        if getattr(self, "h", None):
            s, code, a, b = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
            self.lib.gdem_batch_last_error(self.h, C.byref(s), C.byref(code), C.byref(a), C.byref(b))
            sysid, id0, id1 = s.value, a.value, b.value
        err = _ERRORS.get(rc, GdemError)(rc, f"{what} (system {sysid})", id0, id1)
        err.system = sysid
        raise err

    def add_system(self, wl):
        bs, keep = _batch_system(wl)
        sid = C.c_int32()
        self._check(self.lib.gdem_batch_add_system(self.h, C.byref(bs), C.byref(sid)),
                    "gdem_batch_add_system")
        self.sizes.append(bs.n)
        return sid.value

    def commit(self):
        self._check(self.lib.gdem_batch_commit(self.h), "gdem_batch_commit")

    def set_launch(self, block_size, workgroups=0):
        self._check(self.lib.gdem_batch_set_launch(self.h, int(block_size), int(workgroups)),
                    "gdem_batch_set_launch")

    def prime(self):
        self._check(self.lib.gdem_batch_prime(self.h), "gdem_batch_prime")

    def step(self, nsteps, dt):
        self._check(self.lib.gdem_batch_step(self.h, int(nsteps), float(dt)), "gdem_batch_step")

    def synchronize(self):
        self._check(self.lib.gdem_batch_synchronize(self.h), "gdem_batch_synchronize")

    def get_state(self, sys):
        n = self.sizes[sys]
        out = {k: np.empty((n, 3)) for k in ("x", "v", "w", "F", "T")}
        out["q"] = np.empty((n, 4))
        self._check(self.lib.gdem_batch_get_state(self.h, int(sys), _ptr(out["x"]), _ptr(out["v"]),
                                                  _ptr(out["w"]), _ptr(out["q"]), _ptr(out["F"]),
                                                  _ptr(out["T"])), "gdem_batch_get_state")
        return out

    def set_kernel_timing(self, on):
        self._check(self.lib.gdem_batch_set_kernel_timing(self.h, int(bool(on))),
                    "gdem_batch_set_kernel_timing")

    def counters(self):
        st, ct, ln = C.c_int64(), C.c_int64(), C.c_int64()
        ms = C.c_double()
        self._check(self.lib.gdem_batch_get_counters(self.h, C.byref(st), C.byref(ct), C.byref(ms),
                                                     C.byref(ln)), "gdem_batch_get_counters")
        return {"steps": st.value, "contacts": ct.value, "kernel_ms": ms.value, "launches": ln.value}


def build_batch(systems, device=0, block_size=256, workgroups=0):
    """A committed GdemBatch holding every workload dict of systems."""
    b = GdemBatch(device)
    for wl in systems:
        b.add_system(wl)
    b.commit()
    b.set_launch(block_size, workgroups)
    return b


class GdemDiff:
    """Loss and adjoint gradient of calibration trajectories (include/gdem.h,
    gdem_diff_*): systems as for GdemBatch, Hooke with one (kn, en, mu) per
    system."""

    def __init__(self, systems, nsteps, dt, checkpoint_every, obs_steps, device=0, block_size=256,
                 workgroups=0):
        self.lib = load_library()
        h = C.c_void_p()
        self._check(self.lib.gdem_diff_create(int(device), C.byref(h)), "gdem_diff_create")
        self.h = h
        self.sizes = []
        for wl in systems:
            bs, keep = _batch_system(wl)
            sid = C.c_int32()
            self._check(self.lib.gdem_diff_add_system(self.h, C.byref(bs), C.byref(sid)),
                        "gdem_diff_add_system")
            self.sizes.append(bs.n)
        obs = np.ascontiguousarray(obs_steps, dtype=np.int32)
        self.nobs = len(obs)
        self._check(self.lib.gdem_diff_commit(self.h, int(nsteps), float(dt), int(checkpoint_every),
                                              self.nobs, _ptr(obs)), "gdem_diff_commit")
        self._check(self.lib.gdem_diff_set_launch(self.h, int(block_size), int(workgroups)),
                    "gdem_diff_set_launch")
        self.nsys = len(self.sizes)
        self.ntot = int(sum(self.sizes))
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)

    def close(self):
        if getattr(self, "h", None):
            self.lib.gdem_diff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc == OK:
            return
        sysid = id0 = id1 = -1
        if getattr(self, "h", None):
            s, code, a, b = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
            self.lib.gdem_diff_last_error(self.h, C.byref(s), C.byref(code), C.byref(a), C.byref(b))
            sysid, id0, id1 = s.value, a.value, b.value
        err = _ERRORS.get(rc, GdemError)(rc, f"{what} (system {sysid})", id0, id1)
        err.system = sysid
        raise err

    def set_parameters(self, theta=None, v0=None):
        th = None if theta is None else _f64(theta, (self.nsys, 3))
        vv = None if v0 is None else _f64(v0, (self.ntot, 3))
        self._check(self.lib.gdem_diff_set_parameters(self.h, _ptr(th), _ptr(vv)),
                    "gdem_diff_set_parameters")

    def set_targets(self, target):
        t = _f64(target, (self.nobs, self.ntot, 3))
        self._check(self.lib.gdem_diff_set_targets(self.h, _ptr(t)), "gdem_diff_set_targets")

    def record(self):
        out = np.empty((self.nobs, self.ntot, 3))
        self._check(self.lib.gdem_diff_record(self.h, _ptr(out)), "gdem_diff_record")
        return out

    def loss(self):
        out = np.empty(self.nsys)
        self._check(self.lib.gdem_diff_loss(self.h, _ptr(out)), "gdem_diff_loss")
        return out

    def directional(self, dtheta, dv0):
        """Forward-mode derivative of each system's loss along (dtheta, dv0)."""
        dt = _f64(dtheta, (self.nsys, 3))
        dv = _f64(dv0, (self.ntot, 3))
        loss = np.empty(self.nsys)
        dl = np.empty(self.nsys)
        self._check(self.lib.gdem_diff_directional(self.h, _ptr(dt), _ptr(dv), _ptr(loss), _ptr(dl)),
                    "gdem_diff_directional")
        return loss, dl

    def gradient(self):
        loss = np.empty(self.nsys)
        gth = np.empty((self.nsys, 3))
        gv0 = np.empty((self.ntot, 3))
        self._check(self.lib.gdem_diff_gradient(self.h, _ptr(loss), _ptr(gth), _ptr(gv0)),
                    "gdem_diff_gradient")
        return loss, gth, gv0
