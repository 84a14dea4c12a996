"""Independent decoder of the reference's binary reload records (2014 format,
hybrid "_Particles.result.bin"), used to check the library's reader and
writer: a size_t byte count, then int type, id, tag, unsigned activity,
origin (3 doubles), basis (9, row by row), unsigned transform type, velocity
(3), quaternion (w, x, y, z), angular velocity (3), force (3), torque (3),
int map size, and per map entry int id0, id1, id2, active and 9 doubles."""
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reload2014")


def decode(blob):
    recs, at = [], 0
    while at < len(blob):
        nb, = struct.unpack_from("<Q", blob, at)
        at += 8
        end = at + nb
        r = {}
        r["type"], r["id"], r["tag"], r["active"] = struct.unpack_from("<3iI", blob, at)
        at += 16
        r["origin"] = np.array(struct.unpack_from("<3d", blob, at)); at += 24
        r["basis"] = np.array(struct.unpack_from("<9d", blob, at)).reshape(3, 3); at += 72
        r["ttype"], = struct.unpack_from("<I", blob, at); at += 4
        r["v"] = np.array(struct.unpack_from("<3d", blob, at)); at += 24
        r["q_wxyz"] = np.array(struct.unpack_from("<4d", blob, at)); at += 32
        r["w"] = np.array(struct.unpack_from("<3d", blob, at)); at += 24
        r["F"] = np.array(struct.unpack_from("<3d", blob, at)); at += 24
        r["T"] = np.array(struct.unpack_from("<3d", blob, at)); at += 24
        nm, = struct.unpack_from("<i", blob, at); at += 4
        mem = []
        for _ in range(nm):
            ids = struct.unpack_from("<4i", blob, at); at += 16
            vals = np.array(struct.unpack_from("<9d", blob, at)); at += 72
            mem.append((ids, vals))
        r["mem"] = mem
        assert at == end, (at, end)
        recs.append(r)
    return recs


def rotation(q_wxyz):
    w, x, y, z = q_wxyz
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def load(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()
