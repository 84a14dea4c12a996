"""Reload files in the reference's 2014 format through the C ABI
(gdem_write_particles_2014 / gdem_read_particles_2014, gdem_set_state,
gdem_get_memory / gdem_set_memory).

* the reference's own record (tests/golden/reload2014) is read bit for bit,
  and written back byte for byte except the basis, which the reference keeps
  step by step and the library rebuilds from the quaternion (equal to 1e-12);
* a system with contact memory round-trips through the text and the binary
  file bit for bit (state and memory);
* a restart continues the run: the stored F, T are the leap-frog's a0, so no
  force evaluation (which would advance every contact's spring again) comes
  between; after 20 more steps the restored and the uninterrupted runs agree
  to 1e-12 (they differ only in the order the restored copy sums contacts)."""
import os
import tempfile

import numpy as np
import pytest

from gdem_testutil import rel_err
from pacific_amd import gdem, workloads
from reload2014_util import GOLDEN, decode, load

pytestmark = pytest.mark.gpu


def one_sphere_system():
    """The reference particle of the golden files: r 0.001, density 900,
    crust 0.00025, no walls (the rising sphere's fluid is not part of the step)."""
    wl = {"origin": (0.0, 0.0, 0.0), "size": (0.2, 0.2, 0.2), "periodic": (0, 0, 0),
          "cell_size": 0.002, "gravity": (0.0, 0.0, -9.81), "dt": 1e-5,
          "models": [(0, 0, gdem.MODEL_HOOKE, (1e5, 0.9, -1.0, 0.5, 0.0))], "walls": []}
    x = np.array([[0.1, 0.1, 0.1]])
    p = workloads._particles(x, np.zeros((1, 3)), np.array([1e-3]))
    p["density"] = np.array([900.0])
    p["crust"] = np.array([2.5e-4])
    wl["particles"] = p
    return gdem.build_system(wl)


def test_reads_reference_record_bit_for_bit():
    print("[case] reload: reference record", flush=True)
    rec = decode(load("simulA_Particles.result.bin"))[0]
    s = one_sphere_system()
    gt = s.read_particles_2014(os.path.join(GOLDEN, "simulA_Particles.result.bin"), binary=True,
                               id_offset=1)
    assert gt.tolist() == [0]
    st = s.get_state()
    q = rec["q_wxyz"]
    np.testing.assert_array_equal(st["x"][0], rec["origin"])
    np.testing.assert_array_equal(st["v"][0], rec["v"])
    np.testing.assert_array_equal(st["w"][0], rec["w"])
    np.testing.assert_array_equal(st["q"][0], [q[1], q[2], q[3], q[0]])
    np.testing.assert_array_equal(st["F"][0], rec["F"])
    np.testing.assert_array_equal(np.signbit(st["F"][0]), np.signbit(rec["F"]))
    np.testing.assert_array_equal(st["T"][0], rec["T"])
    # written back: the same bytes, the basis aside
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "p.bin")
        s.write_particles_2014(out, binary=True, id_offset=1)
        got = open(out, "rb").read()
    ref = load("simulA_Particles.result.bin")
    assert len(got) == len(ref)
    basis = slice(8 + 16 + 24, 8 + 16 + 24 + 72)
    assert got[:basis.start] == ref[:basis.start] and got[basis.stop:] == ref[basis.stop:]
    np.testing.assert_allclose(decode(got)[0]["basis"], rec["basis"], rtol=0, atol=1e-12)
    s.close()


def _stepped(name, nsteps):
    wl = (workloads.memory_box(1000, seed=1) if name == "box"
          else workloads.memory_periodic(20000, seed=2))
    r = wl["particles"]["radius"] * (1.08 if name == "box" else 1.02)
    wl["particles"]["radius"] = r
    wl["cell_size"] = 2.0 * (1.2 if any(wl["periodic"]) else 1.0) * r.max()
    s = gdem.build_system(wl)
    s.step(nsteps, wl["dt"])
    return wl, s


@pytest.mark.parametrize("binary", [True, False])
def test_round_trip_with_memory(binary):
    print(f"[case] reload round trip, binary {binary}", flush=True)
    wl, s = _stepped("box", 30)
    mem = s.get_memory()
    assert len(mem["owner"]) > 50
    st = s.get_state()
    r = gdem.build_system(wl)          # fresh copy at the initial state
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "p.bin" if binary else "p.txt")
        s.write_particles_2014(path, binary=binary, id_offset=1, wall_id_offset=5000)
        r.read_particles_2014(path, binary=binary, id_offset=1, wall_id_offset=5000)
    rt = r.get_state()
    for k in ("x", "v", "w", "q", "F", "T"):
        np.testing.assert_array_equal(rt[k], st[k], err_msg=k)
    rm = r.get_memory()
    for k in ("owner", "key", "kd", "nrm", "spr"):
        np.testing.assert_array_equal(rm[k], mem[k], err_msg=k)
    s.close()
    r.close()


def test_restart_continues_without_recomputing_a0():
    print("[case] reload restart", flush=True)
    wl, s = _stepped("periodic", 30)
    assert len(s.get_memory()["owner"]) > 50
    r = gdem.build_system(wl)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "p.bin")
        s.write_particles_2014(path)
        r.read_particles_2014(path)
    s.step(20, wl["dt"])
    r.step(20, wl["dt"])
    a, b = s.get_state(), r.get_state()
    x0 = wl["particles"]["x"]
    L = np.array(wl["size"])
    da, db = a["x"] - x0, b["x"] - x0
    da -= L * np.round(da / L)
    db -= L * np.round(db / L)
    assert rel_err(db, da) <= 1e-12
    for k in ("v", "w"):
        assert rel_err(b[k], a[k]) <= 1e-12, k
    s.close()
    r.close()
