"""The reference's own reload records (tests/golden/reload2014, written by
Grains3D) decode with the layout the library reads and writes: every byte
count matches the record, the basis is the rotation of the stored quaternion
(row by row), the quaternion is a unit one."""
import numpy as np
import pytest

from reload2014_util import decode, load, rotation


@pytest.mark.parametrize("name", ["insertA_Particles.result.bin", "simulA_Particles.result.bin"])
def test_reference_records_decode(name):
    recs = decode(load(name))
    assert len(recs) == 1
    r = recs[0]
    assert (r["type"], r["id"], r["tag"], r["active"]) == (0, 1, 0, 1)
    # translation | rotation (Transform.hh:246-256)
    assert r["ttype"] == 3
    # unit to the rounding the reference accumulates over its steps (q <- dq q)
    assert np.linalg.norm(r["q_wxyz"]) == pytest.approx(1.0, abs=1e-12)
    # the reference updates the basis step by step; it stays R(q) to rounding
    np.testing.assert_allclose(r["basis"], rotation(r["q_wxyz"]), rtol=0, atol=1e-12)
    assert r["mem"] == []


def test_simulated_record_has_moved():
    r = decode(load("simulA_Particles.result.bin"))[0]
    r0 = decode(load("insertA_Particles.result.bin"))[0]
    assert r["origin"][2] > r0["origin"][2]   # the sphere rose
    assert np.all(r0["v"] == 0.0) and r["v"][2] > 0.0
