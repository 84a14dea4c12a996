"""Pins of the oracle's imposed-force walls (orc_add_wall_force: the
reference's ObstacleImposedForce, ConstantTranslation), fixed by kinematics
and statics:

* until the wall first feels a load it moves at vmax against the force's
  direction: a plate above a sphere descends by vmax t, at velocity -vmax d;
* once it presses the sphere onto the floor, the controller brings the
  plate's load to the imposed force A and holds it, and the floor then
  carries A plus the sphere's weight (statics of the squeezed sphere)."""
import math

import numpy as np
import pytest

from pacific_amd import workloads
from pacific_amd.gdem import MODEL_HOOKE

HOOKE = (1.0e5, 0.5, -1.0, 0.3, 0.0)
DT = 5.0e-6


def piston(A, mass, gap=1.0e-4, vmax=0.05):
    wl = workloads.sphere_on_plate((MODEL_HOOKE, HOOKE), x0=(0.0, 0.0, None), dt=DT)
    wl["walls"] = wl["walls"] + [(workloads.IDENTITY, (0.0, 0.0, 2.0e-3 + gap + 1.0e-3),
                                  (0.01, 0.01, 1.0e-3), workloads.CRUST, 1)]
    wl["wall_forces"] = [(1, 0.0, 1.0, (0.0, 0.0, A, mass, vmax))]
    return wl


def weight():
    return workloads.DENSITY * 4.0 * math.pi * 1e-9 / 3.0 * 9.81


def test_approach_at_vmax(oracle_mod):
    A = 10.0 * weight()
    wl = piston(A, 2.5e-5)
    o = oracle_mod.Oracle(wl)
    o.prime()
    z0 = o.wall(1)[0][2]
    for k in range(1, 301):
        o.step(1, DT)
        c, _, v, _ = o.wall(1)
        np.testing.assert_array_equal(v, [0.0, 0.0, -0.05])
        assert c[2] == pytest.approx(z0 - 0.05 * k * DT, abs=1e-15)   # 300 roundings near 3e-3
    assert np.all(o.wall_loads()[0][1] == 0.0)


@pytest.mark.parametrize("factor", [3.0, 10.0])
def test_load_reaches_imposed_force(oracle_mod, factor):
    mg = weight()
    A = factor * mg
    wl = piston(A, 2.5e-5)
    o = oracle_mod.Oracle(wl)
    o.prime()
    o.step(2400, DT)      # 400 steps of approach, then the controller settles
    f, _ = o.wall_loads()
    assert f[1][2] == pytest.approx(A, rel=1e-8)
    assert f[0][2] == pytest.approx(-(A + mg), rel=1e-8)
    np.testing.assert_allclose(f[:, :2], 0.0, atol=1e-12 * A)
    assert abs(o.wall(1)[2][2]) < 1e-9      # the plate has stopped
