"""Pins of the oracle's imposed forces on composite obstacles
(orc_add_composite_force: ObstacleImposedForce linked to a CompositeObstacle),
fixed by kinematics and statics:

* a lid made of two plates in one composite, above a sphere resting on a
  floor: until the composite feels a load it moves at vmax against the
  force's direction, both plates with it (equal heights, bit for bit), the
  composite's and the plates' velocity -vmax d;
* once the lid presses the sphere, the controller brings the composite's
  load, the sum of its plates' loads, to the imposed force A and holds it;
  the floor then carries A plus the sphere's weight;
* the composite's load is the sum its torsor keeps after each step, so at
  the first step it is zero: a lid already pressing the sphere still moves at
  vmax for that step (a wall force on one plate would not)."""
import numpy as np
import pytest

from pacific_amd import workloads
from test_oracle_imposed_force_pins import DT, piston, weight


def composite_piston(A, mass, gap=1.0e-4, vmax=0.05):
    wl = piston(A, mass, gap=gap, vmax=vmax)
    lid = wl["walls"][1]
    z = lid[1][2]
    # two half plates side by side: the sphere (at the origin) under the first
    wl["walls"] = [wl["walls"][0],
                   (workloads.IDENTITY, (0.0, 0.0, z), (0.005, 0.01, 1.0e-3), workloads.CRUST, 1),
                   (workloads.IDENTITY, (0.0101, 0.0, z), (0.005, 0.01, 1.0e-3), workloads.CRUST, 1)]
    wl["wall_forces"] = []
    wl["composites"] = [(-1, [1, 2])]
    wl["composite_forces"] = [(0, 0.0, 1.0, (0.0, 0.0, A, mass, vmax))]
    return wl


def test_approach_at_vmax_together(oracle_mod):
    A = 10.0 * weight()
    o = oracle_mod.Oracle(composite_piston(A, 2.5e-5))
    o.prime()
    z0 = o.wall(1)[0][2]
    for k in range(1, 301):
        o.step(1, DT)
        c1, _, v1, _ = o.wall(1)
        c2, _, v2, _ = o.wall(2)
        assert c1[2] == c2[2]
        np.testing.assert_array_equal(v1, [0.0, 0.0, -0.05])
        np.testing.assert_array_equal(v2, [0.0, 0.0, -0.05])
        assert c1[2] == pytest.approx(z0 - 0.05 * k * DT, abs=1e-15)
    _, cv, _ = o.composite(0)
    np.testing.assert_array_equal(cv, [0.0, 0.0, -0.05])


@pytest.mark.parametrize("factor", [3.0, 10.0])
def test_composite_load_reaches_imposed_force(oracle_mod, factor):
    mg = weight()
    A = factor * mg
    o = oracle_mod.Oracle(composite_piston(A, 2.5e-5))
    o.prime()
    o.step(2400, DT)
    f, _ = o.wall_loads()
    assert f[1][2] + f[2][2] == pytest.approx(A, rel=1e-8)
    assert f[2][2] == 0.0                       # nothing under the second plate
    assert f[0][2] == pytest.approx(-(A + mg), rel=1e-8)
    assert o.wall(1)[0][2] == o.wall(2)[0][2]
    assert abs(o.composite(0)[1][2]) < 1e-9     # stopped


def test_first_step_sees_no_composite_load(oracle_mod):
    mg = weight()
    A = 3.0 * mg
    # the lid starts pressing the sphere (gap -2e-7)
    wl = composite_piston(A, 2.5e-5, gap=-2.0e-7)
    o = oracle_mod.Oracle(wl)
    o.prime()
    assert o.wall_loads()[0][1][2] > 0.0        # the plate is loaded at the start
    z0 = o.wall(1)[0][2]
    o.step(1, DT)
    assert o.wall(1)[0][2] == pytest.approx(z0 - 0.05 * DT, abs=1e-18)
    # a force on the plate itself reads its load at once: no vmax step
    wl1 = piston(A, 2.5e-5, gap=-2.0e-7)
    o1 = oracle_mod.Oracle(wl1)
    o1.prime()
    z1 = o1.wall(1)[0][2]
    o1.step(1, DT)
    assert abs(o1.wall(1)[0][2] - z1) < 0.5 * 0.05 * DT
