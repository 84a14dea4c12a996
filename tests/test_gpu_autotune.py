"""The batched launch's autotuner (pacific_amd/autotune.py): it times every
(cell factor, workgroup size) pair on a sample of the systems, keeps the
fastest, caches the choice per workload, and its workgroup-size knob changes
only the speed (states bit-identical for every size)."""
import time

import numpy as np
import pytest

from pacific_amd import autotune, gdem, workloads

pytestmark = pytest.mark.gpu

DT = 1.0e-5


def _systems():
    return workloads.batch_systems(nsys=24, seed=11, nmin=64, nmax=300)


def test_tune_picks_the_fastest_and_caches_it(tmp_path):
    print("[case] autotune over 2 x 2 launch settings, 8 sampled systems", flush=True)
    systems = _systems()
    cache = str(tmp_path / "tune.json")
    res = autotune.tune(systems, DT, sample=8, steps=3, warmup=1, block_sizes=(64, 256),
                        cell_factors=(1.0, 1.2), cache=cache)
    table = res["table"]
    assert len(table) == 4
    assert {(cf, bs) for cf, bs, _ in table} == {(1.0, 64), (1.0, 256), (1.2, 64), (1.2, 256)}
    assert all(rate > 0 for _, _, rate in table)
    best = max(table, key=lambda r: r[2])
    assert (res["cell_factor"], res["block_size"]) == (best[0], best[1])
    t = time.perf_counter()
    again = autotune.tune(systems, DT, sample=8, steps=3, warmup=1, block_sizes=(64, 256),
                          cell_factors=(1.0, 1.2), cache=cache)
    assert time.perf_counter() - t < 1.0  # read from the cache, nothing timed
    assert (again["cell_factor"], again["block_size"]) == (res["cell_factor"], res["block_size"])
    assert [tuple(r) for r in again["table"]] == [tuple(r) for r in table]


def test_workgroup_size_changes_no_bit():
    print("[case] batch states after 30 steps at workgroups 64, 256 and 512", flush=True)
    systems = autotune.with_cell_factor(_systems(), 1.2)
    states = []
    for bs in (64, 256, 512):
        b = gdem.build_batch(systems)
        b.set_launch(bs, 0)
        b.prime()
        b.step(30, DT)
        b.synchronize()
        states.append([b.get_state(k) for k in range(len(systems))])
        b.close()
    for other in states[1:]:
        for s0, s1 in zip(states[0], other):
            for key in ("x", "v", "w", "q"):
                assert np.array_equal(s0[key], s1[key]), key
