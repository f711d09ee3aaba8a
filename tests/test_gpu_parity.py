"""Parity of the HIP batch engine with the reference's numerics, through the C ABI.

- A traced instance of an 8-instance batch reproduces the reference's golden traces of titanic64/titanic in the
  portable and portable32 numerics at tolerance 0: every dense record, every sampled record, the events and the
  founder time (the reference's cuda_trace_* -> cuda_matches_portable* acceptance, tests/CMakeLists.txt:40-57).
- A sweep (curves, events, final readouts, step counts) equals the oracle's CPU sweep value for value (the
  reference's sweep_gpu_matches_cpu, tests/CMakeLists.txt:59-65), on the reference's smoke sweep and on seeded
  random sweeps over every titanic64 scenario.
- Identical instances of one batch give identical results (no cross-instance interference, no races).
"""
import os
import zlib

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _batch(ship, sim, **kw):
    from sinksim_amd.batch import Batch, BatchOptions
    return Batch(ship, sim, BatchOptions(**kw))


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_trace_matches_golden(titanic, goldens, numerics):
    ship, sim = titanic
    gold = goldens[numerics]
    b = _batch(ship, sim, count=8, mixed_precision=numerics == "portable32", traced_instances=1, dense_steps=gold.denseSteps,
               sample_every=gold.sampleEvery, t_max=5 * 3600.0)
    b.run()
    dense, samples = b.trace(0)
    assert dense.shape == gold.dense.shape
    assert samples.shape == gold.samples.shape
    bad = np.argwhere(dense != gold.dense)
    assert bad.size == 0, f"first dense difference at record {bad[0][0]}, field {bad[0][1]}"
    bad = np.argwhere(samples != gold.samples)
    assert bad.size == 0, f"first sample difference at record {bad[0][0]}, field {bad[0][1]}"
    r = b.result(0)
    assert r.foundered == gold.foundered and r.founderT == gold.founderT and r.steps == gold.steps
    ev = [(t, sim.event_id(k, i)) for t, i, k in r.events]
    assert ev == [(t, i) for t, i, _ in gold.events]
    res = b.results()
    assert np.all(res["steps"] == gold.steps) and np.all(res["founderT"] == gold.founderT)
    assert b.total_steps() == 8 * gold.steps


def _gpu_sweep(ship, sim, sw, numerics):
    from sinksim_amd import sweep as S
    en, area, tOn = S.connection_tables(sw, sim)
    b = _batch(ship, sim, count=len(sw.instances), mixed_precision=numerics == "portable32", traced_instances=0,
               curve_every=sw.every, t_max=sw.tMax)
    b.set_connections(en, area, tOn)
    b.run()
    out = []
    for i in range(len(sw.instances)):
        r = b.result(i)
        out.append({"steps": r.steps, "foundered": r.foundered, "founderT": r.founderT, "tEnd": r.tEnd,
                    "curve": b.curve(i), "events": r.events, "final": b.final_readouts(i)})
    return out, (en, area, tOn)


def _assert_same(gpu, cpu, ids):
    for i, (g, c) in enumerate(zip(gpu, cpu)):
        where = f"instance {ids[i]}"
        assert g["steps"] == c["steps"], where
        assert g["foundered"] == c["foundered"] and g["founderT"] == c["founderT"] and g["tEnd"] == c["tEnd"], where
        assert g["events"] == c["events"], where
        assert g["curve"].shape == c["curve"].shape, where
        bad = np.argwhere(g["curve"] != c["curve"])
        assert bad.size == 0, f"{where}: curve record {bad[0][0]} value {bad[0][1]}"
        assert np.array_equal(g["final"], c["curve"][-1]), where


@pytest.mark.parametrize("numerics", ["portable32", "portable"])
def test_smoke_sweep_matches_oracle(titanic, numerics):
    import pyoracle
    from sinksim_amd import sweep as S
    ship, sim = titanic
    sw = S.load_sweep(os.path.join(GOLDEN, "smoke.sweep.json"), sim)
    gpu, (en, area, tOn) = _gpu_sweep(ship, sim, sw, numerics)
    cpu = pyoracle.sweep(ship, sim, numerics, en, area, tOn, sw.tMax, sw.every, threads=8)
    _assert_same(gpu, cpu, [x.id for x in sw.instances])


@pytest.mark.parametrize("sim_id", ["four", "five", "hawke", "britannic", "britannicClosed", "titanicNoBR5", "titanicFour",
                                    "titanicD", "titanicB", "titanicDoors"])
def test_scenarios_match_oracle(sim_id):
    """Every titanic64 scenario, 6 seeded instances each (kind scales log-uniform in [0.5, 2], doors opened with
    p = 0.2, activation times in [0, 1800] s on 10 % of the connections), 1 h, both numerics."""
    import pyoracle
    from sinksim_amd.data_paths import titanic as load
    from sinksim_amd.workloads import random_tables
    ship, sim = load(sim_id)
    en, area, tOn = random_tables(sim, 6, seed=zlib.crc32(sim_id.encode()), activation_share=0.1)
    for numerics in ("portable", "portable32"):
        b = _batch(ship, sim, count=6, mixed_precision=numerics == "portable32", traced_instances=0, curve_every=30.0,
                   t_max=3600.0)
        b.set_connections(en, area, tOn)
        b.run()
        gpu = []
        for i in range(6):
            r = b.result(i)
            gpu.append({"steps": r.steps, "foundered": r.foundered, "founderT": r.founderT, "tEnd": r.tEnd,
                        "curve": b.curve(i), "events": r.events, "final": b.final_readouts(i)})
        cpu = pyoracle.sweep(ship, sim, numerics, en, area, tOn, 3600.0, 30.0, threads=8)
        _assert_same(gpu, cpu, [f"{sim_id}/{numerics}/{i}" for i in range(6)])


def test_instances_are_independent(titanic):
    """32 copies of one instance among 32 others with different tables: the copies agree bit for bit."""
    from sinksim_amd.workloads import random_tables
    ship, sim = titanic
    en, area, tOn = random_tables(sim, 64, seed=7)
    en[::2], area[::2], tOn[::2] = en[0], area[0], tOn[0]
    b = _batch(ship, sim, count=64, mixed_precision=False, traced_instances=0, curve_every=60.0, t_max=1800.0)
    b.set_connections(en, area, tOn)
    b.run()
    c0 = b.curve(0)
    for i in range(2, 64, 2):
        assert np.array_equal(b.curve(i), c0)
