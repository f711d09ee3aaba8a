"""The uneven persistent batch (a fleet): instances of different hulls, space and connection counts in one launch,
handed out by a work counter and refilled as simulations end.

- Every instance of a fleet of seeded synthetic hulls (1,000-4,000 columns, 8-48 spaces) and titanic64 equals the
  oracle's CPU run of that instance, in both numerics, with the grid forced down to a few workgroups so that each
  workgroup runs many instances in turn (the refill path), and at the default grid.
- A fleet that holds the generator's largest hulls (16,000 columns, 128 spaces, 600 connections) beside smaller ones
  equals the oracle instance for instance over 1,200 s, in both numerics (one of those hulls is pinned against the
  reference's own CPU engine in tests/test_synthetic_reference.py).
- A fleet of one model equals the plain batch of that model.
- The packed 40-byte double records give the same runs as the padded 48-byte ones (the oracle's, for a fleet), and
  the autotuner measures both layouts.
"""
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _models():
    from sinksim_amd.data_paths import titanic
    from sinksim_amd.synth import synthetic_hull
    models = [synthetic_hull(seed, columns=cols, spaces=sp) for seed, cols, sp in
              [(3, 1000, 8), (11, 2500, 24), (17, 4000, 48), (23, 1800, 12)]]
    models.append(titanic("titanic"))
    return models


def _tables(models, inst_model, seed):
    from sinksim_amd.workloads import random_tables
    en, area, tOn = [], [], []
    for i, m in enumerate(inst_model):
        e, a, t = random_tables(models[m][1], 1, seed=seed + i, activation_share=0.1)
        en.append(e[0]); area.append(a[0]); tOn.append(t[0])
    return en, area, tOn


def _check_fleet(models, inst_model, numerics, t_max, every, workgroups, tpb=0, record_bytes=48):
    import pyoracle
    from sinksim_amd.batch import Batch, BatchOptions
    en, area, tOn = _tables(models, inst_model, seed=zlib.crc32(numerics.encode()))
    b = Batch.fleet(models, inst_model, BatchOptions(count=len(inst_model), mixed_precision=numerics == "portable32",
                                                     traced_instances=0, curve_every=every, t_max=t_max))
    b.set_connections(np.concatenate(en), np.concatenate(area), np.concatenate(tOn))
    b.set_launch(tpb, workgroups)
    b.set_record_layout(record_bytes)
    b.run()
    for i, m in enumerate(inst_model):
        ship, sim = models[m]
        cpu = pyoracle.sweep(ship, sim, numerics, en[i][None], area[i][None], tOn[i][None], t_max, every, threads=1)[0]
        r = b.result(i)
        where = f"{numerics} instance {i} (model {m}, {ship.columns} columns, {ship.nodes} spaces)"
        assert r.steps == cpu["steps"], where
        assert r.foundered == cpu["foundered"] and r.founderT == cpu["founderT"] and r.tEnd == cpu["tEnd"], where
        assert r.events == cpu["events"], where
        g = b.curve(i)
        assert g.shape == cpu["curve"].shape, where
        bad = np.argwhere(g != cpu["curve"])
        assert bad.size == 0, f"{where}: curve record {bad[0][0]} value {bad[0][1]}"
    assert b.total_steps() == int(b.results()["steps"].sum())
    b.close()


@pytest.mark.parametrize("tpb", [256, 512, 1024])
@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_fleet_matches_oracle_with_refill(numerics, tpb):
    models = _models()
    inst_model = np.array([4, 0, 1, 2, 3, 0, 2, 1, 3, 4, 0, 2], np.int32)
    _check_fleet(models, inst_model, numerics, t_max=1200.0, every=30.0, workgroups=3, tpb=tpb)


# the generator's largest hulls (the calibration hull is seed 5 at these sizes) and two drawn sizes
LARGEST = [dict(seed=5, columns=16000, spaces=128, connections=600), dict(seed=31, columns=16000, spaces=128, connections=600),
           dict(seed=47, columns=16000, spaces=128, connections=600), dict(seed=8), dict(seed=19)]


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_fleet_of_the_largest_hulls_matches_oracle(numerics):
    from sinksim_amd.synth import synthetic_hull
    models = [synthetic_hull(**kw) for kw in LARGEST]
    assert max(m[0].columns for m in models) >= 16000 and max(m[0].nodes for m in models) == 128
    assert max(m[1].connections for m in models) == 600
    inst_model = np.array([0, 3, 1, 4, 2, 0, 1], np.int32)
    _check_fleet(models, inst_model, numerics, t_max=1200.0, every=30.0, workgroups=3)


def test_fleet_matches_oracle_full_grid():
    models = _models()
    inst_model = np.array([2, 1, 0, 3, 4, 2, 0], np.int32)
    _check_fleet(models, inst_model, "portable", t_max=900.0, every=60.0, workgroups=0)


def test_single_model_fleet_equals_batch():
    from sinksim_amd.batch import Batch, BatchOptions
    from sinksim_amd.data_paths import titanic
    from sinksim_amd.workloads import random_tables
    ship, sim = titanic("titanic")
    en, area, tOn = random_tables(sim, 16, seed=5)
    opt = BatchOptions(count=16, mixed_precision=False, traced_instances=0, curve_every=60.0, t_max=1800.0)
    a = Batch(ship, sim, opt)
    a.set_connections(en, area, tOn)
    a.run()
    f = Batch.fleet([(ship, sim)], np.zeros(16, np.int32), opt)
    f.set_connections(en, area, tOn)
    f.set_workgroups(5)
    f.run()
    for i in range(16):
        assert np.array_equal(a.curve(i), f.curve(i)), i
        ra, rf = a.result(i), f.result(i)
        assert ra.steps == rf.steps and ra.events == rf.events and np.array_equal(ra.vol, rf.vol), i


@pytest.mark.parametrize("fleet", [False, True])
def test_chunked_launches_equal_one_launch(fleet):
    """stepsPerLaunch > 0 replays the captured launch graph once per chunk (counter reset + kernel); every replay
    continues the instances where the last one stopped, and the results equal one launch to the end bit for bit."""
    from sinksim_amd.batch import Batch, BatchOptions
    models = _models()
    inst_model = np.array([4, 0, 1, 2, 3, 1, 4, 0], np.int32) if fleet else np.full(8, 4, np.int32)
    en, area, tOn = _tables(models, inst_model, seed=99)
    runs = []
    for spl in (0, 700):
        opt = BatchOptions(count=len(inst_model), mixed_precision=False, traced_instances=0, curve_every=30.0, t_max=2400.0,
                           steps_per_launch=spl)
        b = Batch.fleet(models, inst_model, opt)
        b.set_connections(np.concatenate(en), np.concatenate(area), np.concatenate(tOn))
        b.set_workgroups(3)
        b.run()
        runs.append(([b.curve(i) for i in range(len(inst_model))], [b.result(i) for i in range(len(inst_model))], b.kernel_time()[1]))
        b.close()
    (c0, r0, l0), (c1, r1, l1) = runs
    assert l0 == 1 and l1 > 3
    for i in range(len(inst_model)):
        assert np.array_equal(c0[i], c1[i]), i
        assert r0[i].steps == r1[i].steps and r0[i].events == r1[i].events and np.array_equal(r0[i].vol, r1[i].vol), i


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_512_threads_reproduce_golden_trace(titanic, goldens, numerics):
    """Threads-per-simulation 512 (buoyancy and level solve side by side) on the reference's golden traces."""
    from sinksim_amd.batch import Batch, BatchOptions
    ship, sim = titanic
    gold = goldens[numerics]
    b = Batch(ship, sim, BatchOptions(count=4, mixed_precision=numerics == "portable32", traced_instances=1,
                                      dense_steps=gold.denseSteps, sample_every=gold.sampleEvery, t_max=5 * 3600.0))
    b.set_launch(512, 0)
    b.run()
    dense, samples = b.trace(0)
    assert np.array_equal(dense, gold.dense) and np.array_equal(samples, gold.samples)
    r = b.result(0)
    assert r.steps == gold.steps and r.founderT == gold.founderT


def test_autotune_picks_a_shape_and_keeps_results():
    """The autotuner measures the three workgroup sizes on a trial copy, sets the fastest shape, and the tuned run
    equals an untuned one instance for instance."""
    from sinksim_amd.batch import Batch, BatchOptions
    models = _models()
    inst_model = np.array([2, 1, 0, 3, 4, 2, 0, 1, 3, 4], np.int32)
    en, area, tOn = _tables(models, inst_model, seed=7)
    out = []
    for tune in (False, True):
        b = Batch.fleet(models, inst_model, BatchOptions(count=len(inst_model), mixed_precision=False, traced_instances=0,
                                                         curve_every=60.0, t_max=900.0))
        b.set_connections(np.concatenate(en), np.concatenate(area), np.concatenate(tOn))
        if tune:
            r = b.autotune(sample=8, trial_steps=200)
            assert r["threads_per_sim"] in (256, 512, 1024) and r["rate_256"] > 0 and r["rate_512"] > 0 and r["rate_1024"] > 0
            assert {t["threads_per_sim"] for t in r["trials"]} == {256, 512, 1024}, r
            # the double records' two layouts are measured (the last trial: the other layout at the fastest shape)
            assert r["record_bytes"] in (48, 40) and {t["record_bytes"] for t in r["trials"]} == {48, 40}, r
        b.run()
        out.append([b.curve(i) for i in range(len(inst_model))])
        b.close()
    for a, c in zip(*out):
        assert np.array_equal(a, c)


def test_autotune_grid_choice_keeps_results():
    """With more instances than are resident at once, the autotuner also measures refilling grids of three quarters
    and half the resident workgroups; whatever it picks, the run equals an untuned one instance for instance."""
    from sinksim_amd.batch import Batch, BatchOptions
    models = _models()[:2]
    n = 6000
    inst_model = (np.arange(n) % 2).astype(np.int32)
    out = []
    for tune in (False, True):
        b = Batch.fleet(models, inst_model, BatchOptions(count=n, mixed_precision=False, traced_instances=0, curve_every=30.0,
                                                         t_max=60.0))
        if tune:
            r = b.autotune(sample=n, trial_steps=60)
            grids = {t["workgroups"] for t in r["trials"]}
            assert len(r["trials"]) >= 3 and any(g > 0 for g in grids), r
            assert r["workgroups"] in grids and r["threads_per_sim"] in (256, 512, 1024)
        b.run()
        out.append((b.results(), [b.curve(i) for i in range(0, n, 97)]))
        b.close()
    (r0, c0), (r1, c1) = out
    assert np.array_equal(r0["steps"], r1["steps"])
    for a, c in zip(c0, c1):
        assert np.array_equal(a, c)


def test_packed_records_equal_the_oracle():
    """The packed 40-byte double geometry records (ss_batch_set_record_layout) give the oracle's runs bit for bit, as
    the padded 48-byte ones do: a fleet through the refilling kernel (three workgroups for ten instances) and at the
    default grid, portable numerics (the float records have one layout)."""
    models = _models()
    inst_model = np.array([2, 1, 0, 3, 4, 2, 0, 1, 3, 4], np.int32)
    _check_fleet(models, inst_model, "portable", 900.0, 60.0, workgroups=3, record_bytes=40)
    _check_fleet(models, inst_model, "portable", 900.0, 60.0, workgroups=0, record_bytes=40)


def test_packed_records_keep_a_titanic_batch_in_epochs():
    """A titanic64 batch in epochs (every instance resident, the 256-, 512- and 1,024-thread forms) with the packed
    records equals the padded records' run: steps, events, curves, and a traced instance's dense and sampled records;
    other sizes are refused, and the layout cannot change after a run."""
    from sinksim_amd import sweep as S
    from sinksim_amd._capi import EngineError
    from sinksim_amd.batch import Batch, BatchOptions
    from sinksim_amd.data_paths import titanic
    from sinksim_amd.workloads import titanic_batch_sweep
    ship, sim = titanic("titanic")
    sw = titanic_batch_sweep(sim, count=300, t_max=3 * 3600.0)
    en, area, tOn = S.connection_tables(sw, sim)
    out = []
    for rec in (48, 40):
        b = Batch(ship, sim, BatchOptions(count=300, mixed_precision=False, traced_instances=1, dense_steps=50, sample_every=300.0,
                                          curve_every=sw.every, t_max=sw.tMax))
        b.set_connections(en, area, tOn)
        with pytest.raises(ValueError):  # SS_INVALID_ARGUMENT
            b.set_record_layout(44)
        b.set_record_layout(rec)
        b.run()
        with pytest.raises(EngineError):  # SS_LOGIC
            b.set_record_layout(48)
        out.append((b.results(), [b.curve(i) for i in range(300)], b.trace(0), [b.result(i).events for i in range(300)]))
        b.close()
    (r0, c0, t0, e0), (r1, c1, t1, e1) = out
    for k in ("steps", "foundered", "founderT", "tEnd"):
        assert np.array_equal(r0[k], r1[k]), k
    assert e0 == e1
    for a, c in zip(c0, c1):
        assert np.array_equal(a, c)
    for a, c in zip(t0, t1):
        assert np.array_equal(a, c)
