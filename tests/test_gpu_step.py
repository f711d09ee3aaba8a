"""The single-step entry point (kernel::step over the batch, ss_batch_step) held to the reference's acceptance
procedure of apps/sinksim_check.cpp:110-126, through the C ABI, at tolerance 0:

- dense single steps: restore dense record i, step once, compare with dense record i + 1 (all 400 steps of the golden,
  one instance each);
- sampled windows: restore sampled record i, step round((t[i+1] - t[i]) / dt) times, compare with sampled record i + 1
  (all 157 windows of 60 s);

both in one batch with per-instance step counts, for the portable and portable32 goldens. A second call continues
from the device state (the dense instances step once more and meet record i + 2). Records are compared field for
field in the trace layout (t, pose, rates, inflow, buoyancy volume, hull top, and vol/level/cx/cy/cz of every node):
the fields the reference's compare() checks.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _restore_records(b, recs, nn):
    for i, r in enumerate(recs):
        b.restore(i, r[:7], r[10:10 + nn], r[10 + nn:10 + 2 * nn], r[10 + 2 * nn:10 + 3 * nn], r[10 + 3 * nn:10 + 4 * nn],
                  r[10 + 4 * nn:10 + 5 * nn])


def _first_difference(got, want, what):
    bad = np.argwhere(got != want)
    if bad.size:
        i, f = bad[0]
        return f"{what}: first difference in check {i}, field {f}: {got[i, f]!r} vs {want[i, f]!r} ({len(np.unique(bad[:, 0]))} checks differ)"
    return None


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_check_procedure_dense_steps_and_windows(titanic, goldens, numerics):
    from sinksim_amd.batch import Batch, BatchOptions
    ship, sim = titanic
    gold = goldens[numerics]
    nn = ship.nodes
    dense, samples = gold.dense, gold.samples
    nd, ns = len(dense) - 1, len(samples) - 1
    windows = np.array([int(round((samples[i + 1, 0] - samples[i, 0]) / sim.dt)) for i in range(ns)], np.int64)
    assert windows.max() == 240 and windows.min() >= 1
    count = nd + ns
    print(f"check procedure, {numerics}: {nd} dense steps and {ns} windows as {count} instances", flush=True)
    b = Batch(ship, sim, BatchOptions(count=count, mixed_precision=numerics == "portable32", traced_instances=0, t_max=5 * 3600.0))
    _restore_records(b, np.concatenate([dense[:nd], samples[:ns]]), nn)
    b.step(per_instance=np.concatenate([np.ones(nd, np.int64), windows]))
    got = b.state_records()
    assert got.shape == (count, 10 + 5 * nn)
    errs = [e for e in (_first_difference(got[:nd], dense[1:], "dense single steps"),
                        _first_difference(got[nd:], samples[1:], "sampled windows")) if e]
    assert not errs, "; ".join(errs)
    # steps count per instance and in total
    res = b.results()
    assert np.array_equal(res["steps"], np.concatenate([np.ones(nd, np.int64), windows]))
    assert b.total_steps() == nd + int(windows.sum())
    # a second call continues from the device state: the dense instances meet record i + 2
    b.step(per_instance=np.concatenate([np.ones(nd - 1, np.int64), np.zeros(1 + ns, np.int64)]))
    again = b.state_records()
    err = _first_difference(again[:nd - 1], dense[2:], "second dense step")
    assert err is None, err
    assert np.array_equal(again[nd:], got[nd:]), "instances given 0 steps changed"
    b.close()


def test_uniform_step_count_and_refusals(titanic, goldens):
    """A uniform count steps every instance alike (here 7 copies of dense record 10, 25 steps: all equal to dense record
    35); run() after a step and restore after a step are refused."""
    from sinksim_amd._capi import EngineError
    from sinksim_amd.batch import Batch, BatchOptions
    ship, sim = titanic
    gold = goldens["portable"]
    nn = ship.nodes
    b = Batch(ship, sim, BatchOptions(count=7, mixed_precision=False, traced_instances=0, t_max=5 * 3600.0))
    _restore_records(b, np.repeat(gold.dense[10:11], 7, axis=0), nn)
    assert b.shared_memory_bytes() > 0  # the forward layout is known before the first launch
    print("uniform step: 7 instances x 25 steps", flush=True)
    b.step(25)
    got = b.state_records()
    assert np.array_equal(got, np.repeat(gold.dense[35:36], 7, axis=0))
    with pytest.raises(EngineError):
        b.run()
    with pytest.raises(EngineError):
        b.restore(0, gold.dense[0, :7], *[gold.dense[0, 10 + k * nn:10 + (k + 1) * nn] for k in range(5)])
    with pytest.raises(ValueError):
        b.step(-1)
    b.close()
