"""A run in epochs (every instance resident) against fixed launch shapes, bit for bit.

ss_batch_run splits a resident run into epochs: before each one the device lists the instances that have not ended,
so that the live ones spread evenly over the compute units, and the automatic shape runs them at 512 threads per
simulation once they all fit resident at that size and at 1,024 once they fit one per compute unit. None of this may
change a result. The batch here holds more instances than fit resident at 512 threads, so the automatic run starts at
256 threads and switches part way (to 512, then to 1,024 as instances end); it is compared with runs held at 256 and
at 512 threads (the latter in the refilling form while the instances exceed the resident count) and with short
epochs, in both numerics: steps, founder flags and times, end times, events, every curve record, the final readouts
and two traced instances' dense and sampled records. The calm-rule batch (96 instances) runs at 1,024 threads
throughout.
"""
import numpy as np
import pytest

COUNT = 640


def _run(numerics, threads=0, steps_per_launch=0, stop_when_stable=False, count=COUNT):
    from sinksim_amd import sweep as S
    from sinksim_amd.batch import Batch, BatchOptions
    from sinksim_amd.data_paths import titanic
    from sinksim_amd.workloads import titanic_batch_sweep
    ship, sim = titanic("titanic")
    sw = titanic_batch_sweep(sim, count=count, t_max=4 * 3600.0)
    en, area, tOn = S.connection_tables(sw, sim)
    b = Batch(ship, sim, BatchOptions(count=count, mixed_precision=numerics == "portable32", traced_instances=2, dense_steps=40,
                                      sample_every=600.0, curve_every=sw.every, t_max=sw.tMax, steps_per_launch=steps_per_launch))
    b.set_connections(en, area, tOn)
    if threads:
        b.set_launch(threads, 0)
    if stop_when_stable:
        b.set_stop_when_stable(True)
    b.run()
    out = {
        "results": [b.result(i) for i in range(count)],
        "curves": [b.curve(i) for i in range(count)],
        "final": [b.final_readouts(i) for i in range(count)],
        "traces": [b.trace(i) for i in range(2)],
        "launches": b.kernel_time()[1],
    }
    b.close()
    return out


def _same(a, b, what):
    for i, (ra, rb) in enumerate(zip(a["results"], b["results"])):
        where = f"{what}: instance {i}"
        assert ra.steps == rb.steps and ra.foundered == rb.foundered, where
        assert ra.founderT == rb.founderT and ra.tEnd == rb.tEnd and ra.events == rb.events, where
        assert np.array_equal(ra.state, rb.state) and np.array_equal(ra.outputs, rb.outputs), where
        assert np.array_equal(ra.vol, rb.vol) and np.array_equal(ra.level, rb.level), where
        assert np.array_equal(a["curves"][i], b["curves"][i]), where
        assert np.array_equal(a["final"][i], b["final"][i]), where
    for i in range(2):
        for x, y in zip(a["traces"][i], b["traces"][i]):
            assert np.array_equal(x, y), f"{what}: trace {i}"


@pytest.mark.gpu
@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_epochs_and_tail_form_keep_results(numerics):
    auto = _run(numerics)
    steps = np.array([r.steps for r in auto["results"]])
    # the instances end at different steps (so the live set shrinks), and some run to the 4 h cap
    assert steps.min() < 0.5 * steps.max() and (steps == steps.max()).sum() >= 2
    assert auto["launches"] >= 8
    print(f"{numerics}: automatic run in {auto['launches']} epochs, steps {steps.min()}..{steps.max()}", flush=True)
    _same(auto, _run(numerics, threads=256), f"{numerics} automatic vs 256 threads")
    _same(auto, _run(numerics, threads=512), f"{numerics} automatic vs 512 threads")
    _same(auto, _run(numerics, steps_per_launch=700), f"{numerics} automatic vs epochs of 700 steps")


@pytest.mark.gpu
def test_epochs_with_the_calm_rule():
    """stopWhenStable ends instances by the calm rule; the epoch plan's live test includes it."""
    auto = _run("portable", stop_when_stable=True, count=96)
    held = _run("portable", threads=256, stop_when_stable=True, count=96)
    _same(auto, held, "calm rule: automatic vs 256 threads")
