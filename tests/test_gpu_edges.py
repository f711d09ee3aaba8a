"""Edge cases of the batch kernel against the oracle, in both numerics (each instance run with the oracle's sweep
semantics, curves every 30 s, compared value for value):

- the box-barge breach rig of the reference's kernel tests (tests/test_kernel.cpp:163-287): two spaces solved singly,
  two connections (a breach from the sea and an orifice between the spaces), no monitors, no marks;
- a run whose time cap is reached before the first step (no step, the closing record only);
- titanic with every connection disabled (no water moves; the ship stays at its initial state);
- a batch of one instance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _compare(ship, sim, numerics, en, area, tOn, t_max, every=30.0):
    import pyoracle
    from sinksim_amd.batch import Batch, BatchOptions
    n = en.shape[0]
    b = Batch(ship, sim, BatchOptions(count=n, mixed_precision=numerics == "portable32", traced_instances=0, curve_every=every,
                                      t_max=t_max))
    b.set_connections(en, area, tOn)
    b.run()
    cpu = pyoracle.sweep(ship, sim, numerics, en, area, tOn, t_max, every, threads=1)
    for i in range(n):
        r, c = b.result(i), cpu[i]
        assert r.steps == c["steps"] and r.foundered == c["foundered"], (i, r.steps, c["steps"])
        assert r.founderT == c["founderT"] and r.tEnd == c["tEnd"] and r.events == c["events"], i
        g = b.curve(i)
        assert g.shape == c["curve"].shape and np.array_equal(g, c["curve"]), i
    res = [b.result(i) for i in range(n)]
    b.close()
    return res, cpu


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_box_barge_breach_rig(numerics):
    from sinksim_amd.synth import box_barge
    ship, sim = box_barge(connections=[(-1, 1, 0, 40.0, -10.0, 1.0, 0.5, 0.6), (1, 0, 0, 0.0, 0.0, 0.0, 1.0, 0.6)])
    assert sim.monitors == 0 and sim.marks == 0 and sim.connections == 2
    en = np.tile(sim.cEn.astype(np.uint8), (3, 1))
    area = np.tile(sim.cArea.astype(np.float64), (3, 1)) * np.array([[1.0], [0.5], [2.0]])
    tOn = np.tile(sim.cTOn.astype(np.float64), (3, 1))
    res, cpu = _compare(ship, sim, numerics, en, area, tOn, t_max=1200.0)
    assert all(r.steps == 4800 for r in res)
    vol = res[0].vol
    assert vol[0] > 0 and vol[1] > vol[0]  # the breach fills space 1, the orifice passes water on to space 0


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_time_cap_before_the_first_step(titanic, numerics):
    ship, sim = titanic
    en, area, tOn = sim.cEn.astype(np.uint8)[None], sim.cArea.astype(np.float64)[None], sim.cTOn.astype(np.float64)[None]
    res, cpu = _compare(ship, sim, numerics, en, area, tOn, t_max=sim.state0[0])
    assert res[0].steps == 0 and not res[0].foundered


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_every_connection_disabled(titanic, numerics):
    ship, sim = titanic
    n = 2
    en = np.zeros((n, sim.connections), np.uint8)
    area = np.tile(sim.cArea.astype(np.float64), (n, 1))
    tOn = np.tile(sim.cTOn.astype(np.float64), (n, 1))
    res, cpu = _compare(ship, sim, numerics, en, area, tOn, t_max=600.0)
    for r in res:
        assert r.steps == 2400 and not r.foundered and r.events == []
        assert np.array_equal(r.vol, np.asarray(sim.vol0, np.float64))


def test_batch_of_one(titanic):
    ship, sim = titanic
    en, area, tOn = sim.cEn.astype(np.uint8)[None], sim.cArea.astype(np.float64)[None], sim.cTOn.astype(np.float64)[None]
    res, cpu = _compare(ship, sim, "portable", en, area, tOn, t_max=1800.0)
    assert res[0].steps == 7200
