"""The 1,024-instance titanic64 batch (the benchmark's workload) against the reference CPU engine's results for
every instance (tests/golden/titanic_batch1024.npz, written by tests/golden/make_batch1024.py from the reference's
own sweep runner).

- CPU: the oracle reproduces the fixture on a few instances (both numerics).
- GPU: the HIP batch reproduces every instance of the fixture exactly (steps, founder flag and time, end time,
  events, every curve record through its digest, final readouts), in the portable and portable32 numerics.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

FIXTURE = os.path.join(GOLDEN, "titanic_batch1024.npz")

pytestmark = pytest.mark.skipif(not os.path.exists(FIXTURE), reason="fixture not generated")


def _fixture():
    return np.load(FIXTURE, allow_pickle=False)


def _sweep():
    from sinksim_amd import sweep as S
    from sinksim_amd.data_paths import titanic
    from sinksim_amd.workloads import titanic_batch_sweep
    ship, sim = titanic("titanic")
    sw = titanic_batch_sweep(sim, count=1024)
    return ship, sim, sw, S.connection_tables(sw, sim)


def _check_instance(fx, numerics, i, steps, foundered, founderT, tEnd, events, curve, final, sim):
    p = numerics + "_"
    where = f"{numerics} instance {i}"
    assert steps == fx[p + "steps"][i], where
    assert bool(foundered) == bool(fx[p + "foundered"][i]), where
    assert founderT == fx[p + "founderT"][i] and tEnd == fx[p + "tEnd"][i], where
    start = int(fx[p + "evCount"][:i].sum())
    n = int(fx[p + "evCount"][i])
    want = [(float(t), str(fx["evNames"][k])) for t, k in zip(fx[p + "evT"][start:start + n], fx[p + "evName"][start:start + n])]
    assert [(t, sim.event_id(k, e)) for t, e, k in events] == want, where
    assert curve.shape[0] == fx[p + "curveLen"][i], where
    assert _digest(curve) == fx[p + "curveDigest"][i], where
    assert np.array_equal(final, fx[p + "final"][i]), where


def _digest(curve):
    import hashlib
    a = np.where(curve == 0.0, 0.0, np.asarray(curve, dtype=np.float64)).astype("<f8")
    return hashlib.blake2b(np.ascontiguousarray(a).tobytes(), digest_size=8).hexdigest()


def test_fixture_is_consistent():
    fx = _fixture()
    for numerics in ("portable", "portable32"):
        p = numerics + "_"
        assert fx[p + "steps"].shape == (1024,)
        assert int(fx[p + "evCount"].sum()) == fx[p + "evT"].shape[0]
        # every instance ends at its founder time or at the 4 h cap, after steps * dt of simulated time
        f = fx[p + "foundered"].astype(bool)
        assert np.all(fx[p + "tEnd"][f] == fx[p + "founderT"][f])
        assert np.all(fx[p + "tEnd"][~f] >= 4 * 3600.0)
        assert np.allclose(fx[p + "tEnd"], fx[p + "steps"] * 0.25)


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_oracle_matches_fixture_on_sample(numerics):
    import pyoracle
    fx = _fixture()
    ship, sim, sw, (en, area, tOn) = _sweep()
    # the two shortest instances (a few seconds of oracle time)
    order = np.argsort(fx[numerics + "_steps"])
    pick = [int(order[0]), int(order[1])]
    out = pyoracle.sweep(ship, sim, numerics, en[pick], area[pick], tOn[pick], sw.tMax, sw.every, threads=2)
    for i, o in zip(pick, out):
        _check_instance(fx, numerics, i, o["steps"], o["foundered"], o["founderT"], o["tEnd"], o["events"], o["curve"],
                        o["curve"][-1], sim)


@pytest.mark.gpu
@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_gpu_batch_matches_reference_results(numerics):
    from sinksim_amd.batch import Batch, BatchOptions
    fx = _fixture()
    ship, sim, sw, (en, area, tOn) = _sweep()
    b = Batch(ship, sim, BatchOptions(count=1024, mixed_precision=numerics == "portable32", traced_instances=0,
                                      curve_every=sw.every, t_max=sw.tMax))
    b.set_connections(en, area, tOn)
    b.run()
    res = b.results()
    p = numerics + "_"
    bad = np.flatnonzero(res["steps"] != fx[p + "steps"])
    assert bad.size == 0, f"{bad.size} instances differ in steps, first {bad[0]}: {res['steps'][bad[0]]} vs {fx[p + 'steps'][bad[0]]}"
    assert b.total_steps() == int(fx[p + "steps"].sum())
    for i in range(1024):
        r = b.result(i)
        _check_instance(fx, numerics, i, r.steps, r.foundered, r.founderT, r.tEnd, r.events, b.curve(i), b.final_readouts(i), sim)
    b.close()
