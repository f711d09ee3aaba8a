"""The oracle against the reference's golden traces of titanic64/titanic, at tolerance 0, in the three numerics the
reference pins (tests/CMakeLists.txt:13-28), with the reference's acceptance procedure (apps/sinksim_check.cpp:
105-169): the initial record, every dense record restarted and stepped once, every sampled window restarted and
stepped to the next sample, and the run end to end (samples, events, founder time, step count)."""
import numpy as np
import pytest

import pyoracle


@pytest.mark.parametrize("numerics", ["portable", "portable32", "oracle"])
def test_oracle_reproduces_golden(titanic, goldens, numerics):
    ship, sim = titanic
    gold = goldens[numerics]
    S = pyoracle.OracleSim(ship, sim, numerics)
    assert np.array_equal(S.record(), gold.dense[0]), "initial state"
    for i in range(gold.dense.shape[0] - 1):
        S.restore(gold.dense[i])
        S.step()
        assert np.array_equal(S.record(), gold.dense[i + 1]), f"dense step {i + 1}"
    for i in range(gold.samples.shape[0] - 1):
        S.restore(gold.samples[i])
        S.step(int(round((gold.samples[i + 1, 0] - gold.samples[i, 0]) / gold.dt)))
        assert np.array_equal(S.record(), gold.samples[i + 1]), f"window {i} -> {i + 1}"
    S.reset()
    dense, samples, steps = S.record_trace(5 * 3600, gold.denseSteps, gold.sampleEvery)
    assert steps == gold.steps
    assert np.array_equal(dense, gold.dense)
    assert np.array_equal(samples, gold.samples)
    events, foundered, founderT = S.events()
    assert foundered == gold.foundered and founderT == gold.founderT
    assert [(t, sim.event_id(k, i)) for t, i, k in events] == [(t, i) for t, i, _ in gold.events]


def test_mass_balance(titanic):
    """Water aboard equals the integral of the sea inflow (oracle/js/HANDOFF.md:30), step by step, to rounding."""
    ship, sim = titanic
    S = pyoracle.OracleSim(ship, sim, "portable")
    water0 = S.record()[10:10 + ship.nodes].sum()
    inflow = 0.0
    for _ in range(4000):
        S.step()
        inflow += S.outputs()[0] * sim.dt
    water = S.record()[10:10 + ship.nodes].sum()
    assert abs((water - water0) - inflow) <= 1e-9 * max(1.0, inflow)
    assert inflow > 1000.0


def test_frame_is_orthonormal(titanic):
    """R = Ry(pitch) Rx(roll) is a rotation; with g0 = 0 the translation is pure heave (test_kernel.cpp:143-160)."""
    body = np.zeros(16)
    F = pyoracle.pose(body, 3.5, 0.1, -0.2)
    R = F[:9].reshape(3, 3)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-15, rtol=0)
    assert F[11] == 3.5
    Q = pyoracle.pose(body, 3.5, 0.1, -0.2, "std")
    assert abs(F[8] - Q[8]) <= 1e-15
