"""The product's host CPU engine (ss_cpu_trace / ss_cpu_sweep in libsinksim_hip.so; sinksim_amd/cpu.py) against the
reference's own outputs, no GPU:

- the reference's three golden traces of titanic64/titanic (oracle, portable and portable32 numerics) at tolerance 0:
  every dense record, every sampled record, the events, the founder time and the step count;
- the reference CPU engine's results for a sample of the 1,024-instance benchmark batch (the fixture
  tests/golden/titanic_batch1024.npz, written from oracle/_ref/sinksim_sweep) in portable and portable32;
- the std numerics (the platform's sin, cos and pow; no golden exists) against the reference's sinksim_sweep on the same
  host, in tests/test_tools.py.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN


@pytest.mark.parametrize("numerics", ["oracle", "portable", "portable32"])
def test_trace_reproduces_golden(titanic, goldens, numerics):
    from sinksim_amd import cpu
    ship, sim = titanic
    g = goldens[numerics]
    dense, samples, result, events = cpu.trace(ship, sim, numerics, t_max=5 * 3600.0, dense=g.denseSteps, sample_every=g.sampleEvery)
    assert dense.shape == g.dense.shape and samples.shape == g.samples.shape
    bad = np.argwhere(dense != g.dense)
    assert bad.size == 0, f"dense record {bad[0][0]} field {bad[0][1]} differs"
    bad = np.argwhere(samples != g.samples)
    assert bad.size == 0, f"sampled record {bad[0][0]} field {bad[0][1]} differs"
    assert result["foundered"] == g.foundered and result["founderT"] == g.founderT and result["steps"] == g.steps
    assert [(t, sim.event_id(k, i)) for t, i, k in events] == [(t, i) for t, i, _ in g.events]


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_sweep_reproduces_reference_batch_sample(numerics):
    """16 instances spread over the benchmark's 1,024 (every 64th), each with its own kind scales and door overrides,
    against the reference CPU engine's results: steps, founder, end time, final readouts, events and curve digests."""
    from sinksim_amd import cpu
    from test_batch1024 import FIXTURE, _check_instance, _fixture, _sweep
    if not os.path.exists(FIXTURE):
        pytest.skip("fixture not generated")
    fx = _fixture()
    ship, sim, sw, (en, area, tOn) = _sweep()
    pick = np.arange(0, 1024, 64)
    out = cpu.sweep(ship, sim, numerics, en[pick], area[pick], tOn[pick], t_max=sw.tMax, every=sw.every, threads=8)
    for j, i in enumerate(pick):
        _check_instance(fx, numerics, int(i), out["steps"][j], out["foundered"][j], out["founderT"][j], out["tEnd"][j], out["events"][j],
                        out["curves"][j], out["final7"][j], sim)


def test_refusals():
    from sinksim_amd import cpu
    from sinksim_amd.data_paths import titanic
    ship, sim = titanic()
    with pytest.raises(ValueError):
        cpu.sweep(ship, sim, "quad")
    with pytest.raises(ValueError):
        cpu.sweep(ship, sim, "oracle", count=1, every=0.0)
    with pytest.raises(ValueError):
        cpu.trace(ship, sim, "portable", t_max=-1.0)
    # the C ABI refuses what the wrapper would let through
    import ctypes as C
    from sinksim_amd import _capi as A
    lib = A.load_library()
    s, m = A.ShipStruct(ship), A.SimStruct(sim)
    opt = A.SsCpuOptions(7, 1, 600.0, 30.0, 0)
    assert lib.ss_cpu_sweep(C.byref(s.s), C.byref(m.s), C.byref(opt), 1, None, None, None, None, None, 0, None, 0, None, None,
                            None) == A.SS_INVALID_ARGUMENT
    assert b"numerics" in lib.ss_last_error()
    bad = A.SimStruct(sim)
    bad.s.nnode = 3
    opt.numerics = A.SS_NUMERICS_ORACLE
    assert lib.ss_cpu_sweep(C.byref(s.s), C.byref(bad.s), C.byref(opt), 1, None, None, None, None, None, 0, None, 0, None, None,
                            None) == A.SS_INVALID_ARGUMENT
