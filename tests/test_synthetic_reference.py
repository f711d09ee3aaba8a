"""Synthetic hulls pin the oracle where titanic64 does not reach: shared surfaces over N = 4 members (the capacity
cascade of kernel/levels.hpp:136-155), Threshold / Never / Single group modes, skipGroup cross-flow, late breach
activation, other node and column counts. The generator writes standard compiled files; the reference's own CPU
engine (oracle/_ref/sinksim_run, built from its sources) records a trace of each, and the oracle must reproduce
every record of it exactly, in the portable and portable32 numerics. Also: the generated files load back through
this package's loaders with their hashes verified."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
from sinksim_amd.model import check_pairing, load_ship, load_sim
from sinksim_amd.synth import synthetic_hull, write_compiled
from sinksim_amd.trace import load_trace

REF_RUN = os.path.join(pyoracle.REF_DIR, "sinksim_run")


def _files(tmp_path, seed, **kw):
    ship, sim = synthetic_hull(seed, **kw)
    sp, mp = os.path.join(tmp_path, f"s{seed}.ship.json"), os.path.join(tmp_path, f"s{seed}.sim.json")
    write_compiled(ship, sim, sp, mp)
    return sp, mp


def test_generated_files_load_and_pair(tmp_path):
    sp, mp = _files(tmp_path, 5, columns=2000, spaces=24)
    ship, sim = load_ship(sp), load_sim(mp)
    check_pairing(ship, sim)
    assert ship.nodes == 24 and 1500 <= ship.columns <= 2600
    assert int(sim.gNodeCount.sum()) == ship.nodes


@pytest.mark.skipif(not os.path.exists(REF_RUN), reason="reference CPU build absent (make -C oracle ref)")
@pytest.mark.parametrize("seed", [1, 4, 9, 12])
@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_oracle_reproduces_reference_on_synthetic_hulls(tmp_path, seed, numerics):
    sp, mp = _files(tmp_path, seed, columns=3000, spaces=40)
    out = os.path.join(tmp_path, "ref.trace.json")
    r = subprocess.run([REF_RUN, sp, mp, "--numerics", numerics, "--tmax", "2400", "--trace", out, "--dense", "200",
                        "--sample-every", "30"], capture_output=True, text=True)
    if numerics == "portable32" and r.returncode == 2:
        pytest.skip("this reference tool accepts oracle|portable|std only")
    assert r.returncode == 0, r.stderr
    ref = load_trace(out)
    ship, sim = load_ship(sp), load_sim(mp)
    S = pyoracle.OracleSim(ship, sim, numerics)
    dense, samples, steps = S.record_trace(2400.0, 200, 30.0)
    assert steps == ref.steps
    assert np.array_equal(dense, ref.dense)
    assert np.array_equal(samples, ref.samples)
    events, foundered, founderT = S.events()
    assert [(t, sim.event_id(k, i)) for t, i, k in events] == [(t, i) for t, i, _ in ref.events]
    assert foundered == ref.foundered and founderT == ref.founderT


@pytest.mark.skipif(not os.path.exists(REF_RUN), reason="reference CPU build absent (make -C oracle ref)")
def test_oracle_reproduces_reference_on_a_largest_hull(tmp_path):
    """A hull at the generator's largest sizes (16,000 columns, 128 spaces, 600 connections; seed 31, one of the GPU
    fleet test's), 1,200 s in portable numerics: every dense and sampled record of the reference's trace."""
    sp, mp = _files(tmp_path, 31, columns=16000, spaces=128, connections=600)
    out = os.path.join(tmp_path, "ref.trace.json")
    r = subprocess.run([REF_RUN, sp, mp, "--numerics", "portable", "--tmax", "1200", "--trace", out, "--dense", "100",
                        "--sample-every", "30"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    ref = load_trace(out)
    ship, sim = load_ship(sp), load_sim(mp)
    assert ship.columns >= 16000 and ship.nodes == 128 and sim.connections == 600
    S = pyoracle.OracleSim(ship, sim, "portable")
    dense, samples, steps = S.record_trace(1200.0, 100, 30.0)
    assert steps == ref.steps
    assert np.array_equal(dense, ref.dense)
    assert np.array_equal(samples, ref.samples)
