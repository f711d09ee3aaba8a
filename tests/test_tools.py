"""The C++ command-line drop-ins on the host (no GPU): usage errors and exit codes, numerics refusals, the exact results
comparison, and sinksim_sweep's CPU runs against the reference's own sinksim_sweep (built from its sources in
oracle/_ref) in all four numerics, compared value for value by both tools' --compare."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
from conftest import ROOT
from sinksim_amd import sweep as S
from sinksim_amd.data_paths import SHIP, sim_path

BIN = os.path.join(ROOT, "sinksim_amd", "bin")
REF_SWEEP = os.path.join(pyoracle.REF_DIR, "sinksim_sweep")


def _tool(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.fail(f"{p} is not built (run build())")
    return p


def _run(*cmd):
    return subprocess.run(list(cmd), capture_output=True, text=True)


def test_usage_errors_exit_2():
    assert _run(_tool("sinksim_hip_run")).returncode == 2
    assert _run(_tool("sinksim_sweep")).returncode == 2
    assert _run(_tool("sinksim_sweep"), "--compare").returncode == 2
    r = _run(_tool("sinksim_sweep"), SHIP, sim_path("titanic"), "x.json", "--numerics", "quad")
    assert r.returncode == 2 and "unknown --numerics" in r.stderr
    r = _run(_tool("sinksim_sweep"), SHIP, sim_path("titanic"), "x.json", "--numerics", "oracle", "--device", "gpu")
    assert r.returncode == 2 and "CPU" in r.stderr
    r = _run(_tool("sinksim_sweep"), SHIP, sim_path("titanic"), "x.json", "--threads", "many")
    assert r.returncode == 2 and "needs a number" in r.stderr
    r = _run(_tool("sinksim_hip_run"), SHIP, sim_path("titanic"), "--precision", "quad")
    assert r.returncode == 2 and "unknown --precision" in r.stderr
    r = _run(_tool("sinksim_hip_run"), SHIP, sim_path("titanic"), "--numerics", "oracle")
    assert r.returncode == 2 and "CPU only" in r.stderr
    r = _run(_tool("sinksim_hip_run"), SHIP, sim_path("titanic"), "--count")
    assert r.returncode == 2 and "needs a value" in r.stderr


def test_missing_files_exit_1(tmp_path):
    r = _run(_tool("sinksim_sweep"), os.path.join(tmp_path, "none.json"), sim_path("titanic"), "x.json")
    assert r.returncode == 1 and r.stderr.startswith("sinksim_sweep: ")
    r = _run(_tool("sinksim_sweep"), "--compare", os.path.join(tmp_path, "a.json"), os.path.join(tmp_path, "b.json"))
    assert r.returncode == 1


def _results(titanic, count=3):
    """Oracle results of a small sweep, as a Results document."""
    ship, sim = titanic
    from sinksim_amd.workloads import titanic_batch_sweep
    sw = titanic_batch_sweep(sim, count=count, t_max=600.0)
    en, area, tOn = S.connection_tables(sw, sim)
    out = pyoracle.sweep(ship, sim, "portable32", en, area, tOn, sw.tMax, sw.every, threads=4)
    r = S.Results(ship.id, sim.id, "portable32", "oracle", sw.tMax, sw.every)
    for inst, o in zip(sw.instances, out):
        ev = [(t, sim.event_id(k, i), sim.event_label(k, i)) for t, i, k in o["events"]]
        r.instances.append(S.InstanceResult(inst.id, o["foundered"], o["founderT"], o["tEnd"], o["steps"],
                                            S.final_dict(o["curve"][-1]), ev, o["curve"]))
    return r


def test_compare_identical_and_different(titanic, tmp_path):
    r = _results(titanic)
    a, b = os.path.join(tmp_path, "a.json"), os.path.join(tmp_path, "b.json")
    S.save_results(r, a)
    S.save_results(r, b)
    res = _run(_tool("sinksim_sweep"), "--compare", a, b)
    assert res.returncode == 0 and "RESULT: IDENTICAL" in res.stdout
    r.instances[1].curve = r.instances[1].curve.copy()
    r.instances[1].curve[3, 1] = np.nextafter(r.instances[1].curve[3, 1], 1e9)
    S.save_results(r, b)
    res = _run(_tool("sinksim_sweep"), "--compare", a, b)
    assert res.returncode == 1 and "RESULT: DIFFERENT" in res.stdout and "curve[3] trim" in res.stdout


@pytest.mark.skipif(not os.path.exists(REF_SWEEP), reason="reference CPU build absent (make -C oracle ref)")
def test_reference_compare_reads_our_results(titanic, tmp_path):
    """The reference's own --compare parses results files written here and agrees on identity and difference."""
    r = _results(titanic, 2)
    a, b = os.path.join(tmp_path, "a.json"), os.path.join(tmp_path, "b.json")
    S.save_results(r, a)
    S.save_results(r, b)
    assert _run(REF_SWEEP, "--compare", a, b).returncode == 0
    r.instances[0].founderT += 0.25
    S.save_results(r, b)
    res = _run(REF_SWEEP, "--compare", a, b)
    assert res.returncode == 1 and "founderT" in res.stdout


@pytest.mark.skipif(not os.path.exists(REF_SWEEP), reason="reference CPU build absent (make -C oracle ref)")
@pytest.mark.parametrize("numerics", ["portable32", "portable", "oracle"])
def test_oracle_sweep_equals_reference_sweep(titanic, tmp_path, numerics):
    """The oracle's sweep runner against the reference's CPU sweep runner on the reference's smoke sweep, value for
    value (the sweep file, apply order, curves, events and final readouts)."""
    from conftest import GOLDEN
    ship, sim = titanic
    sw_path = os.path.join(GOLDEN, "smoke.sweep.json")
    ref_out = os.path.join(tmp_path, "ref.json")
    res = _run(REF_SWEEP, SHIP, sim_path("titanic"), sw_path, "--numerics", numerics, "--threads", "8", "--results", ref_out)
    assert res.returncode == 0, res.stderr
    ref = S.load_results(ref_out)
    sw = S.load_sweep(sw_path, sim)
    en, area, tOn = S.connection_tables(sw, sim)
    out = pyoracle.sweep(ship, sim, numerics, en, area, tOn, sw.tMax, sw.every, threads=8)
    mine = S.Results(ship.id, sim.id, numerics, "oracle", sw.tMax, sw.every)
    for inst, o in zip(sw.instances, out):
        ev = [(t, sim.event_id(k, i), sim.event_label(k, i)) for t, i, k in o["events"]]
        mine.instances.append(S.InstanceResult(inst.id, o["foundered"], o["founderT"], o["tEnd"], o["steps"],
                                               S.final_dict(o["curve"][-1]), ev, o["curve"]))
    d = S.compare(ref, mine)
    assert d.differing == 0, d.first
    assert d.compared > 1000


def test_validate_tool_usage_and_numerics():
    """sinksim_hip_validate without arguments exits 2; the CPU-only numerics are refused with exit 2."""
    assert _run(_tool("sinksim_hip_validate")).returncode == 2
    from sinksim_amd.data_paths import CATALOG
    assert _run(_tool("sinksim_hip_validate"), CATALOG, "--numerics", "oracle").returncode == 2


def _header_and_tail(text):
    lines = text.splitlines()
    return lines[0], [l for l in lines[2:] if l.startswith("  ")]


@pytest.mark.skipif(not os.path.exists(REF_SWEEP), reason="reference CPU build absent (make -C oracle ref)")
@pytest.mark.parametrize("numerics", ["oracle", "portable", "portable32", "std"])
def test_sinksim_sweep_is_the_reference_cpu_tool(tmp_path, numerics):
    """sinksim_sweep on the reference's smoke sweep on the host CPU: the same header and instance lines as the
    reference's sinksim_sweep, and results IDENTICAL by both tools' --compare (the sweep file, apply order, curves,
    events, final readouts and step counts; tests/CMakeLists.txt:36-39 runs this sweep as the reference's fixture)."""
    from conftest import GOLDEN
    sw = os.path.join(GOLDEN, "smoke.sweep.json")
    mine, ref = os.path.join(tmp_path, "mine.json"), os.path.join(tmp_path, "ref.json")
    a = _run(_tool("sinksim_sweep"), SHIP, sim_path("titanic"), sw, "--numerics", numerics, "--threads", "4", "--results", mine)
    assert a.returncode == 0, a.stderr
    b = _run(REF_SWEEP, SHIP, sim_path("titanic"), sw, "--numerics", numerics, "--threads", "4", "--results", ref)
    assert b.returncode == 0, b.stderr
    assert _header_and_tail(a.stdout) == _header_and_tail(b.stdout)
    for tool in (_tool("sinksim_sweep"), REF_SWEEP):
        c = _run(tool, "--compare", ref, mine)
        assert c.returncode == 0 and "RESULT: IDENTICAL" in c.stdout, c.stdout


@pytest.mark.skipif(not os.path.exists(REF_SWEEP), reason="reference CPU build absent (make -C oracle ref)")
def test_sinksim_sweep_default_numerics_and_stop_when_stable(tmp_path, titanic):
    """Without --numerics both tools run the oracle numerics; a sweep with stopWhenStable (the calm rule, CPU-only in
    the reference) over variants of the britannicClosed scenario, which settles afloat and stops on the rule, ends every
    instance at the same step in both tools."""
    import json
    from sinksim_amd.data_paths import titanic as load
    ship, sim = load("britannicClosed")
    doc = {"format": "sinksim.sweep", "formatVersion": 1, "ship": ship.id, "sim": sim.id, "tMax": 4 * 3600.0, "every": 60.0,
           "stopWhenStable": True,
           "instances": [{"id": "as-compiled"}, {"id": "smaller-breach", "scale": {"breach": 0.7}},
                         {"id": "larger-breach", "scale": {"breach": 1.3}}]}
    sw = os.path.join(tmp_path, "calm.sweep.json")
    with open(sw, "w") as f:
        json.dump(doc, f)
    mine, ref = os.path.join(tmp_path, "mine.json"), os.path.join(tmp_path, "ref.json")
    a = _run(_tool("sinksim_sweep"), SHIP, sim_path("britannicClosed"), sw, "--results", mine)
    b = _run(REF_SWEEP, SHIP, sim_path("britannicClosed"), sw, "--results", ref)
    assert a.returncode == 0 and b.returncode == 0, a.stderr + b.stderr
    assert "numerics oracle" in a.stdout.splitlines()[0]
    c = _run(REF_SWEEP, "--compare", ref, mine)
    assert c.returncode == 0, c.stdout
    r = S.load_results(mine)
    assert any(not i.foundered and i.tEnd < 4 * 3600.0 for i in r.instances), "no instance stopped on the calm rule"
