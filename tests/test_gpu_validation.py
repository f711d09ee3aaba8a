"""The validation catalog on the GPU with the device's stopWhenStable (ss_batch_set_stop_when_stable; the rule of
engine/src/simulation.cpp:179-182, which the reference's GPU batch refuses, apps/sinksim_cuda_run.cpp:28):

- the 11 catalog cases as one fleet batch (one instance per scenario, curves every 30 s, 8 h cap) equal the oracle's
  runs with stopWhenStable in the same numerics, bit for bit: steps, verdict, founder and end times, every curve
  record and every event, in portable and portable32;
- the calm time survives chunked launches: the same batch in launches of 4,000 steps gives the same results;
- the catalog runner sinksim_hip_validate (the drop-in for apps/sinksim_validate.cpp) writes a table of the same
  format equal to the oracle's runs, and its comparison with the reference's validation table (oracle numerics) is the
  reference's comparison of those runs.
"""
import gzip
import json
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _fleet(numerics, steps_per_launch=0):
    from sinksim_amd.batch import Batch, BatchOptions
    from sinksim_amd.data_paths import catalog_cases, titanic
    cases = catalog_cases()
    models = [titanic(c["sim"]) for c in cases]
    opt = BatchOptions(count=len(cases), mixed_precision=numerics == "portable32", traced_instances=0, curve_every=30.0,
                       t_max=8 * 3600.0, steps_per_launch=steps_per_launch)
    b = Batch.fleet(models, np.arange(len(cases), dtype=np.int32), opt)
    b.set_stop_when_stable(True)
    b.run()
    out = []
    for i in range(len(cases)):
        r = b.result(i)
        out.append({"steps": r.steps, "foundered": r.foundered, "founderT": r.founderT, "tEnd": r.tEnd, "curve": b.curve(i),
                    "events": [(t, k, e) for t, k, e in r.events]})
    b.close()
    return cases, models, out


def _oracle(models, numerics):
    import pyoracle

    def one(m):
        ship, sim = m
        return pyoracle.sweep(ship, sim, numerics, sim.cEn.astype(np.uint8)[None], sim.cArea[None], sim.cTOn[None], 8 * 3600.0, 30.0,
                              threads=1, stop_when_stable=True)[0]
    with ThreadPoolExecutor(max_workers=8) as ex:
        return list(ex.map(one, models))


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_catalog_fleet_equals_oracle(numerics):
    cases, models, gpu = _fleet(numerics)
    cpu = _oracle(models, numerics)
    stopped = 0
    for c, g, o in zip(cases, gpu, cpu):
        print(f"{numerics} {c['id']}: {g['steps']} steps, foundered {g['foundered']}, end {g['tEnd']}", flush=True)
        assert g["steps"] == o["steps"] and g["foundered"] == o["foundered"] and g["tEnd"] == o["tEnd"], (c["id"], g["steps"], o["steps"])
        if o["foundered"]:
            assert g["founderT"] == o["founderT"], c["id"]
        elif o["tEnd"] < 8 * 3600.0:
            stopped += 1
        assert g["curve"].shape == o["curve"].reshape(-1, 7).shape, c["id"]
        assert np.array_equal(g["curve"], o["curve"].reshape(-1, 7)), c["id"]
        assert [(t, i, k) for t, i, k in g["events"]] == [(t, i, k) for t, i, k in o["events"]], c["id"]
    assert stopped >= 1


def test_calm_time_survives_chunked_launches():
    _, _, one = _fleet("portable")
    _, _, chunked = _fleet("portable", steps_per_launch=4000)
    for a, b in zip(one, chunked):
        assert a["steps"] == b["steps"] and a["tEnd"] == b["tEnd"] and a["foundered"] == b["foundered"]
        assert np.array_equal(a["curve"], b["curve"]) and a["events"] == b["events"]


def _expected_mismatches(table, cases, runs):
    """The reference's comparison (apps/sinksim_validate.cpp:76-103) of oracle runs with the table."""
    bad = []
    for c, o in zip(cases, runs):
        ref = table[c["id"]]
        final = o["curve"].reshape(-1, 7)[-1]
        p = []
        if ref["foundered"] != o["foundered"]:
            p.append("verdict")
        if ref["foundered"] and o["foundered"] and abs(ref["founderT"] - o["founderT"]) > 1.0:
            p.append("founderT")
        tol_t, tol_w = (0.5, 200) if ref["foundered"] else (0.01, 1)
        if abs(ref["trimDeg"] - final[1]) > tol_t:
            p.append("trim")
        if abs(ref["listDeg"] - final[2]) > tol_t:
            p.append("list")
        if abs(ref["waterT"] - final[3]) > tol_w:
            p.append("water")
        if not ref["foundered"] and abs(ref["endT"] - o["tEnd"]) > 60:
            p.append("endT")
        if p:
            bad.append(c["id"])
    return bad


@pytest.mark.parametrize("numerics", ["portable", "portable32"])
def test_validate_tool_against_the_reference_table(tmp_path, numerics):
    """The tool's table equals the oracle's runs in the same numerics exactly, and its verdict against the reference
    table (made in oracle numerics) is the reference's comparison of those runs: the same mismatching cases (where
    the portable numerics move an equilibrium's water by more than 1 t or a founder time by more than 1 s)."""
    from sinksim_amd.data_paths import CATALOG, catalog_cases, titanic
    ref = tmp_path / "validation.oracle.json"
    with gzip.open(os.path.join(ROOT, "tests", "golden", "validation.oracle.json.gz")) as f:
        ref.write_bytes(f.read())
    table = json.loads(ref.read_text())["cases"]
    cases = catalog_cases()
    runs = _oracle([titanic(c["sim"]) for c in cases], numerics)
    expected = _expected_mismatches(table, cases, runs)
    out = tmp_path / "validation.engine.json"
    exe = os.path.join(ROOT, "sinksim_amd", "bin", "sinksim_hip_validate")
    r = subprocess.run([exe, CATALOG, "--numerics", numerics, "--compare", str(ref), "--out", str(out)], capture_output=True, text=True,
                       timeout=600)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == (1 if expected else 0)
    assert f"RESULT: {'FAIL' if expected else 'PASS'} ({len(expected)} mismatching cases)" in r.stdout
    doc = json.loads(out.read_text())
    assert doc["format"] == "sinksim.validation" and len(doc["cases"]) == 11
    for c, o in zip(cases, runs):
        e = doc["cases"][c["id"]]
        final = o["curve"].reshape(-1, 7)[-1]
        assert e["foundered"] == o["foundered"] and e["endT"] == o["tEnd"], c["id"]
        assert (e["founderT"] == o["founderT"]) if o["foundered"] else e["founderT"] is None, c["id"]
        assert [e["trimDeg"], e["listDeg"], e["waterT"]] == [final[1], final[2], final[3]], c["id"]
        assert [x["tSec"] for x in e["events"]] == [t for t, _, _ in o["events"]], c["id"]
