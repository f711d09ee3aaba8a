"""The algorithmic FP64 operation count per simulation step that bench.py prices its roofline with.

The oracle's operation counter is pinned first against counts made independently of it: a counting passer driven
through the reference's own kernel headers over the portable 1912 run (titanic64, scenario titanic, to founder)
measured 37,760 steps and, per step, 4,408 column evaluations of which 92.0 % wet, 71.1 node passes, 9,435 segment
evaluations of which 44.4 % wet and 29.9 % partly wet, 45.9 active connections and 27.3 weir pow calls; priced
per evaluation (18 / 7 per wet / dry column, 22 / 21 / 10 per partial / full / dry segment) these give 75.5k column
and 143.2k segment operations per step. The bench's constant is then re-counted on a sample of the bench workload
(the 1,024-instance titanic64 sweep)."""
import importlib.util
import os

import numpy as np

import pyoracle
from sinksim_amd import sweep as S
from sinksim_amd.data_paths import titanic
from sinksim_amd.workloads import titanic_batch_sweep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_counter_reproduces_the_independent_count_of_the_1912_run():
    ship, sim = titanic()
    en, area, tOn = sim.cEn.astype(np.uint8)[None], sim.cArea[None].copy(), sim.cTOn[None].copy()
    c = pyoracle.op_counts(ship, sim, "portable", en, area, tOn, 1e9, 30.0)
    s = c["steps"]
    assert s == 37760
    cols = (c["cwet"] + c["cdry"]) / s
    segs = (c["spart"] + c["sfull"] + c["sdry"]) / s
    assert round(cols) == 4408 and round(100 * c["cwet"] / (c["cwet"] + c["cdry"]), 1) == 92.0
    assert round(segs) == 9435
    assert round(100 * (c["spart"] + c["sfull"]) / (segs * s), 1) == 44.4
    assert round(100 * c["spart"] / (segs * s), 1) == 29.9
    assert round(c["passes"] / s, 1) == 71.1
    assert round(c["active"] / s, 1) == 45.9
    assert round(c["pows"] / s, 1) == 27.3
    assert round((18 * c["cwet"] + 7 * c["cdry"]) / s / 100) == 755
    assert round((22 * c["spart"] + 21 * c["sfull"] + 10 * c["sdry"]) / s / 100) == 1432


def test_flops_per_step_matches_counted_sample():
    ship, sim = titanic()
    sw = titanic_batch_sweep(sim, 1024)
    en, area, tOn = S.connection_tables(sw, sim)
    c = pyoracle.op_counts(ship, sim, "portable", en[:3], area[:3], tOn[:3], sw.tMax, sw.every)
    assert c["steps"] > 0 and c["passes"] > c["steps"]  # at least one level solve per step
    f = pyoracle.flops_per_step(c, ship.nodes, sim.connections, int((sim.ca < 0).sum()))
    assert abs(f / _bench().FLOPS_PER_STEP - 1) < 0.03, f
