"""The algorithmic FP64 operation count per simulation step that bench.py prices its roofline with, re-counted by the
oracle's operation counter on a sample of the bench workload (the 1,024-instance titanic64 sweep)."""
import importlib.util
import os

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


def test_flops_per_step_matches_counted_sample():
    ship, sim = titanic()
    sw = titanic_batch_sweep(sim, 1024)
    en, area, tOn = S.connection_tables(sw, sim)
    c = pyoracle.op_counts(ship, sim, "portable", en[:3], area[:3], tOn[:3], sw.tMax, sw.every)
    assert c["steps"] > 0 and c["passes"] > c["steps"]  # at least one level solve per step
    f = pyoracle.flops_per_step(c, ship.nodes, sim.connections, int((sim.ca < 0).sum()))
    assert abs(f / _bench().FLOPS_PER_STEP - 1) < 0.06, f
