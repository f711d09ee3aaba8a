"""Seeded batch workloads.

titanic_batch_sweep: the 1,024-instance titanic64 batch the benchmark runs — per instance, a scale factor on the
area of every connection kind present in the scenario drawn log-uniform in [0.5, 2], and each watertight door
(kind "door", idx 2..13) overridden open with probability 0.2; tMax 4 h, curve every 30 s. It is written as a
standard sinksim.sweep document, so the CPU tools run the identical instances.

calibration_hull / calibration_problems: the gradient calibration workload on a synthetic hull of 16,000 columns,
128 spaces and 600 connections over 20,000 steps, with 24 kind-zone area factors as parameters; each problem is a
twin experiment with its own hidden scale factors.
"""
from __future__ import annotations

import math

import numpy as np

from . import sweep as S


def titanic_batch_sweep(sim, count: int = 1024, seed: int = 20261008, t_max: float = 4 * 3600.0, every: float = 30.0,
                        p_open: float = 0.2) -> S.Sweep:
    rng = np.random.default_rng(seed)
    present = sorted({int(k) for k in sim.cKind})
    door = sim.kind_index("door")
    door_idx = sorted({int(sim.cIdx[c]) for c in range(sim.connections) if int(sim.cKind[c]) == door}) if door >= 0 else []
    sw = S.Sweep(ship=sim.shipId, sim=sim.id, tMax=t_max, every=every)
    for i in range(count):
        scale = [(k, float(math.exp(rng.uniform(math.log(0.5), math.log(2.0))))) for k in present]
        opened = [S.ConnectionOverride(kind=door, idx=d, enabled=True) for d in door_idx if rng.uniform() < p_open]
        sw.instances.append(S.Instance(id=f"b{i:04d}", kindScale=scale, overrides=opened))
    return sw


def random_tables(sim, count: int, seed: int, activation_share: float = 0.0, p_open: float = 0.2):
    """[count, nconn] connection tables: kind scales log-uniform in [0.5, 2], every disabled connection enabled
    with probability p_open, and optionally a share of connections activated late (tOn uniform in [0, 1800] s)."""
    rng = np.random.default_rng(seed)
    nc = sim.connections
    en = np.tile(sim.cEn.astype(np.uint8), (count, 1))
    area = np.tile(sim.cArea.astype(np.float64), (count, 1))
    tOn = np.tile(sim.cTOn.astype(np.float64), (count, 1))
    kinds = sorted({int(k) for k in sim.cKind})
    for i in range(count):
        for k in kinds:
            m = sim.cKind == k
            area[i, m] = sim.cArea[m] * math.exp(rng.uniform(math.log(0.5), math.log(2.0)))
        opened = (sim.cEn == 0) & (rng.uniform(size=nc) < p_open)
        en[i, opened] = 1
        if activation_share > 0:
            late = rng.uniform(size=nc) < activation_share
            tOn[i, late] = rng.uniform(0, 1800, size=int(late.sum()))
    return en, area, tOn


def uneven_fleet(count: int = 4096, seed: int = 20261008, processes: int = 0):
    """The uneven batch: `count` seeded synthetic hulls (sinksim_amd.synth.synthetic_hull with seeds seed + i: length
    100-300 m, 1,000-16,000 columns, 8-128 spaces, 20-600 connections, 1-6 breaches), one instance of each running its
    own scenario from its initial state. Returns (models, inst_model). Hulls are built in `processes` worker
    processes (0: one per available core, at most 16)."""
    from .synth import synthetic_hull
    seeds = [seed + i for i in range(count)]
    if processes == 0:
        try:
            import os
            processes = min(16, len(os.sched_getaffinity(0)))
        except (AttributeError, OSError):
            processes = 1
    if processes > 1 and count > 64:
        from multiprocessing import get_context
        with get_context("fork").Pool(processes) as pool:
            models = pool.map(synthetic_hull, seeds, chunksize=32)
    else:
        models = [synthetic_hull(s) for s in seeds]
    return models, np.arange(count, dtype=np.int32)


def twin_experiment(count: int = 256, nparams: int = 12, seed: int = 20261008, span: float = 0.3):
    """The gradient workload's parameters: per instance hidden true scale factors and an independent starting guess,
    both log-uniform in [exp(-span), exp(span)] (the reference's search box, tools/py/calibrate_gpu.py:84,106),
    from one seed. Returns (theta_true, theta_guess), each [count, nparams]."""
    rng = np.random.default_rng(seed)
    truth = np.exp(rng.uniform(-span, span, (count, nparams)))
    guess = np.exp(rng.uniform(-span, span, (count, nparams)))
    return truth, guess


# the calibration hull: seed 5 of the synthetic generator at 16,000 columns, 128 spaces and 600 connections (16,680
# columns after the plan's rounding, 61 free-surface groups of up to four spaces, 31 monitors, 3 marks); its scenario
# stays afloat for the whole 20,000-step window (six events fire in it), so every run has the same length
CALIBRATION_HULL = dict(seed=5, columns=16000, spaces=128, connections=600)
CALIBRATION_STEPS = 20000
CALIBRATION_TRIM_STEPS = (4800, 9600, 14400, 19200)  # 20, 40, 60 and 80 minutes at dt = 0.25 s


def calibration_hull():
    """(ship, sim) of the calibration workload."""
    from .synth import synthetic_hull
    return synthetic_hull(**CALIBRATION_HULL)


def calibration_problems(count: int, seed: int = 20261008, span: float = 0.3, nparams: int = 24):
    """Hidden true log scale factors of `count` twin experiments, uniform in [-span, span] (the reference's search
    box around scales 1, tools/py/calibrate_gpu.py:84,106): [count, nparams]."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-span, span, (count, nparams))
