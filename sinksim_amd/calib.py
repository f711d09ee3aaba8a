"""Calibration parameters and the differentiable loss (no counterpart in the reference; they extend the reference's
kind scales, engine/include/sinksim/sweep.hpp:49-65, and its calibration objective, tools/py/calibrate_gpu.py:35-64).

Parameters are positive scale factors theta. Each connection c maps to at most one area factor pa[c] and one
discharge-coefficient factor pc[c] (-1: none); its area becomes area[c] * theta[pa[c]] and its coefficient
coef[c] * theta[pc[c]]. The flow of a connection is proportional to coef * area wherever the limiter does not act, so
the two factors of one kind move the flow alike; they are kept apart because the kind scales of a sweep act on the
area only.

The loss of a run is a weighted sum of squared errors:

- trim terms: trimDeg = pitch * 180 / pi after step k_j (the curve's trim, engine/include/sinksim/kernel/readouts.hpp),
  against an observation; a step after the end of the run (the ship foundered first) scores trim_after with no
  derivative, as the reference objective scores trimAfterFounder;
- event terms: the time at which a monitor's flow sum first exceeds the threshold, or a mark first goes under (or
  above), against an observation. A step count has no derivative, so the time is the linear interpolation of the
  crossing between the step before and the step at which the event fires: with s_k the event's signed signal
  (monitor: sum - threshold; mark: -z, or +z for an 'above' mark), t* = t_k - dt * s_k / (s_k - s_{k-1}). An event
  that fires at the first step scores t_1; one that never fires scores the run's end time, both with no derivative.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# The four calibration kinds of the reference's objective (tools/py/calibrate_gpu.py:30: breach, over, down, top).
CALIBRATION_KINDS = ("breach", "over", "down", "top")

# The observation weights of data/observations/titanic-1912.json (objective): trim sigma 0.3 deg up to 45 min and
# 0.4 deg up to 100 min; events well 130 +- 8 min, bridge 155 +- 5 min, overF 75 +- 8 min.
TITANIC_TRIM_MIN = (15.0, 30.0, 45.0, 75.0)
TITANIC_TRIM_SIGMA = (0.3, 0.3, 0.3, 0.4)
TITANIC_EVENTS = (("overF", 0, 8.0), ("well", 1, 8.0), ("bridge", 1, 5.0))  # (id, kind 0 monitor / 1 mark, sigma min)
TRIM_AFTER_FOUNDER = 40.0


@dataclass
class ParamMap:
    names: list
    pa: np.ndarray   # int32 [nconn]
    pc: np.ndarray   # int32 [nconn]

    @property
    def count(self) -> int:
        return len(self.names)


def kind_params(sim, area_kinds=CALIBRATION_KINDS, coef_kinds=None) -> ParamMap:
    """Area factors for area_kinds, then coefficient factors for coef_kinds (default: every kind name of the
    scenario except 'open'). With titanic's nine kind names this gives 4 + 8 = 12 parameters."""
    if coef_kinds is None:
        coef_kinds = [k for k in sim.kindNames if k != "open"]
    names, pa, pc = [], np.full(sim.connections, -1, np.int32), np.full(sim.connections, -1, np.int32)
    for k in area_kinds:
        b = sim.kind_index(k)
        if b < 0:
            raise ValueError(f"unknown connection kind {k!r}")
        pa[sim.cKind == b] = len(names)
        names.append("area:" + k)
    for k in coef_kinds:
        b = sim.kind_index(k)
        if b < 0:
            raise ValueError(f"unknown connection kind {k!r}")
        pc[sim.cKind == b] = len(names)
        names.append("coef:" + k)
    return ParamMap(names, pa, pc)


def group_params(sim, groups: int, seed: int = 0) -> ParamMap:
    """Area factors over `groups` fixed groups of connections (each connection to one group, by a seeded draw), for
    scenarios whose kinds are too few to give the wanted parameter count."""
    rng = np.random.default_rng(seed)
    g = rng.integers(0, groups, sim.connections).astype(np.int32)
    return ParamMap([f"area:g{k}" for k in range(groups)], g, np.full(sim.connections, -1, np.int32))


@dataclass
class LossSpec:
    trim_steps: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    trim_obs: np.ndarray = field(default_factory=lambda: np.zeros(0))
    trim_w: np.ndarray = field(default_factory=lambda: np.zeros(0))
    trim_after: float = TRIM_AFTER_FOUNDER
    ev_kind: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    ev_idx: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    ev_obs: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ev_w: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def signals(self) -> int:
        return len(self.trim_steps) + len(self.ev_kind)

    def with_observations(self, obs) -> "LossSpec":
        """The same terms against observations obs [signals] (trim values, then event times)."""
        obs = np.asarray(obs, np.float64)
        n = len(self.trim_steps)
        return LossSpec(self.trim_steps.copy(), obs[:n].copy(), self.trim_w.copy(), self.trim_after, self.ev_kind.copy(),
                        self.ev_idx.copy(), obs[n:].copy(), self.ev_w.copy())

    def evaluate(self, signals) -> float:
        """The loss from the signals of a run (trim values then event times)."""
        s = np.asarray(signals, np.float64)
        n = len(self.trim_steps)
        return float(np.sum(self.trim_w * (s[:n] - self.trim_obs) ** 2) + np.sum(self.ev_w * (s[n:] - self.ev_obs) ** 2))


def titanic_loss(sim, trim_minutes=TITANIC_TRIM_MIN, trim_sigma=TITANIC_TRIM_SIGMA, events=TITANIC_EVENTS) -> LossSpec:
    """Trim at 15, 30, 45 and 75 min and the times of overF, well and bridge, weighted as the reference objective
    weights them (1 / sigma^2, times in seconds). The observations are zero; fill them with with_observations."""
    steps = np.array([int(round(m * 60.0 / sim.dt)) for m in trim_minutes], np.int64)
    kinds, idx, w = [], [], []
    for name, kind, sigma in events:
        ids = list(sim.monitorId) if kind == 0 else list(sim.markId)
        if name not in ids:
            raise ValueError(f"the scenario has no {'monitor' if kind == 0 else 'mark'} {name!r}")
        kinds.append(kind)
        idx.append(ids.index(name))
        w.append(1.0 / (60.0 * sigma) ** 2)
    return LossSpec(steps, np.zeros(len(steps)), 1.0 / np.asarray(trim_sigma, np.float64) ** 2, TRIM_AFTER_FOUNDER,
                    np.array(kinds, np.int32), np.array(idx, np.int32), np.zeros(len(kinds)), np.array(w))


def generic_loss(sim, trim_steps, trim_sigma_deg: float = 0.3, event_sigma_s: float = 300.0) -> LossSpec:
    """Trim at the given steps and the times of every monitor and mark of the scenario, all equally weighted."""
    steps = np.asarray(trim_steps, np.int64)
    kinds = [0] * sim.monitors + [1] * sim.marks
    idx = list(range(sim.monitors)) + list(range(sim.marks))
    return LossSpec(steps, np.zeros(len(steps)), np.full(len(steps), 1.0 / trim_sigma_deg ** 2), TRIM_AFTER_FOUNDER,
                    np.array(kinds, np.int32), np.array(idx, np.int32), np.zeros(len(kinds)),
                    np.full(len(kinds), 1.0 / event_sigma_s ** 2))
