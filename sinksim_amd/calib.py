"""Calibration parameters and the differentiable loss (no counterpart in the reference; they extend the reference's
kind scales, engine/include/sinksim/sweep.hpp:49-65, and its calibration objective, tools/py/calibrate_gpu.py:35-64).

Parameters are positive scale factors theta. Each connection c maps to at most one area factor pa[c] and one
discharge-coefficient factor pc[c] (-1: none); its area becomes area[c] * theta[pa[c]] and its coefficient
coef[c] * theta[pc[c]]. The flow of a connection is proportional to coef * area wherever the limiter does not act, so
the two factors of one kind move the flow alike; they are kept apart because the kind scales of a sweep act on the
area only.

The loss of a run is a weighted sum of squared errors against observations:

- readout terms: the trim (kind 0: pitch * 180 / pi) or list (kind 1: roll * 180 / pi) after a given step (the curve's
  trim and list, engine/include/sinksim/kernel/readouts.hpp); a step after the end of the run (the ship foundered
  first) scores the term's `after` value with no derivative, as the reference objective scores trimAfterFounder and
  listAfterFounder;
- event terms: the time at which a monitor's flow sum first exceeds the threshold (kind 0), a mark first goes under
  or above (kind 1), or the ship founders (kind 2). A step count has no derivative, so the time is the linear
  interpolation of the crossing between the step before and the step at which the event fires: with s_k the event's
  signed signal (monitor: sum - threshold; mark: -z, or +z for an 'above' mark; foundering: the largest margin of the
  founder rule -- hullTopBelow - hull top, |pitch| - pitchAbsAbove, after minT buoyancyBelow - V),
  t* = t_k - dt * s_k / (s_k - s_{k-1}). An event that fires at the first step scores t_1; one that never fires the
  founder time if the ship foundered and the term's `missing` value if it stayed afloat (the reference objective's
  afloatFounderMin), all with no derivative.

With the reference's observations (data/observations/titanic-1912.json) reference_objective() builds the reference's
objective J this way; it differs from the reference's evaluation of J only in taking event and founder times
interpolated within their step (less than dt = 0.25 s from the recorded times).
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field

import numpy as np

# The four calibration kinds of the reference's objective (tools/py/calibrate_gpu.py:25: breach, over, down, top).
CALIBRATION_KINDS = ("breach", "over", "down", "top")

# The observation weights of data/observations/titanic-1912.json (objective): trim sigma 0.3 deg up to 45 min and
# 0.4 deg up to 100 min; events well 130 +- 8 min, bridge 155 +- 5 min, overF 75 +- 8 min.
TITANIC_TRIM_MIN = (15.0, 30.0, 45.0, 75.0)
TITANIC_TRIM_SIGMA = (0.3, 0.3, 0.3, 0.4)
TITANIC_EVENTS = (("overF", 0, 8.0), ("well", 1, 8.0), ("bridge", 1, 5.0))  # (id, kind 0 monitor / 1 mark, sigma min)
TRIM_AFTER_FOUNDER = 40.0
AFLOAT_FOUNDER_S = 300.0 * 60.0


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
    scenario except 'open'). With titanic's nine kind names this gives 4 + 8 = 12 parameters; with coef_kinds=()
    the reference's four calibration scales."""
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


def zone_params(sim, kinds=None, zones: int = 3) -> ParamMap:
    """Area factors per (connection kind, longitudinal zone): the connections of each kind (default: every kind name
    but 'open') split into `zones` equal bands of their x position. Eight kinds in three zones give 24 parameters."""
    if kinds is None:
        kinds = [k for k in sim.kindNames if k != "open"]
    x = np.asarray(sim.cx, np.float64)
    edges = np.linspace(x.min(), x.max(), zones + 1)
    band = np.clip(np.searchsorted(edges, x, side="right") - 1, 0, zones - 1)
    names, pa = [], np.full(sim.connections, -1, np.int32)
    for k in kinds:
        b = sim.kind_index(k)
        if b < 0:
            raise ValueError(f"unknown connection kind {k!r}")
        for z in range(zones):
            pa[(sim.cKind == b) & (band == z)] = len(names)
            names.append(f"area:{k}:{z}")
    return ParamMap(names, pa, np.full(sim.connections, -1, np.int32))


@dataclass
class LossSpec:
    read_steps: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    read_kind: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))   # 0 trim, 1 list
    read_obs: np.ndarray = field(default_factory=lambda: np.zeros(0))
    read_w: np.ndarray = field(default_factory=lambda: np.zeros(0))
    read_after: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ev_kind: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))     # 0 monitor, 1 mark, 2 founder
    ev_idx: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    ev_obs: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ev_w: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ev_missing: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def signals(self) -> int:
        return len(self.read_steps) + len(self.ev_kind)

    @property
    def observations(self) -> np.ndarray:
        return np.concatenate([np.asarray(self.read_obs, np.float64), np.asarray(self.ev_obs, np.float64)])

    def with_observations(self, obs) -> "LossSpec":
        """The same terms against observations obs [signals] (readout values, then event times)."""
        obs = np.asarray(obs, np.float64)
        n = len(self.read_steps)
        return LossSpec(self.read_steps.copy(), self.read_kind.copy(), obs[:n].copy(), self.read_w.copy(), self.read_after.copy(),
                        self.ev_kind.copy(), self.ev_idx.copy(), obs[n:].copy(), self.ev_w.copy(), self.ev_missing.copy())

    def evaluate(self, signals) -> float:
        """The loss from the signals of a run (readout values then event times)."""
        s = np.asarray(signals, np.float64)
        n = len(self.read_steps)
        return float(np.sum(self.read_w * (s[:n] - self.read_obs) ** 2) + np.sum(self.ev_w * (s[n:] - self.ev_obs) ** 2))


def _spec(steps, kinds, w, after, ev_kind, ev_idx, ev_w, ev_missing) -> LossSpec:
    n, m = len(steps), len(ev_kind)
    return LossSpec(np.asarray(steps, np.int64), np.asarray(kinds, np.int32), np.zeros(n), np.asarray(w, np.float64),
                    np.asarray(after, np.float64), np.asarray(ev_kind, np.int32), np.asarray(ev_idx, np.int32), np.zeros(m),
                    np.asarray(ev_w, np.float64), np.asarray(ev_missing, np.float64))


def _event_index(sim, name: str, kind: int) -> int:
    ids = list(sim.monitorId) if kind == 0 else list(sim.markId)
    if name not in ids:
        raise ValueError(f"the scenario has no {'monitor' if kind == 0 else 'mark'} {name!r}")
    return ids.index(name)


def titanic_loss(sim, trim_minutes=TITANIC_TRIM_MIN, trim_sigma=TITANIC_TRIM_SIGMA, events=TITANIC_EVENTS) -> LossSpec:
    """Trim at 15, 30, 45 and 75 min and the times of overF, well and bridge, weighted as the reference objective
    weights them (1 / sigma^2, times in seconds). The observations are zero; fill them with with_observations."""
    steps = [int(round(m * 60.0 / sim.dt)) for m in trim_minutes]
    ev_kind = [k for _, k, _ in events]
    ev_idx = [_event_index(sim, name, k) for name, k, _ in events]
    ev_w = [1.0 / (60.0 * s) ** 2 for _, _, s in events]
    return _spec(steps, [0] * len(steps), 1.0 / np.asarray(trim_sigma, np.float64) ** 2, [TRIM_AFTER_FOUNDER] * len(steps),
                 ev_kind, ev_idx, ev_w, [AFLOAT_FOUNDER_S] * len(events))


def generic_loss(sim, trim_steps, trim_sigma_deg: float = 0.3, event_sigma_s: float = 300.0) -> LossSpec:
    """Trim at the given steps and the times of every monitor and mark of the scenario, all equally weighted."""
    kinds = [0] * sim.monitors + [1] * sim.marks
    idx = list(range(sim.monitors)) + list(range(sim.marks))
    n = len(trim_steps)
    return _spec(trim_steps, [0] * n, [1.0 / trim_sigma_deg ** 2] * n, [TRIM_AFTER_FOUNDER] * n, kinds, idx,
                 [1.0 / event_sigma_s ** 2] * len(kinds), [AFLOAT_FOUNDER_S] * len(kinds))


def load_observations(path: str | None = None) -> dict:
    """The reference's observation file (data/observations/titanic-1912.json, shipped with the package data)."""
    if path is None:
        from .data_paths import OBSERVATIONS
        path = OBSERVATIONS
    with open(path, "r", encoding="utf-8") as f:
        return json.load(f)


def reference_objective(sim, obs: dict | None = None) -> LossSpec:
    """The reference's calibration objective J (tools/py/calibrate_gpu.py:35-64 with the weights of the observation
    file): trim at each observed minute but 0 ((v - deg) / sigma by band)^2, trimAfterFounder after foundering; the
    founder time ((tf - founderMin) / founderSigmaMin)^2, afloatFounderMin when afloat; listWeight * ((v - deg) /
    listSigmaDeg)^2 at each observed list minute, listAfterFounder after foundering; each objective event
    ((t - obsMin) / sigmaMin)^2, the founder time when it never fired. Times in seconds (the weights carry the 60^2)."""
    obs = obs or load_observations()
    ob = obs["objective"]
    steps, kinds, w, after, rd_obs = [], [], [], [], []
    for p in obs["trim"]:
        m = p["min"]
        if m == 0:
            continue
        sigma = next(b["sigma"] for b in ob["trimSigmaByMinute"] if m <= b["upTo"])
        steps.append(int(round(m * 60.0 / sim.dt))); kinds.append(0); w.append(1.0 / sigma ** 2)
        after.append(float(ob["trimAfterFounder"])); rd_obs.append(float(p["deg"]))
    for p in obs["list"]:
        m = p["min"]
        steps.append(int(round(m * 60.0 / sim.dt))); kinds.append(1); w.append(ob["listWeight"] / ob["listSigmaDeg"] ** 2)
        after.append(float(ob["listAfterFounder"])); rd_obs.append(float(p["deg"]))
    ev_kind, ev_idx, ev_w, ev_missing, ev_obs = [2], [0], [1.0 / (60.0 * ob["founderSigmaMin"]) ** 2], [60.0 * ob["afloatFounderMin"]], [
        60.0 * obs["founderMin"]]
    for e in ob["events"]:
        kind = 0 if e["id"] in list(sim.monitorId) else 1
        ev_kind.append(kind); ev_idx.append(_event_index(sim, e["id"], kind)); ev_w.append(1.0 / (60.0 * e["sigmaMin"]) ** 2)
        ev_missing.append(60.0 * ob["afloatFounderMin"]); ev_obs.append(60.0 * e["obsMin"])
    spec = _spec(steps, kinds, w, after, ev_kind, ev_idx, ev_w, ev_missing)
    return spec.with_observations(np.r_[rd_obs, ev_obs])


def reference_J(curve: np.ndarray, events: list, foundered: bool, founderT: float, sim, obs: dict | None = None) -> float:
    """The reference's objective evaluated as the reference evaluates it (tools/py/calibrate_gpu.py:35-64) from a
    run's 30 s readout curve (records t trim list ...; linear interpolation in t) and event log [(t, id, kind)]."""
    obs = obs or load_observations()
    ob = obs["objective"]
    t = curve[:, 0]

    def interp(col, ts):
        return float(np.interp(ts, t, curve[:, col]))

    tf_min = founderT / 60 if foundered else ob["afloatFounderMin"]
    founder_sec = founderT if foundered else 1e9
    J = 0.0
    for p in obs["trim"]:
        m = p["min"]
        if m == 0:
            continue
        sigma = next(b["sigma"] for b in ob["trimSigmaByMinute"] if m <= b["upTo"])
        v = ob["trimAfterFounder"] if m * 60 > founder_sec else interp(1, m * 60)
        J += ((v - p["deg"]) / sigma) ** 2
    J += ((tf_min - obs["founderMin"]) / ob["founderSigmaMin"]) ** 2
    for p in obs["list"]:
        m = p["min"]
        v = ob["listAfterFounder"] if m * 60 > founder_sec else interp(2, m * 60)
        J += ob["listWeight"] * ((v - p["deg"]) / ob["listSigmaDeg"]) ** 2
    fired = {}
    for te, i, k in events:
        name = sim.event_id(k, i) if k in (0, 1) else None
        if name is not None and name not in fired:
            fired[name] = te / 60
    for e in ob["events"]:
        v = fired.get(e["id"], tf_min)
        J += ((v - e["obsMin"]) / e["sigmaMin"]) ** 2
    return float(J)
