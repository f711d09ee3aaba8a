"""Calibration of connection-kind scale factors on the GPU: the reference's batched search and its gradient-based
replacement, on one objective (no counterpart in the reference for the gradient method).

- search(): the reference's calibration (tools/py/calibrate_gpu.py:95-124): rounds of `batch` candidate points
  log-uniform in a box around the best point so far (the first round includes scales 1), the box shrinking each round,
  every point one instance of a GPU batch run (the plain forward kernel, curves every 30 s), scored with the
  reference's objective J (calib.reference_J) on its curve and events. The candidate points are the reference's own:
  the same Python random.Random(seed) draws in the same order.
- lbfgs(): batched L-BFGS over log scale factors from several starts at once, each iteration one call of the adjoint
  (calib.reference_objective, the same J with event and founder times interpolated within their step, and its
  gradient) through the torch.autograd.Function of sinksim_amd.adjoint, with a backtracking line search evaluated for
  every start in the same call.

Both report the device time of their kernel launches and the simulation steps they ran.

On many problems at once (twin experiments on the calibration hull: 24 kind-zone area factors, each problem its
own hidden factors and the observations they produce):

- twin_search(): the same search per problem (5 rounds x 264 points, the first round holding scales 1, the box
  halving around the problem's best), every round of every problem in one batch run of the plain forward kernel,
  each run scored from its 30 s curve and event log (twin_J);
- twin_lbfgs(): lbfgs_core with one start per problem at scales 1, every iteration one adjoint call over all problems;
- twin_polish(): after the gradient phase, rounds of forward-only runs at points drawn in a small box around each
  problem's best point (the box halving where a round finds nothing better). Below J ~ 1e-3 this objective is rugged
  (event times of weir onsets move by seconds as levels held within the level solve's tolerance catch up), so its
  lowest values are found by sampling, which is how the search reaches them too.

twin_compare() runs both and reports each method's device time per problem, the search on a subset of the problems
(its cost per problem does not depend on how many share a launch once a round fills the device).
"""
from __future__ import annotations

import math
import random
import time
from dataclasses import dataclass, field

import numpy as np

from .batch import Batch, BatchOptions
from .calib import CALIBRATION_KINDS, kind_params, load_observations, reference_J, reference_objective


@dataclass
class Trace:
    """Progress of a calibration: after each round / iteration, the best objective so far, the cumulative device time
    (ms), the cumulative simulation steps and evaluations, and the best point (log scales); row_best: each start's
    best objective so far [starts] (lbfgs_core)."""
    best_J: list = field(default_factory=list)
    device_ms: list = field(default_factory=list)
    steps: list = field(default_factory=list)
    evaluations: list = field(default_factory=list)
    best_x: list = field(default_factory=list)
    row_best: list = field(default_factory=list)
    final_x: object = None  # every start's best point [starts, P] (lbfgs_core, twin_polish)

    def add(self, J, ms, steps, evals, x, rows=None):
        self.best_J.append(float(J)); self.device_ms.append(float(ms)); self.steps.append(int(steps)); self.evaluations.append(int(evals))
        self.best_x.append(np.asarray(x, np.float64).tolist())
        if rows is not None:
            self.row_best.append(np.asarray(rows, np.float64).copy())

    def time_to(self, target: float) -> float | None:
        """Device time (ms) at which the best objective first reached `target` (None if never)."""
        for J, ms in zip(self.best_J, self.device_ms):
            if J <= target:
                return ms
        return None


def _kind_area(sim, kinds, x):
    """The connection areas with kind k's area scaled by exp(x[k]) (sweep::apply, engine/include/sinksim/sweep.hpp:49-65)."""
    area = np.asarray(sim.cArea, np.float64).copy()
    for k, v in zip(kinds, x):
        b = sim.kind_index(k)
        m = np.asarray(sim.cKind) == b
        area[m] = area[m] * math.exp(v)
    return area


def search(ship, sim, rounds: int = 5, batch: int = 264, span: float = 0.3, shrink: float = 0.5, seed: int = 20261008,
           kinds=CALIBRATION_KINDS, numerics: str = "portable", obs: dict | None = None, log=None) -> Trace:
    """The reference's batched search (tools/py/calibrate_gpu.py:95-124) on the GPU batch engine."""
    obs = obs or load_observations()
    ob = obs["objective"]
    t_max, every = ob["tMaxMin"] * 60.0, float(ob["historyEverySec"])
    rng = random.Random(seed)
    center = {k: 0.0 for k in kinds}
    best = (math.inf, None)
    trace = Trace()
    ms_total = steps_total = evals = 0
    en, tOn = np.asarray(sim.cEn, np.uint8), np.asarray(sim.cTOn, np.float64)
    for r in range(rounds):
        points = []
        if r == 0:
            points.append({k: 0.0 for k in kinds})
        while len(points) < batch:
            points.append({k: center[k] + rng.uniform(-span, span) for k in kinds})
        n = len(points)
        area = np.stack([_kind_area(sim, kinds, [p[k] for k in kinds]) for p in points])
        b = Batch(ship, sim, BatchOptions(count=n, mixed_precision=numerics == "portable32", traced_instances=0, curve_every=every,
                                          t_max=t_max))
        b.set_connections(np.tile(en, (n, 1)), area, np.tile(tOn, (n, 1)))
        b.run()
        ms, _ = b.kernel_time()
        ms_total += ms
        steps_total += b.total_steps()
        evals += n
        scored = []
        for i, p in enumerate(points):
            res = b.result(i)
            J = reference_J(b.curve(i), res.events, res.foundered, res.founderT, sim, obs)
            scored.append((J, p))
        b.close()
        scored.sort(key=lambda s: s[0])
        if scored[0][0] < best[0]:
            best = scored[0]
        center = dict(scored[0][1])
        span *= shrink
        trace.add(best[0], ms_total, steps_total, evals, [best[1][k] for k in kinds])
        if log:
            log(f"search round {r}: {n} points, round best J {scored[0][0]:.4f}, best {best[0]:.4f}, device {ms_total / 1e3:.2f} s")
    return trace


def lbfgs_core(evaluate, x0, iterations: int = 12, memory: int = 6, max_step: float = 0.3, lo: float = -1.6, hi: float = 1.6,
               target=None, progress=None, stop=None) -> Trace:
    """Batched L-BFGS: evaluate(x [starts, P]) -> (f [starts], g [starts, P]) evaluates every start at once.

    target: a number, or one per start (then the run stops once every start has reached its own). stop(trace), when
    given, is asked after every iteration and ends the run when it returns True.

    Each iteration evaluates, for every start, a trial point x + a d (d from the two-loop recursion over the start's
    last `memory` accepted steps, its largest entry at most max_step; the first direction steepest descent scaled to
    max_step) and keeps it where the Armijo condition holds; a start whose trial fails halves a and retries at the next
    call (after six halvings its memory is cleared). Points are clipped to [lo, hi]. Stops after `iterations` calls or
    once the best objective reaches `target`; progress(trace) is called after every evaluation and returns
    (device_ms, steps) for the trace."""
    trace = Trace()
    x = np.clip(np.atleast_2d(np.asarray(x0, np.float64)).copy(), lo, hi)
    S, P = x.shape
    f, g = evaluate(x)
    best_i = int(np.argmin(f))
    best = (float(f[best_i]), x[best_i].copy())
    rows = np.where(np.isfinite(f), f, np.inf)
    rows_x = x.copy()
    ms, st = progress() if progress else (0.0, 0)
    trace.add(best[0], ms, st, S, best[1], rows)
    s_hist = [[] for _ in range(S)]
    y_hist = [[] for _ in range(S)]
    alpha = np.ones(S)
    d = np.zeros_like(x)

    def direction(i):
        q = g[i].copy()
        a_s = []
        for s_, y_ in zip(reversed(s_hist[i]), reversed(y_hist[i])):
            rho = 1.0 / float(y_ @ s_)
            a = rho * float(s_ @ q)
            a_s.append((rho, a, s_, y_))
            q -= a * y_
        if s_hist[i]:
            s_, y_ = s_hist[i][-1], y_hist[i][-1]
            q *= float(s_ @ y_) / float(y_ @ y_)
        for rho, a, s_, y_ in reversed(a_s):
            q += s_ * (a - rho * float(y_ @ q))
        di = -q
        if not s_hist[i] or float(di @ g[i]) >= 0:  # first step, or not a descent direction: steepest descent
            di = -g[i] * (max_step / max(np.abs(g[i]).max(), 1e-300))
        n = np.abs(di).max()
        return di * (max_step / n) if n > max_step else di

    def reached():
        if target is None:
            return False
        if np.ndim(target) == 0:
            return best[0] <= float(target)
        return bool(np.all(rows <= np.asarray(target, np.float64)))

    for it in range(iterations):
        if reached():
            break
        for i in range(S):
            if alpha[i] == 1.0:
                d[i] = direction(i)
        xt = np.clip(x + alpha[:, None] * d, lo, hi)
        ft, gt = evaluate(xt)
        for i in range(S):
            step = xt[i] - x[i]
            if np.isfinite(ft[i]) and ft[i] <= f[i] + 1e-4 * float(g[i] @ step):
                y_ = gt[i] - g[i]
                if float(y_ @ step) > 1e-12:
                    s_hist[i].append(step); y_hist[i].append(y_)
                    if len(s_hist[i]) > memory:
                        s_hist[i].pop(0); y_hist[i].pop(0)
                x[i], f[i], g[i] = xt[i], ft[i], gt[i]
                alpha[i] = 1.0
            else:
                alpha[i] *= 0.5
                if alpha[i] < 1.0 / 64:
                    alpha[i] = 1.0
                    s_hist[i].clear(); y_hist[i].clear()
            if np.isfinite(ft[i]) and ft[i] < best[0]:
                best = (float(ft[i]), xt[i].copy())
            if np.isfinite(ft[i]) and ft[i] < rows[i]:
                rows[i] = ft[i]
                rows_x[i] = xt[i]
        ms, st = progress() if progress else (0.0, 0)
        trace.add(best[0], ms, st, S * (it + 2), best[1], rows)
        if stop is not None and stop(trace):
            break
    trace.final_x = rows_x
    return trace


def lbfgs(ship, sim, x0, iterations: int = 12, memory: int = 6, kinds=CALIBRATION_KINDS, max_step: float = 0.3, lo: float = -1.6,
          hi: float = 1.6, target: float | None = None, checkpoint_every: int = 0, obs: dict | None = None, log=None) -> Trace:
    """lbfgs_core over x = log scale factors of `kinds` on the reference objective, every evaluation one call of the
    adjoint (relaxed gradients) through its torch.autograd.Function."""
    import torch

    from .adjoint import GradientOptions, GradientRunner, flood_loss
    obs = obs or load_observations()
    t_max = obs["objective"]["tMaxMin"] * 60.0
    pm = kind_params(sim, area_kinds=kinds, coef_kinds=())
    spec = reference_objective(sim, obs)
    S = np.atleast_2d(x0).shape[0]
    runner = GradientRunner([(ship, sim)], np.zeros(S, np.int32), pm, spec, t_max,
                            GradientOptions(checkpoint_every=checkpoint_every, relaxed=True))

    def evaluate(x):
        xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
        loss = flood_loss(torch.exp(xt), runner)
        loss.sum().backward()
        return loss.detach().numpy().copy(), xt.grad.numpy().copy()

    def progress():
        if log:
            log(f"lbfgs: call {runner.calls}, device {runner.kernel_ms / 1e3:.2f} s")
        return runner.kernel_ms, runner.sim_steps

    try:
        tr = lbfgs_core(evaluate, x0, iterations=iterations, memory=memory, max_step=max_step, lo=lo, hi=hi, target=target, progress=progress)
    finally:
        runner.close()
    if log:
        log(f"lbfgs: best J {tr.best_J[-1]:.4f} after {runner.calls} calls, device {runner.kernel_ms / 1e3:.2f} s")
    return tr


def compare(ship, sim, starts: int = 256, seed: int = 20261008, iterations: int = 12, rounds: int = 5, batch: int = 264, log=None) -> dict:
    """The reference's search, then L-BFGS (from scales 1 and starts - 1 seeded points of the search's first box) until
    it reaches the search's best objective; device times and steps of both."""
    t0 = time.perf_counter()
    sw = search(ship, sim, rounds=rounds, batch=batch, seed=seed, log=log)
    rng = np.random.default_rng(seed)
    x0 = np.vstack([np.zeros(len(CALIBRATION_KINDS)), rng.uniform(-0.3, 0.3, (starts - 1, len(CALIBRATION_KINDS)))])
    lb = lbfgs(ship, sim, x0, iterations=iterations, target=sw.best_J[-1], log=log)
    reached = lb.time_to(sw.best_J[-1])
    return {"search": sw, "lbfgs": lb, "search_best_J": sw.best_J[-1], "lbfgs_best_J": lb.best_J[-1],
            "search_ms": sw.device_ms[-1], "lbfgs_ms_to_search_J": reached, "lbfgs_ms": lb.device_ms[-1],
            "wall_s": time.perf_counter() - t0}


# ------------------------------------------------------------------------------------------------ twin problems


def twin_J(curve: np.ndarray, events: list, foundered: bool, founderT: float, spec, obs_row, dt: float) -> float:
    """The loss of calib.LossSpec `spec` against observations obs_row, evaluated from a run's readout curve (records t
    trim list ..., a record at each readout step's time) and event log [(t, id, kind)]: a readout after the run's end
    takes its after-founder value; an event time is the time of the step it fired in (the differentiable loss
    interpolates within that step), an event that never fired the founder time, or its missing value if afloat."""
    n = len(spec.read_steps)
    obs_row = np.asarray(obs_row, np.float64)
    J = 0.0
    t = curve[:, 0]
    for j in range(n):
        tj = float(spec.read_steps[j]) * dt
        if foundered and tj > founderT:
            v = float(spec.read_after[j])
        else:
            v = float(np.interp(tj, t, curve[:, 1 + int(spec.read_kind[j])]))
        J += float(spec.read_w[j]) * (v - obs_row[j]) ** 2
    first = {}
    for te, i, k in events:
        first.setdefault((int(k), int(i)), float(te))
    for e in range(len(spec.ev_kind)):
        k, i = int(spec.ev_kind[e]), int(spec.ev_idx[e])
        if k == 2:
            v = founderT if foundered else float(spec.ev_missing[e])
        else:
            v = first.get((k, i), founderT if foundered else float(spec.ev_missing[e]))
        J += float(spec.ev_w[e]) * (v - obs_row[n + e]) ** 2
    return float(J)


@dataclass
class TwinProblems:
    """Twin experiments on one hull: parameter map, loss, hidden log factors [P, np] and their observations [P, signals]."""
    ship: object
    sim: object
    params: object
    spec: object
    truth: np.ndarray
    obs: np.ndarray
    steps: int

    @property
    def count(self) -> int:
        return int(self.truth.shape[0])


def _area_rows(sim, params, x):
    """Connection areas scaled by exp(x[pa]) per row of x [n, np] (area-only parameter maps)."""
    pa = np.asarray(params.pa)
    fac = np.where(pa >= 0, np.exp(np.asarray(x)[:, np.maximum(pa, 0)]), 1.0)
    return np.asarray(sim.cArea, np.float64)[None, :] * fac


def objective(tp: TwinProblems, x, rows=None, numerics: str = "portable") -> np.ndarray:
    """The differentiable loss (the adjoint's forward, no reverse sweep) of problems `rows` (default all) at log
    factors x [len(rows), np]."""
    from .adjoint import GradientOptions, GradientRunner
    rows = np.arange(tp.count) if rows is None else np.asarray(rows)
    run = GradientRunner([(tp.ship, tp.sim)], np.zeros(len(rows), np.int32), tp.params, tp.spec, tp.steps * tp.sim.dt,
                         GradientOptions(checkpoint_every=1 << 16, relaxed=True, max_steps=tp.steps, numerics=numerics))
    try:
        r = run(np.exp(np.asarray(x, np.float64)), observations=tp.obs[rows], want_grad=False)
    finally:
        run.close()
    return r["loss"]


def twin_problems(count: int, seed: int = 20261008, span: float = 0.3) -> TwinProblems:
    """`count` twin experiments on the calibration hull (sinksim_amd.workloads): hidden log factors uniform in
    [-span, span], observations = the signals of the hidden factors' runs (trim at 20/40/60/80 min, every monitor and
    mark event time; generic_loss weights)."""
    from .adjoint import GradientOptions, GradientRunner
    from .calib import generic_loss, zone_params
    from .workloads import CALIBRATION_STEPS, CALIBRATION_TRIM_STEPS, calibration_hull, calibration_problems
    ship, sim = calibration_hull()
    pm = zone_params(sim)
    spec = generic_loss(sim, CALIBRATION_TRIM_STEPS)
    truth = calibration_problems(count, seed=seed, span=span, nparams=pm.count)
    run = GradientRunner([(ship, sim)], np.zeros(count, np.int32), pm, spec, CALIBRATION_STEPS * sim.dt,
                         GradientOptions(checkpoint_every=1 << 16, relaxed=True, max_steps=CALIBRATION_STEPS))
    try:
        r = run(np.exp(truth), want_grad=False)
    finally:
        run.close()
    return TwinProblems(ship, sim, pm, spec, truth, r["signals"].copy(), CALIBRATION_STEPS)


def twin_search(tp: TwinProblems, rows=None, rounds: int = 5, batch: int = 264, span: float = 0.3, shrink: float = 0.5,
                seed: int = 20261008, every: float = 30.0, log=None) -> dict:
    """The reference's search (tools/py/calibrate_gpu.py:95-124) for each problem of `rows`: rounds of `batch` points
    uniform in a box of half-width span (log factors) around the problem's best so far, the first round holding 0
    (scales 1), the box shrinking by `shrink` each round; each round of every problem in one batch run of the plain
    forward kernel, scored by twin_J. Returns per problem the best J and point after each round, the device time
    (ms) and simulation steps of all rounds."""
    rows = np.arange(tp.count) if rows is None else np.asarray(rows)
    P, npar = len(rows), tp.params.count
    rng = np.random.default_rng(seed)
    center = np.zeros((P, npar))
    best_J = np.full(P, np.inf)
    best_x = np.zeros((P, npar))
    J_by_round, ms_total, steps_total = [], 0.0, 0
    sim = tp.sim
    en, tOn = np.asarray(sim.cEn, np.uint8), np.asarray(sim.cTOn, np.float64)
    for r in range(rounds):
        pts = center[:, None, :] + rng.uniform(-span, span, (P, batch, npar))
        if r == 0:
            pts[:, 0, :] = 0.0
        flat = pts.reshape(P * batch, npar)
        n = flat.shape[0]
        b = Batch(tp.ship, sim, BatchOptions(count=n, mixed_precision=False, traced_instances=0, curve_every=every,
                                             t_max=tp.steps * sim.dt))
        b.set_connections(np.tile(en, (n, 1)), _area_rows(sim, tp.params, flat), np.tile(tOn, (n, 1)))
        b.run()
        ms, _ = b.kernel_time()
        ms_total += ms
        steps_total += b.total_steps()
        J = np.empty(n)
        for i in range(n):
            res = b.result(i)
            J[i] = twin_J(b.curve(i), res.events, res.foundered, res.founderT, tp.spec, tp.obs[rows[i // batch]], sim.dt)
        b.close()
        J = J.reshape(P, batch)
        k = np.argmin(J, axis=1)
        rb = J[np.arange(P), k]
        better = rb < best_J
        best_J[better] = rb[better]
        best_x[better] = pts[np.arange(P), k][better]
        center = pts[np.arange(P), k]
        span *= shrink
        J_by_round.append(best_J.copy())
        if log:
            log(f"twin search round {r}: {n} runs, median best J {np.median(best_J):.4g}, device {ms_total / 1e3:.2f} s")
    return {"best_J": best_J, "best_x": best_x, "J_by_round": np.array(J_by_round), "device_ms": ms_total, "steps": steps_total,
            "evaluations": rounds * batch}


def twin_lbfgs(tp: TwinProblems, rows=None, iterations: int = 12, target=None, checkpoint_every: int = 0, memory: int = 6,
               max_step: float = 0.3, lo: float = -1.6, hi: float = 1.6, stop=None, log=None, numerics: str = "portable") -> Trace:
    """lbfgs_core for each problem of `rows` from scales 1 (x = 0), every iteration one adjoint call over all of them
    (relaxed gradients, through the torch.autograd.Function); target: per-problem objectives to stop at; numerics:
    the forward's (portable32: FP32 partials in the forward and its recomputations, the reverse in FP64)."""
    import torch

    from .adjoint import GradientOptions, GradientRunner, flood_loss
    rows = np.arange(tp.count) if rows is None else np.asarray(rows)
    P = len(rows)
    runner = GradientRunner([(tp.ship, tp.sim)], np.zeros(P, np.int32), tp.params, tp.spec, tp.steps * tp.sim.dt,
                            GradientOptions(checkpoint_every=checkpoint_every, relaxed=True, max_steps=tp.steps, numerics=numerics))
    obs = tp.obs[rows]

    def evaluate(x):
        xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
        loss = flood_loss(torch.exp(xt), runner, obs)
        loss.sum().backward()
        return loss.detach().numpy().copy(), xt.grad.numpy().copy()

    def progress():
        if log:
            log(f"twin lbfgs: call {runner.calls}, device {runner.kernel_ms / 1e3:.2f} s")
        return runner.kernel_ms, runner.sim_steps

    try:
        return lbfgs_core(evaluate, np.zeros((P, tp.params.count)), iterations=iterations, memory=memory, max_step=max_step, lo=lo,
                          hi=hi, target=target, progress=progress, stop=stop)
    finally:
        runner.close()


def twin_polish(tp: TwinProblems, x0, f0, rows=None, rounds: int = 8, samples: int = 16, box: float = 0.01, shrink: float = 0.5,
                target=None, seed: int = 20261008, log=None) -> Trace:
    """Rounds of `samples` forward-only runs per problem at points uniform in a box of half-width `box` (log factors)
    around the problem's best point so far (x0, f0 to start); a problem whose round finds nothing better halves its
    box. Stops after `rounds` rounds or once every problem is at or below its target."""
    from .adjoint import GradientOptions, GradientRunner
    rows = np.arange(tp.count) if rows is None else np.asarray(rows)
    P, npar = len(rows), tp.params.count
    rng = np.random.default_rng(seed)
    best_x, best_f = np.asarray(x0, np.float64).copy(), np.asarray(f0, np.float64).copy()
    half = np.full(P, float(box))
    tgt = np.full(P, -np.inf) if target is None else np.asarray(target, np.float64)
    runner = GradientRunner([(tp.ship, tp.sim)], np.zeros(P * samples, np.int32), tp.params, tp.spec, tp.steps * tp.sim.dt,
                            GradientOptions(checkpoint_every=1 << 16, relaxed=True, max_steps=tp.steps))
    obs = np.repeat(tp.obs[rows], samples, axis=0)
    trace = Trace()
    try:
        for r in range(rounds):
            if np.all(best_f <= tgt):
                break
            pts = best_x[:, None, :] + rng.uniform(-1.0, 1.0, (P, samples, npar)) * half[:, None, None]
            f = runner(np.exp(pts.reshape(P * samples, npar)), observations=obs, want_grad=False)["loss"].reshape(P, samples)
            f = np.where(np.isfinite(f), f, np.inf)
            k = np.argmin(f, axis=1)
            fb = f[np.arange(P), k]
            better = fb < best_f
            best_f[better] = fb[better]
            best_x[better] = pts[np.arange(P), k][better]
            half[~better] *= shrink
            i = int(np.argmin(best_f))
            trace.add(best_f[i], runner.kernel_ms, runner.sim_steps, (r + 1) * samples, best_x[i], best_f)
            if log:
                log(f"twin polish round {r}: median best J {np.median(best_f):.4g}, device {runner.kernel_ms / 1e3:.2f} s")
    finally:
        runner.close()
    trace.final_x = best_x
    return trace


def starts_compare(starts: int = 8192, iterations: int = 2, numerics: str = "portable32", problem: int = 0, rounds: int = 5,
                   batch: int = 264, span: float = 0.3, seed: int = 20261008, log=None) -> dict:
    """One twin problem on the calibration hull (the hidden factors of problem `problem` of twin_problems), the
    reference's search on it (rounds x batch runs) against batched L-BFGS from `starts` starts at once (scales 1 and
    starts - 1 points uniform in the search's first box), the forward in `numerics` and the reverse in FP64, stopping
    once a start reaches the search's best objective (re-scored with the differentiable loss) or after `iterations`
    iterations. Device times of both; the L-BFGS starts share every adjoint call."""
    t0 = time.perf_counter()
    tp = twin_problems(problem + 1, seed=seed)
    tp = TwinProblems(tp.ship, tp.sim, tp.params, tp.spec, tp.truth[problem:problem + 1], tp.obs[problem:problem + 1], tp.steps)
    sw = twin_search(tp, rounds=rounds, batch=batch, span=span, seed=seed, log=log)
    target = float(objective(tp, sw["best_x"], numerics=numerics)[0])
    rng = np.random.default_rng(seed + 1)
    x0 = np.vstack([np.zeros(tp.params.count), rng.uniform(-span, span, (starts - 1, tp.params.count))])
    rep = TwinProblems(tp.ship, tp.sim, tp.params, tp.spec, np.repeat(tp.truth, starts, axis=0), np.repeat(tp.obs, starts, axis=0),
                       tp.steps)
    from .adjoint import GradientOptions, GradientRunner, flood_loss
    import torch
    runner = GradientRunner([(rep.ship, rep.sim)], np.zeros(starts, np.int32), rep.params, rep.spec, rep.steps * rep.sim.dt,
                            GradientOptions(checkpoint_every=0, relaxed=True, max_steps=rep.steps, numerics=numerics))

    def evaluate(x):
        xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
        loss = flood_loss(torch.exp(xt), runner, rep.obs)
        loss.sum().backward()
        return loss.detach().numpy().copy(), xt.grad.numpy().copy()

    def progress():
        if log:
            log(f"starts lbfgs: call {runner.calls}, device {runner.kernel_ms / 1e3:.2f} s")
        return runner.kernel_ms, runner.sim_steps

    try:
        lb = lbfgs_core(evaluate, x0, iterations=iterations, target=target, progress=progress)
    finally:
        runner.close()
    return {"problems": tp, "search": sw, "lbfgs": lb, "target": target, "search_ms": sw["device_ms"],
            "lbfgs_ms_to_target": lb.time_to(target), "wall_s": time.perf_counter() - t0}


def twin_compare(problems: int = 256, search_problems: int = 8, iterations: int = 16, rounds: int = 5, batch: int = 264,
                 polish_rounds: int = 8, polish_samples: int = 16, seed: int = 20261008, log=None) -> dict:
    """twin_search on the first `search_problems` problems, then on all `problems` twin_lbfgs (stopping early once the
    median best objective improves by less than 5 % over two calls) and twin_polish, until each of the searched
    problems reaches its search's best objective (re-scored with the differentiable loss). Device time per problem of
    each method, and the share of searched problems reached after each L-BFGS call and polish round."""
    t0 = time.perf_counter()
    tp = twin_problems(problems, seed=seed)
    sub = np.arange(min(search_problems, problems))
    sw = twin_search(tp, sub, rounds=rounds, batch=batch, seed=seed, log=log)
    target = objective(tp, sw["best_x"], sub)
    full_target = np.full(problems, np.inf)  # unsearched problems: always reached
    full_target[sub] = target

    meds = []

    def stalled(tr):
        meds.append(float(np.median(tr.row_best[-1])))
        return len(meds) >= 3 and meds[-1] > 0.95 * meds[-3]

    lb = twin_lbfgs(tp, iterations=iterations, target=full_target, stop=stalled, log=log)
    rows = list(np.array(lb.row_best))
    ms = list(lb.device_ms)
    phase = ["lbfgs"] * len(rows)
    last = rows[-1]
    po = None
    if not np.all(last[sub] <= target) and polish_rounds > 0:
        po = twin_polish(tp, lb.final_x, last, rounds=polish_rounds, samples=polish_samples, target=full_target, seed=seed + 1, log=log)
        rows += list(np.array(po.row_best))
        ms += [lb.device_ms[-1] + m for m in po.device_ms]
        phase += ["polish"] * len(po.row_best)
    share = [float(np.mean(r[sub] <= target)) for r in rows]
    all_at = next((k for k, v in enumerate(share) if v >= 1.0), None)
    half_at = next((k for k, v in enumerate(share) if v >= 0.5), None)
    return {"problems": tp, "search": sw, "lbfgs": lb, "polish": po, "target": target, "share_reached": share, "phase": phase,
            "device_ms": ms, "median_J": [float(np.median(r)) for r in rows],
            "search_ms_per_problem": sw["device_ms"] / len(sub),
            "ms_per_problem_all": ms[all_at] / problems if all_at is not None else None,
            "ms_per_problem_half": ms[half_at] / problems if half_at is not None else None,
            "wall_s": time.perf_counter() - t0}
