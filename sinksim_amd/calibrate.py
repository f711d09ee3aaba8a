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
    (ms), the cumulative simulation steps and evaluations, and the best point (log scales)."""
    best_J: list = field(default_factory=list)
    device_ms: list = field(default_factory=list)
    steps: list = field(default_factory=list)
    evaluations: list = field(default_factory=list)
    best_x: list = field(default_factory=list)

    def add(self, J, ms, steps, evals, x):
        self.best_J.append(float(J)); self.device_ms.append(float(ms)); self.steps.append(int(steps)); self.evaluations.append(int(evals))
        self.best_x.append(np.asarray(x, np.float64).tolist())

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
               target: float | None = None, progress=None) -> Trace:
    """Batched L-BFGS: evaluate(x [starts, P]) -> (f [starts], g [starts, P]) evaluates every start at once.

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
    ms, st = progress() if progress else (0.0, 0)
    trace.add(best[0], ms, st, S, best[1])
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

    for it in range(iterations):
        if target is not None and best[0] <= target:
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
        ms, st = progress() if progress else (0.0, 0)
        trace.add(best[0], ms, st, S * (it + 2), best[1])
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


def compare(ship, sim, starts: int = 8, seed: int = 20261008, iterations: int = 12, rounds: int = 5, batch: int = 264, log=None) -> dict:
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
