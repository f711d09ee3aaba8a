"""The calibration driver's host logic on the CPU (sinksim_amd/calibrate.py): the batched L-BFGS on functions with
known minima, and the search's candidate points."""
import math
import random

import numpy as np


def test_lbfgs_core_minimises_a_batch_of_quadratics_and_rosenbrock():
    from sinksim_amd.calibrate import lbfgs_core
    rng = np.random.default_rng(0)
    A = rng.standard_normal((4, 4))
    H = A @ A.T + 4 * np.eye(4)
    xs = np.array([0.3, -0.2, 0.1, 0.4])

    def quad(x):
        r = x - xs
        return np.einsum("si,ij,sj->s", r, H, r), 2 * r @ H

    tr = lbfgs_core(quad, rng.uniform(-1, 1, (5, 4)), iterations=40, max_step=0.5, lo=-3, hi=3)
    assert tr.best_J[-1] < 1e-10
    assert np.allclose(tr.best_x[-1], xs, atol=1e-5)
    assert all(a >= b for a, b in zip(tr.best_J, tr.best_J[1:]))  # the best objective never rises

    def rosen(x):
        a, b = x[:, 0], x[:, 1]
        f = (1 - a) ** 2 + 100 * (b - a * a) ** 2
        g = np.stack([-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)], 1)
        return f, g

    tr = lbfgs_core(rosen, np.array([[-1.2, 1.0], [0.0, 0.0]]), iterations=300, max_step=0.5, lo=-3, hi=3, target=1e-8)
    assert tr.best_J[-1] <= 1e-8


def test_lbfgs_core_stops_at_target():
    from sinksim_amd.calibrate import lbfgs_core
    calls = []

    def f(x):
        calls.append(1)
        return np.sum(x ** 2, 1), 2 * x

    tr = lbfgs_core(f, np.ones((3, 2)), iterations=50, max_step=0.4, target=0.5)
    assert tr.best_J[-1] <= 0.5 and len(calls) < 10


def test_search_candidates_follow_the_reference_draws(monkeypatch):
    """The search draws its candidates as the reference does: random.Random(seed), scales 1 first, then per point and
    kind center + uniform(-span, span), the box shrinking by `shrink` around the round's best."""
    from sinksim_amd import calibrate as CAL
    seen = []

    class FakeBatch:
        def __init__(self, ship, sim, opt):
            self.n = opt.count

        def set_connections(self, en, area, tOn):
            seen.append(np.asarray(area).copy())

        def run(self):
            return 0.0

        def kernel_time(self):
            return 1.0, 1

        def total_steps(self):
            return self.n

        def result(self, i):
            class R:
                events, foundered, founderT = [], False, -1.0
            return R()

        def curve(self, i):
            return np.zeros((2, 7))

        def close(self):
            pass

    from sinksim_amd.data_paths import titanic
    ship, sim = titanic("titanic")
    monkeypatch.setattr(CAL, "Batch", FakeBatch)
    # score each point by its distance to a fixed point, so that the centre moves
    target = np.array([0.1, -0.05, 0.2, 0.0])
    calls = {"i": 0}

    def fake_J(curve, events, foundered, founderT, sim_, obs=None):
        k = calls["i"]
        calls["i"] += 1
        area = seen[-1][k % seen[-1].shape[0]]
        x = [math.log(area[np.asarray(sim.cKind) == sim.kind_index(kd)][0] / sim.cArea[np.asarray(sim.cKind) == sim.kind_index(kd)][0])
             for kd in CAL.CALIBRATION_KINDS]
        return float(np.sum((np.array(x) - target) ** 2))

    monkeypatch.setattr(CAL, "reference_J", fake_J)
    tr = CAL.search(ship, sim, rounds=2, batch=6, span=0.3, shrink=0.5, seed=20261008)
    rng = random.Random(20261008)
    kinds = CAL.CALIBRATION_KINDS
    first = [[0.0] * 4] + [[rng.uniform(-0.3, 0.3) for _ in kinds] for _ in range(5)]
    got = [[math.log(seen[0][i][np.asarray(sim.cKind) == sim.kind_index(k)][0] / sim.cArea[np.asarray(sim.cKind) == sim.kind_index(k)][0])
            for k in kinds] for i in range(6)]
    assert np.allclose(got, first, atol=1e-12)
    assert len(tr.best_J) == 2 and tr.evaluations == [6, 12]


def test_lbfgs_core_per_start_targets():
    """With one target per start the run stops once every start has reached its own; row_best tracks each start."""
    from sinksim_amd.calibrate import lbfgs_core
    c = np.array([[0.5, -0.2], [-0.1, 0.3], [0.2, 0.2]])

    def evaluate(x):
        d = x - c
        return np.sum(d * d, axis=1), 2 * d

    target = np.array([1e-3, 1e-3, np.inf])
    tr = lbfgs_core(evaluate, np.zeros((3, 2)), iterations=30, target=target)
    rows = np.array(tr.row_best)
    assert rows.shape[1] == 3 and len(rows) == len(tr.best_J)
    assert np.all(np.diff(rows, axis=0) <= 0)
    assert np.all(rows[-1][:2] <= 1e-3)
    assert not np.all(rows[-2][:2] <= 1e-3)  # it stopped at the first call where both had reached


def test_twin_J_tracks_the_differentiable_loss():
    """twin_J (curve and event log, event times at their firing step) against the dual build's loss (event times
    interpolated within the step) on the calibration hull, 20,000 steps: equal but for the interpolation, which moves
    each event term w r^2 by at most w (2 |r| dt + dt^2)."""
    import pyoracle
    from sinksim_amd.calib import generic_loss, zone_params
    from sinksim_amd.calibrate import _area_rows, twin_J
    from sinksim_amd.workloads import CALIBRATION_STEPS, CALIBRATION_TRIM_STEPS, calibration_hull
    ship, sim = calibration_hull()
    pm = zone_params(sim)
    spec = generic_loss(sim, CALIBRATION_TRIM_STEPS)
    rng = np.random.default_rng(9)
    x_true, x = rng.uniform(-0.3, 0.3, (2, pm.count))
    model = pyoracle.Model(ship, sim)
    en, tOn = sim.cEn.astype(np.uint8), sim.cTOn.copy()
    truth = pyoracle.dual_run(model, en, sim.cArea.copy(), tOn, pm.pa, pm.pc, np.exp(x_true), seed=np.zeros((1, pm.count)),
                              max_steps=CALIBRATION_STEPS, spec=spec)
    obs = truth["signals"]
    assert truth["steps"] == CALIBRATION_STEPS and np.sum(obs[4:] < 18000) >= 3
    d = pyoracle.dual_run(model, en, sim.cArea.copy(), tOn, pm.pa, pm.pc, np.exp(x), seed=np.zeros((1, pm.count)),
                          max_steps=CALIBRATION_STEPS, spec=spec.with_observations(obs))
    area = _area_rows(sim, pm, x[None])
    out = pyoracle.sweep(ship, sim, "portable", en[None], area, tOn[None], CALIBRATION_STEPS * sim.dt, 30.0, threads=1)[0]
    assert out["steps"] == CALIBRATION_STEPS
    J = twin_J(out["curve"], out["events"], out["foundered"], out["founderT"], spec, obs, sim.dt)
    n = len(spec.read_steps)
    r = np.abs(d["signals"][n:] - obs[n:])
    bound = float(np.sum(spec.ev_w * (2 * r * sim.dt + sim.dt ** 2)))
    assert abs(J - d["loss"]) <= bound, (J, d["loss"], bound)
    assert np.allclose(d["signals"][:n], [np.interp(s * sim.dt, out["curve"][:, 0], out["curve"][:, 1]) for s in spec.read_steps], rtol=0, atol=0)
