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
