"""The forward-mode dual-number build of the portable step (oracle/dual_step.hpp), the checker of the adjoint.

Pins, none of which retypes the dual's own formulas:
- values: the dual build's state after N steps equals the portable oracle's record bit for bit (titanic and two
  synthetic hulls, one with a four-space group), so its branches are the oracle's;
- derivatives: its tangent along a parameter direction equals central finite differences (h = 1e-6 relative) of the
  ORACLE run with perturbed area and coefficient tables, to 1e-6 relative on every sizeable component, on horizons
  before the levels begin to chatter and where the perturbation flips no branch;
- the full gradient by N = 12 tangents equals the directional derivatives taken one at a time;
- event times: each interpolated crossing time lies in the step interval the oracle's event log names;
- relaxed derivatives keep the values and give a descent direction of the full-run twin loss.
"""
from types import SimpleNamespace

import numpy as np
import pytest


def _model(ship, sim):
    import pyoracle
    return pyoracle.Model(ship, sim)


def _tables(sim):
    return sim.cEn.astype(np.uint8), sim.cArea.astype(np.float64).copy(), sim.cTOn.astype(np.float64).copy()


def _oracle_record(ship, sim, pm, theta, steps):
    """The portable oracle's state record after `steps` steps with the tables scaled as the parameters say."""
    import pyoracle
    area = sim.cArea * np.where(pm.pa >= 0, theta[np.maximum(pm.pa, 0)], 1.0)
    coef = sim.cCoef * np.where(pm.pc >= 0, theta[np.maximum(pm.pc, 0)], 1.0)
    scaled = SimpleNamespace(**{k: getattr(sim, k) for k in dir(sim) if not k.startswith("_")})
    scaled.cCoef = coef
    o = pyoracle.OracleSim(ship, scaled, "portable", pyoracle.Model(ship, scaled))
    for c in range(sim.connections):
        o.set_conn(c, bool(sim.cEn[c]), float(area[c]), float(sim.cTOn[c]))
    o.reset()
    o.step(steps)
    return o.record()


@pytest.fixture(scope="module")
def hulls(titanic):
    """titanic; a synthetic hull of 1,500 columns, 16 spaces and eight two-space groups (seed 11); one with a
    four-space group, two never-merged groups and singles (seed 12)."""
    from sinksim_amd.synth import synthetic_hull
    return {"titanic": titanic, "synthetic": synthetic_hull(11, columns=1500, spaces=16),
            "synth4": synthetic_hull(12, columns=1500, spaces=16)}


@pytest.mark.parametrize("which", ["titanic", "synthetic", "synth4"])
def test_dual_values_equal_the_oracle(hulls, which):
    import pyoracle
    from sinksim_amd.calib import kind_params
    ship, sim = hulls[which]
    pm = kind_params(sim)
    en, area, tOn = _tables(sim)
    theta = np.exp(np.random.default_rng(3).uniform(-0.4, 0.4, pm.count))
    steps = 600
    r = pyoracle.dual_run(_model(ship, sim), en, area, tOn, pm.pa, pm.pc, theta, seed=np.ones((1, pm.count)), max_steps=steps)
    want = _oracle_record(ship, sim, pm, theta, steps)
    got = r["record"]
    keep = np.ones(len(want), bool)
    keep[7] = False  # inflow is not carried by the dual build's record
    assert np.array_equal(got[keep], want[keep])


# Horizons before the levels of titanic's midship spaces begin to chatter, and at which a relative perturbation of
# 1e-6 flips no branch of the discrete program (a flip, e.g. one more Newton iteration somewhere, shows as a
# finite difference that changes with h; at 3,000 and 6,000 steps of titanic h = 1e-6 crosses one, h = 1e-7 does not;
# the synthetic hulls chatter much earlier).
@pytest.mark.parametrize("which,steps", [("titanic", 2000), ("titanic", 5000), ("synthetic", 1000), ("synth4", 300)])
def test_dual_tangent_matches_oracle_finite_differences(hulls, which, steps):
    import pyoracle
    from sinksim_amd.calib import kind_params
    ship, sim = hulls[which]
    pm = kind_params(sim)
    en, area, tOn = _tables(sim)
    rng = np.random.default_rng(steps)
    theta = np.exp(rng.uniform(-0.2, 0.2, pm.count))
    v = rng.standard_normal(pm.count) * theta
    r = pyoracle.dual_run(_model(ship, sim), en, area, tOn, pm.pa, pm.pc, theta, seed=v[None], max_steps=steps)
    h = 1e-6
    fd = (_oracle_record(ship, sim, pm, theta + h * v, steps) - _oracle_record(ship, sim, pm, theta - h * v, steps)) / (2 * h)
    an = r["drecord"][0]
    sel = np.abs(fd) > 1e-5 * np.abs(fd).max()
    sel[[0, 7, 8, 9]] = False  # t; inflow, buoyancy and hull top are outputs, not state
    rel = np.abs(an[sel] - fd[sel]) / np.abs(fd[sel])
    assert sel.sum() > 20
    # the median is the FD noise floor (~1e-9); single components carry the FD truncation and rounding of h = 1e-6
    assert np.median(rel) < 1e-7 and rel.max() < 1e-4, (np.median(rel), rel.max())
    big = np.abs(fd[sel]) > 1e-2 * np.abs(fd).max()
    assert rel[big].max() < 1e-6, rel[big].max()


def test_dual_full_gradient_equals_directional_derivatives(titanic):
    import pyoracle
    from sinksim_amd.calib import kind_params, titanic_loss
    ship, sim = titanic
    pm = kind_params(sim)
    en, area, tOn = _tables(sim)
    model = _model(ship, sim)
    spec = titanic_loss(sim, trim_minutes=(2.0, 5.0, 8.0), trim_sigma=(0.3, 0.3, 0.3), events=())
    theta = np.exp(np.random.default_rng(9).uniform(-0.2, 0.2, pm.count))
    base = pyoracle.dual_run(model, en, area, tOn, pm.pa, pm.pc, theta, seed=np.zeros((1, pm.count)), max_steps=2000, spec=spec)
    spec = spec.with_observations(base["signals"] + 0.05)
    full = pyoracle.dual_run(model, en, area, tOn, pm.pa, pm.pc, theta, max_steps=2000, spec=spec)
    assert full["dloss"].shape == (pm.count,)
    for k in (0, 1, 3, 11):
        e = np.zeros(pm.count)
        e[k] = 1.0
        one = pyoracle.dual_run(model, en, area, tOn, pm.pa, pm.pc, theta, seed=e[None], max_steps=2000, spec=spec)
        assert one["loss"] == full["loss"]
        assert abs(one["dloss"][0] - full["dloss"][k]) <= 1e-12 * max(1.0, abs(full["dloss"][k]))


def test_event_times_lie_in_the_firing_step(titanic):
    import pyoracle
    from sinksim_amd.calib import generic_loss, kind_params
    ship, sim = titanic
    pm = kind_params(sim)
    en, area, tOn = _tables(sim)
    model = _model(ship, sim)
    spec = generic_loss(sim, trim_steps=[100])
    r = pyoracle.dual_run(model, en, area, tOn, pm.pa, pm.pc, np.ones(pm.count), seed=np.zeros((1, pm.count)), t_max=260 * 60.0,
                          spec=spec)
    o = pyoracle.OracleSim(ship, sim, "portable", model)
    o.reset()
    o.step(r["steps"])
    events, _, _ = o.events()
    t_end = r["record"][0]
    tstar = r["signals"][1:]
    fired = 0
    for e, (kind, idx) in enumerate(zip(spec.ev_kind, spec.ev_idx)):
        log = [t for t, i, k in events if k == kind and i == idx]
        if log:
            fired += 1
            assert log[0] - sim.dt < tstar[e] <= log[0], (kind, idx, tstar[e], log[0])
        else:
            assert tstar[e] == t_end
    assert fired >= 8


def test_relaxed_derivatives_keep_values_and_descend(titanic):
    import pyoracle
    from sinksim_amd.calib import kind_params, titanic_loss
    ship, sim = titanic
    pm = kind_params(sim)
    en, area, tOn = _tables(sim)
    model = _model(ship, sim)
    spec = titanic_loss(sim)
    truth = np.exp(np.random.default_rng(4).uniform(-0.3, 0.3, pm.count))
    obs = pyoracle.dual_run(model, en, area, tOn, pm.pa, pm.pc, truth, seed=np.zeros((1, pm.count)), t_max=260 * 60, spec=spec)
    spec = spec.with_observations(obs["signals"])
    theta = np.ones(pm.count)
    v = np.random.default_rng(8).standard_normal(pm.count)
    ex = pyoracle.dual_run(model, en, area, tOn, pm.pa, pm.pc, theta, seed=v[None], t_max=260 * 60, spec=spec)
    rx = pyoracle.dual_run(model, en, area, tOn, pm.pa, pm.pc, theta, seed=v[None], t_max=260 * 60, spec=spec, relaxed=True)
    assert np.array_equal(ex["record"], rx["record"]) and ex["loss"] == rx["loss"]
    # the exact derivative of the discrete program is useless over the full run (chattering levels); the relaxed
    # one is of the size of coarse finite differences and points downhill
    h = 0.01
    lp = pyoracle.dual_run(model, en, area, tOn, pm.pa, pm.pc, theta + h * v, seed=np.zeros((1, pm.count)), t_max=260 * 60, spec=spec)["loss"]
    lm = pyoracle.dual_run(model, en, area, tOn, pm.pa, pm.pc, theta - h * v, seed=np.zeros((1, pm.count)), t_max=260 * 60, spec=spec)["loss"]
    fd = (lp - lm) / (2 * h)
    d = rx["dloss"][0]
    assert np.sign(d) == np.sign(fd) and 0.5 < d / fd < 2.0, (d, fd)
    assert abs(ex["dloss"][0]) > 1e6 * abs(fd)


def test_plain_loss_run_equals_the_dual_builds_value(titanic):
    """The oracle's plain loss run (any numerics; the portable32 gradient's checker) in portable numerics gives the
    dual build's loss and signals bit for bit: the same events, interpolations and terms, with area and coefficient
    factors applied as the gradient applies them. In portable32 numerics it differs (FP32 partials), by little."""
    import pyoracle
    from sinksim_amd.calib import kind_params, titanic_loss
    ship, sim = titanic
    pm = kind_params(sim)
    spec = titanic_loss(sim)
    en, area, tOn = _tables(sim)
    theta = np.exp(np.random.default_rng(3).uniform(-0.3, 0.3, pm.count))
    model = _model(ship, sim)
    d = pyoracle.dual_run(model, en, area, tOn, pm.pa, pm.pc, theta, seed=np.zeros((1, pm.count)), max_steps=12000, spec=spec)
    r = pyoracle.loss_run(model, "portable", en, area, tOn, pm.pa, pm.pc, theta, spec, max_steps=12000)
    assert r["steps"] == d["steps"] == 12000
    assert r["loss"] == d["loss"] and np.array_equal(r["signals"], d["signals"])
    assert np.isfinite(r["signals"]).all() and (r["signals"][:3] != 0).all()  # the three trim readouts were taken
    r32 = pyoracle.loss_run(model, "portable32", en, area, tOn, pm.pa, pm.pc, theta, spec, max_steps=12000)
    assert r32["loss"] != r["loss"] and abs(r32["loss"] / r["loss"] - 1) < 1e-4


@pytest.mark.parametrize("amin,slope", [(4.0, 1000.0), (2500.0, 2500.0)])
def test_relaxed_level_tangent_is_one_over_the_surface_area(amin, slope):
    """Closed form of the relaxed level derivative on the box-barge rig (tests/test_kernel.cpp:163-227 of the
    reference: node 0 holds 3,000 m^3 at level -1 m over a 1,000 m^2 free surface, at even keel, no connections). A
    unit of volume added to node 0 raises its level by 1 / 1,000 m, or by 1 / amin where the space's minimum effective
    area amin exceeds the free surface (the floor); the volume solve at the stored level converges without a Newton
    step, so the discrete program's (exact) level does not move at all."""
    import pyoracle
    from sinksim_amd.synth import box_barge
    ship, sim = box_barge()
    ship.amin[0] = amin
    sim.vol0[0] = 50 * 20 * 3
    sim.level0[0] = -1.0
    model = pyoracle.Model(ship, sim)
    o = pyoracle.OracleSim(ship, sim, "portable", model)
    rec = o.record()
    nn = ship.nodes
    drec = np.zeros_like(rec)
    drec[10] = 1.0  # d vol[0]
    empty = np.zeros(0)
    for relaxed, want in ((True, 1.0 / slope), (False, 0.0)):
        r2, dr2, _, _ = pyoracle.dual_step(model, np.zeros(0, np.uint8), empty, empty, np.zeros(0, np.int32), np.zeros(0, np.int32),
                                           np.ones(1), np.zeros(1), rec, drec, relaxed=relaxed)
        assert r2[10 + nn] == -1.0 and r2[10] == 3000.0  # the level and the volume themselves
        assert dr2[10] == 1.0                             # the volume carries its tangent
        assert dr2[10 + nn] == pytest.approx(want, rel=1e-12, abs=0.0), (relaxed, dr2[10 + nn])
