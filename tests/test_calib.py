"""The calibration objective and parameter maps (sinksim_amd/calib.py) on the CPU.

- reference_J, the reference's objective restated (tools/py/calibrate_gpu.py:35-64), evaluated on the oracle-numerics
  run at scales 1 gives the value the reference documents (J = 10.411 in oracle numerics, docs/DETERMINISM.md:16-17).
- The differentiable form of the same objective (reference_objective: readouts, founder and event times interpolated
  within their step) differs from reference_J on the same run by less than the interpolation can move it.
- Parameter maps: 12 titanic kind parameters, 4 reference calibration scales, 24 kind-zone factors on a synthetic hull.
"""
import numpy as np
import pytest


def test_reference_J_at_scales_one_is_the_documented_value(titanic):
    import pyoracle
    from sinksim_amd.calib import reference_J
    ship, sim = titanic
    en, area, tOn = sim.cEn.astype(np.uint8)[None], sim.cArea[None], sim.cTOn[None]
    out = pyoracle.sweep(ship, sim, "oracle", en, area, tOn, 260 * 60.0, 30.0, threads=1)[0]
    J = reference_J(out["curve"], out["events"], out["foundered"], out["founderT"], sim)
    assert abs(J - 10.411) < 5e-4, J


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_differentiable_objective_tracks_reference_J(titanic, seed):
    import pyoracle
    from sinksim_amd.calib import kind_params, reference_J, reference_objective
    ship, sim = titanic
    pm = kind_params(sim, coef_kinds=())
    assert pm.count == 4 and pm.names == ["area:breach", "area:over", "area:down", "area:top"]
    theta = np.ones(4) if seed == 0 else np.exp(np.random.default_rng(seed).uniform(-0.3, 0.3, 4))
    area = sim.cArea * np.where(pm.pa >= 0, theta[np.maximum(pm.pa, 0)], 1.0)
    en, tOn = sim.cEn.astype(np.uint8), sim.cTOn
    out = pyoracle.sweep(ship, sim, "portable", en[None], area[None], tOn[None], 260 * 60.0, 30.0, threads=1)[0]
    J = reference_J(out["curve"], out["events"], out["foundered"], out["founderT"], sim)
    spec = reference_objective(sim)
    d = pyoracle.dual_run(pyoracle.Model(ship, sim), en, sim.cArea.copy(), tOn, pm.pa, pm.pc, theta, seed=np.zeros((1, 4)),
                          t_max=260 * 60.0, spec=spec)
    assert d["steps"] == out["steps"]
    # readouts at whole minutes are curve records exactly; an event or founder time moves by less than one step dt,
    # which moves its term w r^2 by at most w (2 |r| dt + dt^2)
    n = len(spec.read_steps)
    ev_times = d["signals"][n:]
    bound = float(np.sum(spec.ev_w * (2 * np.abs(ev_times - spec.ev_obs) * sim.dt + sim.dt ** 2)))
    assert abs(d["loss"] - J) <= bound, (d["loss"], J, bound)
    assert abs(ev_times[0] - out["founderT"]) < sim.dt if out["foundered"] else ev_times[0] == 18000.0


def test_parameter_maps(titanic):
    from sinksim_amd.calib import kind_params, zone_params
    from sinksim_amd.synth import synthetic_hull
    ship, sim = titanic
    pm = kind_params(sim)
    assert pm.count == 12 and pm.names[:4] == ["area:breach", "area:over", "area:down", "area:top"]
    assert set(np.unique(pm.pa)) <= set(range(-1, 4)) and set(np.unique(pm.pc)) <= set(range(-1, 12))
    breach = sim.kind_index("breach")
    assert np.all(pm.pa[sim.cKind == breach] == 0) and np.all(pm.pc[sim.cKind == breach] == 4 + 7)
    _, hsim = synthetic_hull(5, columns=16000, spaces=128, connections=600)
    z = zone_params(hsim)
    assert z.count == 24 and np.all(z.pc == -1)
    used = np.unique(z.pa[z.pa >= 0])
    assert used.size >= 12  # most kind-zone pairs have connections
