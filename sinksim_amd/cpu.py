"""The host CPU engine of libsinksim_hip.so from Python: the reference's four numerics (oracle, portable, portable32,
std) on the CPU, with the semantics of its sweep::run_cpu (engine/src/sweep.cpp:121-161) and Simulation::record_trace
(engine/src/simulation.cpp:194-221). No device is used.

    from sinksim_amd import cpu
    out = cpu.sweep(ship, sim, "oracle", en, area, tOn, t_max=14400, every=30, threads=8)
    dense, samples, result, events = cpu.trace(ship, sim, "portable32", t_max=18000, dense=400, sample_every=60)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A


def _numerics(name: str) -> int:
    if name not in A.NUMERICS:
        raise ValueError(f"unknown numerics {name!r} (use oracle, portable, portable32 or std)")
    return A.NUMERICS[name]


def _tables(sim, count, en, area, tOn):
    nc = sim.connections
    out = []
    for arr, conv, ctype in ((en, A.u8, C.c_uint8), (area, A.f64, C.c_double), (tOn, A.f64, C.c_double)):
        if arr is None:
            out.append((None, None))
            continue
        a = conv(arr).reshape(-1)
        if a.size != count * nc:
            raise ValueError(f"connection tables must hold {count * nc} values")
        out.append((a, A.ptr(a, ctype)))
    return out


def sweep(ship, sim, numerics: str, en=None, area=None, tOn=None, count: int | None = None, t_max: float = 5 * 3600.0,
          every: float = 30.0, threads: int = 0, stop_when_stable: bool = False) -> dict:
    """Runs count instances (en/area/tOn: [count, nconn] tables, None for the scenario's) like Simulation::run; returns
    per-instance arrays: steps, foundered, founderT, tEnd, final7 [count, 7], curves (a list of [n, 7] arrays),
    events (a list of (t, id, kind) lists) and the wall time (s)."""
    lib = A.load_library()
    if count is None:
        count = next((np.asarray(x).reshape(-1).size // sim.connections for x in (en, area, tOn) if x is not None), 1)
    tabs = _tables(sim, count, en, area, tOn)
    s, m = A.ShipStruct(ship), A.SimStruct(sim)
    opt = A.SsCpuOptions(_numerics(numerics), int(threads), float(t_max), float(every), 1 if stop_when_stable else 0)
    if not (every > 0 and t_max > 0):
        raise ValueError("t_max and every must be positive")
    cap = int(np.floor(t_max / every)) + 3
    evcap = sim.monitors + sim.marks + 1
    res = (A.SsInstanceResult * count)()
    ev = (A.SsEvent * (count * evcap))()
    curves = np.zeros((count, cap, 7))
    ncurve = np.zeros(count, np.int32)
    final7 = np.zeros((count, 7))
    secs = C.c_double()
    A.check(lib.ss_cpu_sweep(C.byref(s.s), C.byref(m.s), C.byref(opt), count, tabs[0][1], tabs[1][1], tabs[2][1], res, ev, evcap,
                             A.ptr(curves, C.c_double), cap, A.ptr(ncurve, C.c_int32), A.ptr(final7, C.c_double), C.byref(secs)))
    if (ncurve > cap).any():
        raise RuntimeError("curve capacity exceeded")
    return {
        "steps": np.array([r.steps for r in res], np.int64),
        "foundered": np.array([bool(r.foundered) for r in res]),
        "founderT": np.array([r.founderT for r in res]),
        "tEnd": np.array([r.tEnd for r in res]),
        "final7": final7,
        "curves": [curves[i, :ncurve[i]].copy() for i in range(count)],
        "events": [[(ev[i * evcap + k].t, ev[i * evcap + k].id, ev[i * evcap + k].kind) for k in range(min(res[i].events, evcap))]
                   for i in range(count)],
        "seconds": secs.value,
    }


def trace(ship, sim, numerics: str, t_max: float = 5 * 3600.0, dense: int = 400, sample_every: float = 60.0, en=None, area=None,
          tOn=None):
    """One instance's trace like Simulation::record_trace: dense [n, 10 + 5 * nodes], samples [m, ...], the result as a
    dict and the events (t, id, kind)."""
    lib = A.load_library()
    tabs = _tables(sim, 1, en, area, tOn)
    s, m = A.ShipStruct(ship), A.SimStruct(sim)
    if not (t_max > 0 and sample_every > 0 and dense >= 0):
        raise ValueError("t_max and sample_every must be positive, dense not negative")
    rs = 10 + 5 * ship.nodes
    nd_cap, ns_cap = dense + 1, int(np.floor(t_max / sample_every)) + 2
    d, smp = np.zeros((nd_cap, rs)), np.zeros((ns_cap, rs))
    nd, ns = C.c_int32(), C.c_int32()
    r = A.SsInstanceResult()
    evcap = sim.monitors + sim.marks + 1
    ev = (A.SsEvent * evcap)()
    A.check(lib.ss_cpu_trace(C.byref(s.s), C.byref(m.s), _numerics(numerics), float(t_max), int(dense), float(sample_every),
                             tabs[0][1], tabs[1][1], tabs[2][1], A.ptr(d, C.c_double), nd_cap, C.byref(nd), A.ptr(smp, C.c_double),
                             ns_cap, C.byref(ns), C.byref(r), ev, evcap))
    result = {"foundered": bool(r.foundered), "founderT": r.founderT, "tEnd": r.tEnd, "steps": r.steps}
    events = [(ev[k].t, ev[k].id, ev[k].kind) for k in range(min(r.events, evcap))]
    return d[:nd.value], smp[:ns.value], result, events
