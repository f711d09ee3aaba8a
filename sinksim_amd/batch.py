"""The batch engine from Python: the surface of the reference's sinksim::cuda::Batch
(engine/include/sinksim/cuda/batch.hpp:56-87) over the C ABI of include/sinksim_hip.h.

    from sinksim_amd.model import load_ship, load_sim
    from sinksim_amd.batch import Batch, BatchOptions
    b = Batch(ship, sim, BatchOptions(count=1024, mixed_precision=False, curve_every=30, t_max=14400))
    b.set_connection_area(3, 289, 2.4)
    seconds = b.run()
    b.total_steps(), b.result(3), b.curve(3)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi as A


@dataclass
class BatchOptions:
    """BatchOptions (engine/include/sinksim/cuda/batch.hpp:27-37)."""
    count: int = 1
    mixed_precision: bool = True     # portable32 numerics; False computes the portable (double) numerics
    traced_instances: int = 1
    dense_steps: int = 400
    sample_every: float = 60.0
    curve_every: float = 0.0
    t_max: float = 5 * 3600.0
    steps_per_launch: int = 0        # steps per launch or epoch; <= 0: the library's choice


@dataclass
class InstanceResult:
    """InstanceResult (engine/include/sinksim/cuda/batch.hpp:45-54)."""
    foundered: bool
    founderT: float
    tEnd: float
    steps: int
    state: np.ndarray                # t zO pitch roll vz wth wph
    outputs: np.ndarray              # inflow Vb hullTop mw bx by bz
    vol: np.ndarray
    level: np.ndarray
    cx: np.ndarray
    cy: np.ndarray
    cz: np.ndarray
    events: list = field(default_factory=list)   # (t, id, kind)


def available() -> tuple[bool, str]:
    """cuda::available (engine/cuda/batch.cu:71-83)."""
    lib = A.load_library()
    buf = C.create_string_buffer(512)
    ok = lib.ss_available(buf, 512) == A.SS_OK
    return ok, buf.value.decode()


def device_info() -> dict:
    """cuda::device_info (engine/cuda/batch.cu:85-95)."""
    lib = A.load_library()
    p = A.SsDeviceProperties()
    A.check(lib.ss_device_info(C.byref(p)))
    return {"name": p.name.decode(), "arch": p.arch.decode(), "computeUnits": p.computeUnits, "major": p.major,
            "minor": p.minor, "globalMemory": p.globalMemory, "sharedPerBlock": p.sharedPerBlock,
            "runtimeVersion": p.runtimeVersion, "driverVersion": p.driverVersion}


class Batch:
    """Many simulations of one ship and scenario on the GPU, computing the portable numerics bit for bit."""

    def __init__(self, ship, sim, opt: BatchOptions, _fleet=None):
        self._lib = A.load_library()
        self.ship, self.sim, self.opt = ship, sim, opt
        o = A.SsBatchOptions(int(opt.count), A.SS_NUMERICS_PORTABLE32 if opt.mixed_precision else A.SS_NUMERICS_PORTABLE,
                             int(opt.traced_instances), int(opt.dense_steps), float(opt.sample_every),
                             float(opt.curve_every), float(opt.t_max), int(opt.steps_per_launch))
        h = C.c_void_p()
        if _fleet is None:
            self.models = [(ship, sim)]
            self.inst_model = np.zeros(int(opt.count), np.int32)
            self._structs = [(A.ShipStruct(ship), A.SimStruct(sim))]
            A.check(self._lib.ss_batch_create(C.byref(self._structs[0][0].s), C.byref(self._structs[0][1].s), C.byref(o), C.byref(h)))
        else:
            self.models, self.inst_model = _fleet
            self._structs = [(A.ShipStruct(sh), A.SimStruct(sm)) for sh, sm in self.models]
            ships = (A.SsShip * len(self.models))(*[a.s for a, _ in self._structs])
            sims = (A.SsSim * len(self.models))(*[b.s for _, b in self._structs])
            A.check(self._lib.ss_batch_create_fleet(ships, sims, len(self.models), A.ptr(self.inst_model, C.c_int32), C.byref(o),
                                                    C.byref(h)))
        self._h = h
        self.count = int(opt.count)
        self.traced = min(int(opt.traced_instances), self.count)

    @classmethod
    def fleet(cls, models, inst_model, opt: BatchOptions) -> "Batch":
        """An uneven batch: `models` a list of (ship, sim), instance i running models[inst_model[i]] (no counterpart
        in the reference, whose batch runs one hull). One persistent launch refills its workgroups as simulations
        end."""
        im = np.ascontiguousarray(inst_model, dtype=np.int32)
        if im.shape != (int(opt.count),):
            raise ValueError("inst_model must hold one model index per instance")
        return cls(models[0][0], models[0][1], opt, _fleet=(list(models), im))

    def model_of(self, instance: int):
        """(ship, sim) of an instance."""
        return self.models[int(self.inst_model[instance])]

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.ss_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ set-up (before run)

    def restore(self, instance: int, state7, vol, level, cx, cy, cz) -> None:
        arrs = [A.f64(state7), A.f64(vol), A.f64(level), A.f64(cx), A.f64(cy), A.f64(cz)]
        A.check(self._lib.ss_batch_restore(self._h, instance, *[A.ptr(a, C.c_double) for a in arrs]))

    def set_connection_enabled(self, instance: int, c: int, en: bool) -> None:
        A.check(self._lib.ss_batch_set_connection_enabled(self._h, instance, c, 1 if en else 0))

    def set_connection_area(self, instance: int, c: int, area: float) -> None:
        A.check(self._lib.ss_batch_set_connection_area(self._h, instance, c, float(area)))

    def set_connection_activation(self, instance: int, c: int, tOn: float) -> None:
        A.check(self._lib.ss_batch_set_connection_activation(self._h, instance, c, float(tOn)))

    def set_connections(self, en=None, area=None, tOn=None) -> None:
        """Every instance's connection table at once: [count, nconn] arrays, or for a fleet each instance's row of
        its own model's connection count concatenated (None keeps a field)."""
        n = int(sum(self.models[m][1].connections for m in self.inst_model))
        e = A.u8(en).reshape(-1) if en is not None else None
        a = A.f64(area).reshape(-1) if area is not None else None
        t = A.f64(tOn).reshape(-1) if tOn is not None else None
        for arr in (e, a, t):
            if arr is not None and arr.size != n:
                raise ValueError(f"connection tables must hold {n} values (every instance's connections)")
        A.check(self._lib.ss_batch_set_connections(
            self._h, A.ptr(e, C.c_uint8) if e is not None else None, A.ptr(a, C.c_double) if a is not None else None,
            A.ptr(t, C.c_double) if t is not None else None))
        self._keep = (e, a, t)

    def set_launch(self, threads_per_sim: int = 0, workgroups: int = 0) -> None:
        """The launch shape: 256, 512 or 1,024 threads per simulation (0, the default: 256, and 512 and 1,024 for a
        run's last epochs once its live instances all fit resident at those sizes), and the workgroup count (0: as many
        as are resident at once; fewer makes the kernel refill them). The shape changes the time, never the results."""
        A.check(self._lib.ss_batch_set_launch(self._h, int(threads_per_sim), int(workgroups)))

    def set_stop_when_stable(self, on: bool = True) -> None:
        """The oracle's stopWhenStable rule (engine/src/simulation.cpp:179-182) for every instance: past 30 min, stop
        after 20 min of calm (small inflow, pitch rate and heave velocity). The reference's GPU batch refuses it."""
        A.check(self._lib.ss_batch_set_stop_when_stable(self._h, 1 if on else 0))

    def set_workgroups(self, n: int) -> None:
        self.set_launch(0, n)

    def set_record_layout(self, record_bytes: int) -> None:
        """The double geometry records' device layout: 48 bytes (padded, the default) or 40 (packed); the time
        changes, the results do not (ss_batch_set_record_layout)."""
        A.check(self._lib.ss_batch_set_record_layout(self._h, int(record_bytes)))

    def autotune(self, sample: int = 1024, trial_steps: int = 400) -> dict:
        """Measures each launch shape on the first `sample` instances of the work order for trial_steps steps (on a
        trial copy) and sets the fastest; returns the choice and the measured rates (steps/s)."""
        r = A.SsTuneResult()
        A.check(self._lib.ss_batch_autotune(self._h, int(sample), int(trial_steps), C.byref(r)))
        trials = [{"threads_per_sim": r.trialThreads[k], "workgroups": r.trialWorkgroups[k], "record_bytes": r.trialRecordBytes[k],
                   "rate": r.trialRate[k]} for k in range(r.trials)]
        return {"threads_per_sim": r.threadsPerSim, "workgroups": r.workgroups, "record_bytes": r.recordBytes, "rate_256": r.rate[0],
                "rate_512": r.rate[1], "rate_1024": r.rate[2], "trials": trials}

    # ------------------------------------------------------------------ run and results

    def run(self) -> float:
        """Batch::run: every instance to its founder or tMax (in epochs when every instance is resident; see
        ss_batch_run). Returns the wall time of the device work (s)."""
        s = C.c_double()
        A.check(self._lib.ss_batch_run(self._h, C.byref(s)))
        return s.value

    def step(self, steps: int = 1, per_instance=None) -> float:
        """kernel::step over the batch (engine/include/sinksim/kernel/step.hpp:41-54): every instance takes exactly
        `steps` steps (per_instance[i] for instance i), whether or not it founders or passes tMax. The first call starts
        from the staged or restored states, later calls continue. Returns the launch's wall time (s)."""
        s = C.c_double()
        if per_instance is None:
            A.check(self._lib.ss_batch_step(self._h, int(steps), None, C.byref(s)))
        else:
            p = np.ascontiguousarray(per_instance, dtype=np.int64)
            if p.shape != (self.count,):
                raise ValueError("per_instance must hold one step count per instance")
            A.check(self._lib.ss_batch_step(self._h, 0, p.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(s)))
        return s.value

    def state_records(self) -> np.ndarray:
        """Every instance's state in the trace layout (10 + 5 * nodes values), [count, record] for a single-model batch
        (a flat array of each instance's record for a fleet)."""
        sizes = [10 + 5 * self.model_of(i)[0].nodes for i in range(self.count)]
        out = np.zeros(int(sum(sizes)))
        A.check(self._lib.ss_batch_state_records(self._h, A.ptr(out, C.c_double)))
        return out.reshape(self.count, sizes[0]) if len(set(sizes)) == 1 else out

    def total_steps(self) -> int:
        v = C.c_int64()
        A.check(self._lib.ss_batch_total_steps(self._h, C.byref(v)))
        return v.value

    def shared_memory_bytes(self) -> int:
        v = C.c_size_t()
        A.check(self._lib.ss_batch_shared_memory_bytes(self._h, C.byref(v)))
        return v.value

    def kernel_time(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int32()
        A.check(self._lib.ss_batch_kernel_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def results(self) -> dict:
        """Every instance's steps, founder flag, founder time and end time (one call, [count] arrays)."""
        n = self.count
        steps, f, ft, te = np.zeros(n, np.int64), np.zeros(n, np.int32), np.zeros(n), np.zeros(n)
        A.check(self._lib.ss_batch_results(self._h, steps.ctypes.data_as(C.POINTER(C.c_int64)), A.ptr(f, C.c_int32),
                                           A.ptr(ft, C.c_double), A.ptr(te, C.c_double)))
        return {"steps": steps, "foundered": f.astype(bool), "founderT": ft, "tEnd": te}

    def result(self, instance: int) -> InstanceResult:
        ship, sim = self.model_of(instance)
        nn = ship.nodes
        r = A.SsInstanceResult()
        arrs = [np.zeros(nn) for _ in range(5)]
        cap = sim.monitors + sim.marks + 1
        ev = (A.SsEvent * cap)()
        A.check(self._lib.ss_batch_result(self._h, instance, C.byref(r), *[A.ptr(a, C.c_double) for a in arrs], ev, cap))
        events = [(ev[k].t, ev[k].id, ev[k].kind) for k in range(min(r.events, cap))]
        return InstanceResult(bool(r.foundered), r.founderT, r.tEnd, r.steps, np.array(r.state[:]),
                              np.array(r.outputs[:]), *arrs, events=events)

    def trace(self, instance: int) -> tuple[np.ndarray, np.ndarray]:
        """Dense and sampled records of a traced instance, [n, 10 + 5 * nodes] each."""
        nd, ns = C.c_int32(), C.c_int32()
        A.check(self._lib.ss_batch_trace(self._h, instance, None, 0, C.byref(nd), None, 0, C.byref(ns)))
        rs = 10 + 5 * self.model_of(instance)[0].nodes
        d, s = np.zeros((nd.value, rs)), np.zeros((ns.value, rs))
        A.check(self._lib.ss_batch_trace(self._h, instance, A.ptr(d, C.c_double), nd.value, C.byref(nd),
                                         A.ptr(s, C.c_double), ns.value, C.byref(ns)))
        return d, s

    def curve(self, instance: int) -> np.ndarray:
        """[n, 7] records (t trim list water inflow draftF draftA), closed by the final record."""
        n = C.c_int32()
        A.check(self._lib.ss_batch_curve(self._h, instance, None, 0, C.byref(n)))
        out = np.zeros((n.value, 7))
        A.check(self._lib.ss_batch_curve(self._h, instance, A.ptr(out, C.c_double), n.value, C.byref(n)))
        return out

    def final_readouts(self, instance: int) -> np.ndarray:
        out = np.zeros(7)
        A.check(self._lib.ss_batch_final_readouts(self._h, instance, A.ptr(out, C.c_double)))
        return out

    def event_id(self, kind: int, id_: int) -> str:
        return self.sim.event_id(kind, id_)

    def event_label(self, kind: int, id_: int) -> str:
        return self.sim.event_label(kind, id_)

    # ------------------------------------------------------------------ the gradient

    def _param_tables(self, params, theta):
        """Per-instance connection tables of parameter indices (laid out like the connection tables) and theta
        [count, np]; params: one calib.ParamMap, or one per model of a fleet."""
        maps = [params[m] if isinstance(params, (list, tuple)) else params for m in self.inst_model]
        npar = maps[0].count
        if any(p.count != npar for p in maps):
            raise ValueError("every model's parameter map must have the same parameter count")
        pa = A.i32(np.concatenate([np.asarray(p.pa, np.int32) for p in maps]))
        pc = A.i32(np.concatenate([np.asarray(p.pc, np.int32) for p in maps]))
        th = A.f64(np.broadcast_to(np.asarray(theta, np.float64), (self.count, npar)))
        return pa, pc, th, npar

    def gradient(self, params, theta, spec, observations=None, checkpoint_every: int = 256, relaxed: bool = False,
                 t_max: float = 0.0, max_steps: int = 0, want_grad: bool = True) -> dict:
        """Every instance's loss (spec: a calib.LossSpec, whose observations are replaced by `observations`
        [count, signals] when given) and its gradient with respect to theta [count, np], by the adjoint kernel with a
        checkpoint every `checkpoint_every` steps. params: a calib.ParamMap (one model) or one per model of a fleet.
        Returns loss [count], grad [count, np] (None without want_grad), signals [count, signals], steps [count],
        mismatches [count] (recomputed steps whose values differed from the forward's: 0) and the kernel time (ms).
        The forward runs in the batch's numerics; a portable32 batch's reverse is FP64 at the portable32 states."""
        pa, pc, th, npar = self._param_tables(params, theta)
        nT, nE = len(spec.read_steps), len(spec.ev_kind)
        ts = np.ascontiguousarray(spec.read_steps, np.int64)
        tk, tw, ta = A.i32(spec.read_kind), A.f64(spec.read_w), A.f64(spec.read_after)
        ek, ei, ew, em = A.i32(spec.ev_kind), A.i32(spec.ev_idx), A.f64(spec.ev_w), A.f64(spec.ev_missing)
        L = A.SsLoss(nT, ts.ctypes.data_as(C.POINTER(C.c_int64)), A.ptr(tk, C.c_int32), A.ptr(tw, C.c_double), A.ptr(ta, C.c_double), nE,
                     A.ptr(ek, C.c_int32), A.ptr(ei, C.c_int32), A.ptr(ew, C.c_double), A.ptr(em, C.c_double))
        nsig = nT + nE
        if observations is None:
            obs = np.tile(spec.observations, (self.count, 1))
        else:
            obs = A.f64(observations).reshape(self.count, nsig)
        obs = A.f64(obs)
        opt = A.SsGradientOptions(int(checkpoint_every), 1 if relaxed else 0, float(t_max), int(max_steps))
        loss = np.zeros(self.count)
        grad = np.zeros((self.count, npar)) if want_grad else None
        sig = np.zeros((self.count, max(nsig, 1)))
        steps = np.zeros(self.count, np.int64)
        mism = np.zeros(self.count, np.int32)
        ms = C.c_double()
        A.check(self._lib.ss_batch_gradient(self._h, npar, A.ptr(pa, C.c_int32), A.ptr(pc, C.c_int32), A.ptr(th, C.c_double), C.byref(L),
                                            A.ptr(obs, C.c_double), C.byref(opt), A.ptr(loss, C.c_double),
                                            A.ptr(grad, C.c_double) if grad is not None else None, A.ptr(sig, C.c_double),
                                            steps.ctypes.data_as(C.POINTER(C.c_int64)), A.ptr(mism, C.c_int32), C.byref(ms)))
        return {"loss": loss, "grad": grad, "signals": sig[:, :nsig], "steps": steps, "mismatches": mism, "ms": ms.value}

    def gradient_launch(self) -> dict:
        """The last gradient call's launch shape: threads per simulation, whether the tape's large arrays were in
        global memory, LDS bytes per simulation."""
        t, g, l = C.c_int32(), C.c_int32(), C.c_int64()
        A.check(self._lib.ss_batch_gradient_launch(self._h, C.byref(t), C.byref(g), C.byref(l)))
        return {"threads": t.value, "global_tape": bool(g.value), "lds_bytes": l.value}

    def adjoint_step(self, params, theta, record, lam, monitor_seed=None, mark_seed=None, relaxed: bool = False):
        """One step of instance 0 in reverse (the adjoint's test hook): from a state record (trace layout) and the
        adjoint of the state after the step, the adjoint of the state before it, d loss / d log theta, and the
        recomputed state after the step."""
        pa, pc, th, npar = self._param_tables(params, theta)
        rec, lm = A.f64(record), A.f64(lam)
        ms = A.f64(monitor_seed) if monitor_seed is not None else None
        mk = A.f64(mark_seed) if mark_seed is not None else None
        lo, gl, ro = np.zeros_like(rec), np.zeros(npar), np.zeros_like(rec)
        A.check(self._lib.ss_batch_adjoint_step(self._h, npar, A.ptr(pa, C.c_int32), A.ptr(pc, C.c_int32), A.ptr(th, C.c_double),
                                                1 if relaxed else 0, A.ptr(rec, C.c_double), A.ptr(lm, C.c_double),
                                                A.ptr(ms, C.c_double) if ms is not None else None,
                                                A.ptr(mk, C.c_double) if mk is not None else None, A.ptr(lo, C.c_double),
                                                A.ptr(gl, C.c_double), A.ptr(ro, C.c_double)))
        return lo, gl, ro

