"""ctypes binding of the C ABI in include/sinksim_hip.h (libsinksim_hip.so, built in-tree next to this file).

Loading fails loudly when the library has not been built: there is no CPU fallback for the batch engine.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsinksim_hip.so")

SS_OK, SS_INVALID_ARGUMENT, SS_RUNTIME, SS_LOGIC, SS_OUT_OF_RANGE = 0, 1, 2, 3, 4
SS_NUMERICS_ORACLE, SS_NUMERICS_PORTABLE, SS_NUMERICS_PORTABLE32, SS_NUMERICS_STD = 0, 1, 2, 3
NUMERICS = {"oracle": SS_NUMERICS_ORACLE, "portable": SS_NUMERICS_PORTABLE, "portable32": SS_NUMERICS_PORTABLE32, "std": SS_NUMERICS_STD}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_uint8)


class SsShip(C.Structure):
    _fields_ = [("ncol", C.c_int32), ("x", _dp), ("y", _dp), ("dxdy", _dp), ("zlo", _dp), ("zhi", _dp),
                ("nseg", C.c_int32), ("segCol", _ip), ("segA", _dp), ("segE", _dp),
                ("nnode", C.c_int32), ("mu", _dp), ("vmax", _dp), ("amin", _dp), ("aov", _dp),
                ("segStart", _ip), ("segCount", _ip)]


class SsSim(C.Structure):
    _fields_ = [("nnode", C.c_int32), ("rho", C.c_double), ("g", C.c_double), ("dt", C.c_double),
                ("ngroup", C.c_int32), ("gNodeStart", _ip), ("gNodeCount", _ip), ("gNodes", _ip), ("gMode", _bp),
                ("gpx", _dp), ("gpy", _dp), ("gpz", _dp), ("gMargin", C.c_double),
                ("nconn", C.c_int32), ("ca", _ip), ("cb", _ip), ("law", _bp), ("cx", _dp), ("cy", _dp), ("cz", _dp),
                ("area", _dp), ("coef", _dp), ("kind", _bp), ("idx", _ip), ("en", _bp), ("tOn", _dp),
                ("skipGroup", _ip), ("monitor", _ip),
                ("nmonitor", C.c_int32), ("monitorThreshold", C.c_double),
                ("nmark", C.c_int32), ("mkx", _dp), ("mky", _dp), ("mkz", _dp), ("mkWhen", _bp),
                ("founder", C.c_double * 4), ("body", C.c_double * 16), ("readout", C.c_double * 2),
                ("state0", C.c_double * 7),
                ("vol0", _dp), ("level0", _dp), ("cx0", _dp), ("cy0", _dp), ("cz0", _dp)]


class SsBatchOptions(C.Structure):
    _fields_ = [("count", C.c_int32), ("numerics", C.c_int32), ("traced", C.c_int32), ("denseSteps", C.c_int32),
                ("sampleEvery", C.c_double), ("curveEvery", C.c_double), ("tMax", C.c_double),
                ("stepsPerLaunch", C.c_int64)]


class SsDeviceProperties(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("computeUnits", C.c_int32), ("major", C.c_int32),
                ("minor", C.c_int32), ("globalMemory", C.c_size_t), ("sharedPerBlock", C.c_size_t),
                ("runtimeVersion", C.c_int32), ("driverVersion", C.c_int32)]


class SsInstanceResult(C.Structure):
    _fields_ = [("foundered", C.c_int32), ("founderT", C.c_double), ("tEnd", C.c_double), ("steps", C.c_int64),
                ("state", C.c_double * 7), ("outputs", C.c_double * 7), ("events", C.c_int32)]


class SsTuneResult(C.Structure):
    _fields_ = [("threadsPerSim", C.c_int32), ("workgroups", C.c_int32), ("rate", C.c_double * 3), ("trials", C.c_int32),
                ("trialThreads", C.c_int32 * 8), ("trialWorkgroups", C.c_int32 * 8), ("trialRate", C.c_double * 8),
                ("recordBytes", C.c_int32), ("trialRecordBytes", C.c_int32 * 8)]


class SsEvent(C.Structure):
    _fields_ = [("t", C.c_double), ("id", C.c_int32), ("kind", C.c_int32)]


class SsLoss(C.Structure):
    _fields_ = [("nRead", C.c_int32), ("rdStep", C.POINTER(C.c_int64)), ("rdKind", _ip), ("rdW", _dp), ("rdAfter", _dp),
                ("nEvent", C.c_int32), ("evKind", _ip), ("evIdx", _ip), ("evW", _dp), ("evMissing", _dp)]


class SsGradientOptions(C.Structure):
    _fields_ = [("checkpointEvery", C.c_int32), ("relaxed", C.c_int32), ("tMax", C.c_double), ("maxSteps", C.c_int64)]


class SsCpuOptions(C.Structure):
    _fields_ = [("numerics", C.c_int32), ("threads", C.c_int32), ("tMax", C.c_double), ("curveEvery", C.c_double),
                ("stopWhenStable", C.c_int32)]


# every exported entry point and its signature (the test of the library's exports reads this table)
SIGNATURES = {
    "ss_last_error": (C.c_char_p, []),
    "ss_available": (C.c_int, [C.c_char_p, C.c_size_t]),
    "ss_device_info": (C.c_int, [C.POINTER(SsDeviceProperties)]),
    "ss_batch_create": (C.c_int, [C.POINTER(SsShip), C.POINTER(SsSim), C.POINTER(SsBatchOptions), C.POINTER(C.c_void_p)]),
    "ss_batch_create_fleet": (C.c_int, [C.POINTER(SsShip), C.POINTER(SsSim), C.c_int32, _ip, C.POINTER(SsBatchOptions),
                                        C.POINTER(C.c_void_p)]),
    "ss_batch_destroy": (None, [C.c_void_p]),
    "ss_batch_restore": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp]),
    "ss_batch_set_connection_enabled": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "ss_batch_set_connection_area": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "ss_batch_set_connection_activation": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "ss_batch_set_connections": (C.c_int, [C.c_void_p, _bp, _dp, _dp]),
    "ss_batch_set_launch": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "ss_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.POINTER(SsTuneResult)]),
    "ss_batch_set_record_layout": (C.c_int, [C.c_void_p, C.c_int32]),
    "ss_batch_run": (C.c_int, [C.c_void_p, _dp]),
    "ss_batch_step": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64), _dp]),
    "ss_batch_state_records": (C.c_int, [C.c_void_p, _dp]),
    "ss_batch_step_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]),
    "ss_batch_total_steps": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "ss_batch_shared_memory_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "ss_batch_result": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(SsInstanceResult), _dp, _dp, _dp, _dp, _dp,
                                  C.POINTER(SsEvent), C.c_int32]),
    "ss_batch_results": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), _ip, _dp, _dp]),
    "ss_batch_trace": (C.c_int, [C.c_void_p, C.c_int32, _dp, C.c_int32, _ip, _dp, C.c_int32, _ip]),
    "ss_batch_curve": (C.c_int, [C.c_void_p, C.c_int32, _dp, C.c_int32, _ip]),
    "ss_batch_final_readouts": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "ss_batch_kernel_time": (C.c_int, [C.c_void_p, _dp, _ip]),
    "ss_batch_gradient": (C.c_int, [C.c_void_p, C.c_int32, _ip, _ip, _dp, C.POINTER(SsLoss), _dp, C.POINTER(SsGradientOptions), _dp,
                                    _dp, _dp, C.POINTER(C.c_int64), _ip, _dp]),
    "ss_batch_adjoint_step": (C.c_int, [C.c_void_p, C.c_int32, _ip, _ip, _dp, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "ss_batch_gradient_launch": (C.c_int, [C.c_void_p, _ip, _ip, C.POINTER(C.c_int64)]),
    "ss_batch_set_stop_when_stable": (C.c_int, [C.c_void_p, C.c_int32]),
    "ss_cpu_sweep": (C.c_int, [C.POINTER(SsShip), C.POINTER(SsSim), C.POINTER(SsCpuOptions), C.c_int32, _bp, _dp, _dp,
                               C.POINTER(SsInstanceResult), C.POINTER(SsEvent), C.c_int32, _dp, C.c_int32, _ip, _dp, _dp]),
    "ss_cpu_trace": (C.c_int, [C.POINTER(SsShip), C.POINTER(SsSim), C.c_int32, C.c_double, C.c_int32, C.c_double, _bp, _dp, _dp,
                               _dp, C.c_int32, _ip, _dp, C.c_int32, _ip, C.POINTER(SsInstanceResult), C.POINTER(SsEvent), C.c_int32]),
}

_lib = None


def load_library() -> C.CDLL:
    """The in-tree libsinksim_hip.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class EngineError(RuntimeError):
    """A failed C-ABI call; `status` is the ss_status class."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def check(status: int) -> None:
    if status != SS_OK:
        msg = load_library().ss_last_error().decode("utf-8", "replace")
        if status == SS_INVALID_ARGUMENT:
            raise ValueError(msg)
        if status == SS_OUT_OF_RANGE:
            raise IndexError(msg)
        raise EngineError(status, msg)


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def u8(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint8)


class ShipStruct:
    """An ss_ship over contiguous copies of a CompiledShip's arrays (kept alive with the struct)."""

    def __init__(self, ship):
        self.keep = [f64(ship.x), f64(ship.y), f64(ship.dxdy), f64(ship.zlo), f64(ship.zhi), i32(ship.segCol),
                     f64(ship.segA), f64(ship.segE), f64(ship.mu), f64(ship.vmax), f64(ship.amin), f64(ship.aov),
                     i32(ship.segStart), i32(ship.segCount)]
        x, y, dxdy, zlo, zhi, col, a, e, mu, vmax, amin, aov, s0, sc = self.keep
        self.s = SsShip(ship.columns, ptr(x, C.c_double), ptr(y, C.c_double), ptr(dxdy, C.c_double),
                        ptr(zlo, C.c_double), ptr(zhi, C.c_double), ship.segments, ptr(col, C.c_int32),
                        ptr(a, C.c_double), ptr(e, C.c_double), ship.nodes, ptr(mu, C.c_double),
                        ptr(vmax, C.c_double), ptr(amin, C.c_double), ptr(aov, C.c_double), ptr(s0, C.c_int32),
                        ptr(sc, C.c_int32))


class SimStruct:
    """An ss_sim over contiguous copies of a CompiledSim's arrays."""

    def __init__(self, sim):
        k = {}

        def keep(name, arr):
            k[name] = arr
            return arr

        s = SsSim()
        s.nnode = sim.nodes
        s.rho, s.g, s.dt = sim.rho, sim.g, sim.dt
        s.ngroup = sim.groups
        s.gNodeStart = ptr(keep("gs", i32(sim.gNodeStart)), C.c_int32)
        s.gNodeCount = ptr(keep("gc", i32(sim.gNodeCount)), C.c_int32)
        s.gNodes = ptr(keep("gn", i32(sim.gNodes)), C.c_int32)
        s.gMode = ptr(keep("gm", u8(sim.gMode)), C.c_uint8)
        s.gpx = ptr(keep("gpx", f64(sim.gpx)), C.c_double)
        s.gpy = ptr(keep("gpy", f64(sim.gpy)), C.c_double)
        s.gpz = ptr(keep("gpz", f64(sim.gpz)), C.c_double)
        s.gMargin = sim.gMargin
        s.nconn = sim.connections
        s.ca = ptr(keep("ca", i32(sim.ca)), C.c_int32)
        s.cb = ptr(keep("cb", i32(sim.cb)), C.c_int32)
        s.law = ptr(keep("law", u8(sim.cLaw)), C.c_uint8)
        s.cx = ptr(keep("cx", f64(sim.cx)), C.c_double)
        s.cy = ptr(keep("cy", f64(sim.cy)), C.c_double)
        s.cz = ptr(keep("cz", f64(sim.cz)), C.c_double)
        s.area = ptr(keep("area", f64(sim.cArea)), C.c_double)
        s.coef = ptr(keep("coef", f64(sim.cCoef)), C.c_double)
        s.kind = ptr(keep("kind", u8(sim.cKind)), C.c_uint8)
        s.idx = ptr(keep("idx", i32(sim.cIdx)), C.c_int32)
        s.en = ptr(keep("en", u8(sim.cEn)), C.c_uint8)
        s.tOn = ptr(keep("tOn", f64(sim.cTOn)), C.c_double)
        s.skipGroup = ptr(keep("skip", i32(sim.cSkipGroup)), C.c_int32)
        s.monitor = ptr(keep("mon", i32(sim.cMonitor)), C.c_int32)
        s.nmonitor = sim.monitors
        s.monitorThreshold = sim.monitorThreshold
        s.nmark = sim.marks
        s.mkx = ptr(keep("mkx", f64(sim.mkx)), C.c_double)
        s.mky = ptr(keep("mky", f64(sim.mky)), C.c_double)
        s.mkz = ptr(keep("mkz", f64(sim.mkz)), C.c_double)
        s.mkWhen = ptr(keep("mkw", u8(sim.mkWhen)), C.c_uint8)
        s.founder[:] = list(sim.founder)
        s.body[:] = list(sim.body)
        s.readout[:] = list(sim.readout)
        s.state0[:] = list(sim.state0)
        s.vol0 = ptr(keep("vol0", f64(sim.vol0)), C.c_double)
        s.level0 = ptr(keep("level0", f64(sim.level0)), C.c_double)
        s.cx0 = ptr(keep("cx0", f64(sim.cx0)), C.c_double)
        s.cy0 = ptr(keep("cy0", f64(sim.cy0)), C.c_double)
        s.cz0 = ptr(keep("cz0", f64(sim.cz0)), C.c_double)
        self.keep = k
        self.s = s
