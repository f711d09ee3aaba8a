"""Build recipe of the device-arithmetic probe (tests/probe/libsinksim_probe.so): probe.hip compiled with the engine's
own HIP flags and headers, so that the arithmetic it evaluates is the arithmetic the step compiles to. A test hook
(tests/test_gpu_math.py) and the plain copy kernel that calibrates a profiler's byte counters; not part of the
engine's C ABI."""
from __future__ import annotations

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PROBE = os.path.join(HERE, "libsinksim_probe.so")


def build_probe(force: bool = True) -> str:
    sys.path.insert(0, ROOT)
    from sinksim_amd import build as B

    src = os.path.join(HERE, "probe.hip")
    deps = [src] + [os.path.join(B.CSRC, f) for f in ("step_dev.hpp", "fdlibm_dev.hpp")]
    if not force and os.path.exists(PROBE) and all(os.path.getmtime(PROBE) >= os.path.getmtime(d) for d in deps):
        return PROBE
    B._run([B.HIPCC, *B.HIP_FLAGS, f"-I{B.CSRC}", "-shared", "-o", PROBE + ".tmp", src])
    os.replace(PROBE + ".tmp", PROBE)
    return PROBE


if __name__ == "__main__":
    print("built", build_probe(force="--if-stale" not in sys.argv))
