"""Build recipe of the native parts, in-tree (the built files travel with the repository to a GPU host):

  sinksim_amd/libsinksim_hip.so   the gfx950 batch kernel and the C ABI (include/sinksim_hip.h)

Every HIP translation unit is compiled without contraction (-ffp-contract=off: each product and sum rounds on its
own, as the reference's strict floating-point target requires), with IEEE single-precision denormals and correctly
rounded single-precision division, and without packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}",
]

LIB = os.path.join(HERE, "libsinksim_hip.so")
LIB_SOURCES = [os.path.join(CSRC, "batch.hip")]
LIB_DEPS = LIB_SOURCES + [os.path.join(CSRC, f) for f in ("step_dev.hpp", "fdlibm_dev.hpp")] + [
    os.path.join(ROOT, "include", "sinksim_hip.h")]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature; that is expected
    err = "\n".join(l for l in proc.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"build failed ({proc.returncode}): {' '.join(cmd)}\n{err}")
    if err:
        sys.stderr.write(err + "\n")


def build_library(force: bool = True) -> str:
    if not force and os.path.exists(LIB) and all(os.path.getmtime(LIB) >= os.path.getmtime(d) for d in LIB_DEPS):
        return LIB
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *LIB_SOURCES])
    os.replace(tmp, LIB)
    return LIB


def build_all(force: bool = True) -> list[str]:
    return [build_library(force)]


if __name__ == "__main__":
    for p in build_all(force="--if-stale" not in sys.argv):
        print("built", p)
