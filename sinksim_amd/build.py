"""Build recipe of the native parts, in-tree (the built files travel with the repository to a GPU host):

  sinksim_amd/libsinksim_hip.so   the gfx950 batch and adjoint kernels and the C ABI (include/sinksim_hip.h)
  sinksim_amd/bin/sinksim_hip_run, sinksim_amd/bin/sinksim_sweep
                                  the C++ command-line drop-ins over the C ABI (host code: g++, no HIP headers)

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
LIB_SOURCES = [os.path.join(CSRC, "batch.hip"), os.path.join(CSRC, "adjoint.hip"), os.path.join(CSRC, "cpu_engine.cpp")]
LIB_DEPS = LIB_SOURCES + [os.path.join(CSRC, f) for f in ("step_dev.hpp", "fdlibm_dev.hpp", "adjoint_dev.hpp", "batch_host.hpp")] + [
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
    # every source is compiled as HIP: cpu_engine.cpp is host code that shares the device's fdlibm (host and device)
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, "-x", "hip", *LIB_SOURCES])
    os.replace(tmp, LIB)
    return LIB


HOST = os.path.join(CSRC, "host")
APPS = os.path.join(CSRC, "apps")
BIN = os.path.join(HERE, "bin")
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-Wall", "-Wextra", "-ffp-contract=off", "-fno-fast-math"]
HOST_SOURCES = [os.path.join(HOST, f) for f in ("json.cpp", "model.cpp", "runfiles.cpp")]


def build_tools(force: bool = True) -> list[str]:
    """The command-line drop-ins, linked against the in-tree libsinksim_hip.so (rpath $ORIGIN/..)."""
    os.makedirs(BIN, exist_ok=True)
    deps = HOST_SOURCES + [os.path.join(HOST, f) for f in ("json.hpp", "model.hpp", "runfiles.hpp", "batch.hpp")] + [
        os.path.join(APPS, f) for f in ("options.hpp", "runs.hpp")] + [os.path.join(ROOT, "include", "sinksim_hip.h")]
    out = []
    for app in ("sinksim_hip_run", "sinksim_sweep", "sinksim_hip_validate"):
        exe = os.path.join(BIN, app)
        src = os.path.join(APPS, app + ".cpp")
        if not force and os.path.exists(exe) and all(os.path.getmtime(exe) >= os.path.getmtime(d) for d in deps + [src, LIB]):
            out.append(exe)
            continue
        _run([CXX, *HOST_FLAGS, "-o", exe + ".tmp", src, *HOST_SOURCES, f"-L{HERE}", "-lsinksim_hip",
              "-Wl,-rpath,$ORIGIN/.."])
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def build_all(force: bool = True) -> list[str]:
    return [build_library(force), *build_tools(force)]


if __name__ == "__main__":
    for p in build_all(force="--if-stale" not in sys.argv):
        print("built", p)
