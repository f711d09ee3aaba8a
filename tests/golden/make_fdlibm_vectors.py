"""Writes tests/golden/fdlibm_vectors.npz: inputs and the expected outputs of netlib fdlibm 5.3 (sin, cos, pow,
scalbn), computed by the reference's vendored fdlibm sources compiled unchanged (oracle/_ref/libfdlibm_ref.so,
built by `make -C oracle ref`).

The sample recipe is the reference's tests/test_portable_math.cpp:40-90 (SplitMix64 seed 20261008; sin/cos on
[-1.5, 1.5], [-3, 3], [-100, 100], [-8e5, 8e5] and special values; pow on the weir exponents 1.5 and 0.385 over
(0, 1], general positive bases, wide exponents, negative bases with integer exponents, extreme magnitudes and a
grid of special values; scalbn on [-10, 10] x [-1100, 1100]), thinned to a compact fixture.

    python tests/golden/make_fdlibm_vectors.py [--per-domain 1000]
"""
import argparse
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libfdlibm_ref.so")


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def range(self, lo, hi):
        return lo + (hi - lo) * self.uniform()


def samples(per_domain: int):
    rng = SplitMix64(20261008)
    trig = []
    for i in range(4 * per_domain):
        lo, hi = [(-1.5, 1.5), (-3.0, 3.0), (-100.0, 100.0), (-8.0e5, 8.0e5)][i % 4]
        trig.append(rng.range(lo, hi))
    trig += [0.0, -0.0, 1e-30, -1e-30, 0.785398163397448279, 1.5707963267948966, 3.141592653589793, 1e-10]
    px, py = [], []
    for i in range(6 * per_domain):
        k = i % 6
        if k == 0:
            x, y = rng.uniform(), 1.5
        elif k == 1:
            x, y = rng.uniform(), 0.385
        elif k == 2:
            x, y = rng.range(1e-6, 1e3), rng.range(-3.0, 3.0)
        elif k == 3:
            x, y = rng.range(0.5, 2.0), rng.range(-200.0, 200.0)
        elif k == 4:
            x, y = -rng.range(0.1, 10.0), float(int(rng.range(-6.0, 6.0)))
        else:
            x, y = rng.range(1e-300, 1e300), rng.range(-1.0, 1.0)
        px.append(x)
        py.append(y)
    for x in (0.0, -0.0, 1.0, -1.0, 2.0, 0.5, 1e-310, 1e308, -3.0):
        for y in (0.0, 1.0, -1.0, 2.0, 0.5, 3.0, -2.0, 0.385, 1.5, 1e10, -1e10, 1024.0, -1075.0):
            px.append(x)
            py.append(y)
    sx, sn = [], []
    for _ in range(per_domain):
        sx.append(rng.range(-10.0, 10.0))
        sn.append(int(rng.range(-1100.0, 1100.0)))
    return np.array(trig), np.array(px), np.array(py), np.array(sx), np.array(sn, dtype=np.int32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--per-domain", type=int, default=1000)
    args = ap.parse_args()
    L = C.CDLL(REF_LIB)
    for n in ("fdlibm_sin", "fdlibm_cos"):
        getattr(L, n).restype = C.c_double
        getattr(L, n).argtypes = [C.c_double]
    L.fdlibm_pow.restype = C.c_double
    L.fdlibm_pow.argtypes = [C.c_double, C.c_double]
    L.fdlibm_scalbn.restype = C.c_double
    L.fdlibm_scalbn.argtypes = [C.c_double, C.c_int]
    trig, px, py, sx, sn = samples(args.per_domain)
    out = {
        "trig_x": trig,
        "sin": np.array([L.fdlibm_sin(float(x)) for x in trig]),
        "cos": np.array([L.fdlibm_cos(float(x)) for x in trig]),
        "pow_x": px, "pow_y": py,
        "pow": np.array([L.fdlibm_pow(float(x), float(y)) for x, y in zip(px, py)]),
        "scalbn_x": sx, "scalbn_n": sn,
        "scalbn": np.array([L.fdlibm_scalbn(float(x), int(n)) for x, n in zip(sx, sn)]),
    }
    path = os.path.join(HERE, "fdlibm_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {trig.size} sin/cos, {px.size} pow, {sx.size} scalbn samples")


if __name__ == "__main__":
    main()
