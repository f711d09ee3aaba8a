"""The device arithmetic under the step, bit for bit against the CPU: the device fdlibm (sin, cos, pow, scalbn)
against the committed netlib vectors and against the oracle on the reference's sample recipe, f64 division and
square root (IEEE correctly rounded, as x86 SSE2 computes them) and f32 division and conversion on 10^8 samples."""
import ctypes as C
import os

import numpy as np
import pytest

import pyoracle
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

PROBE = os.path.join(ROOT, "tests", "probe", "libsinksim_probe.so")


def _lib():
    L = C.CDLL(PROBE)
    L.ssp_eval.restype = C.c_int
    L.ssp_eval.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_longlong]
    L.ssp_last_error.restype = C.c_char_p
    return L


def _eval(op, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.zeros_like(a) if b is None else np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty_like(a)
    L = _lib()
    p = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    assert L.ssp_eval(op, p(a), p(b), p(out), a.size) == 0, L.ssp_last_error()
    return out


def _same(x, y):
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    return (x.view(np.uint64) == y.view(np.uint64)) | (np.isnan(x) & np.isnan(y))


def test_fdlibm_against_netlib_vectors():
    v = np.load(os.path.join(GOLDEN, "fdlibm_vectors.npz"))
    assert _same(_eval(0, v["trig_x"]), v["sin"]).all()
    assert _same(_eval(1, v["trig_x"]), v["cos"]).all()
    assert _same(_eval(2, v["pow_x"], v["pow_y"]), v["pow"]).all()
    assert _same(_eval(3, v["scalbn_x"], v["scalbn_n"].astype(np.float64)), v["scalbn"]).all()


def test_fdlibm_against_oracle_on_the_kernel_domains():
    rng = np.random.default_rng(20261008)
    x = np.concatenate([rng.uniform(-1.4, 1.4, 200000), rng.uniform(-3.0, 3.0, 50000), rng.uniform(-8e5, 8e5, 50000)])
    s, c = _eval(0, x), _eval(1, x)
    ts, tc = _eval(9, x, np.zeros_like(x)), _eval(9, x, np.ones_like(x))
    assert _same(s, ts).all() and _same(c, tc).all()
    sel = rng.choice(x.size, 20000, replace=False)
    assert _same(s[sel], [pyoracle.fdl_sin(float(t)) for t in x[sel]]).all()
    assert _same(c[sel], [pyoracle.fdl_cos(float(t)) for t in x[sel]]).all()
    r = rng.uniform(0, 1, 40000)
    y = np.where(np.arange(r.size) % 2 == 0, 1.5, 0.385)
    p = _eval(8, r, y)
    assert _same(p, [pyoracle.fdl_pow(float(a), float(b)) for a, b in zip(r, y)]).all()
    # the Villemonte factor (1 - r^1.5)^0.385, base clipped at 0, as one call with both exponents folded in
    r = np.concatenate([rng.uniform(0, 1, 20000), [0.0, 1.0, 1e-300, 0.999999999]])
    v = _eval(8, r, np.full(r.size, -1.0))
    want = [pyoracle.fdl_pow(max(0.0, 1 - pyoracle.fdl_pow(float(a), 1.5)), 0.385) for a in r]
    assert _same(v, want).all()


def test_f64_division_and_sqrt_are_correctly_rounded():
    rng = np.random.default_rng(7)
    n = 1 << 24
    for k in range(6):  # 10^8 samples over wide and the step's own ranges
        e = rng.uniform(-300, 300, n) if k % 2 else rng.uniform(-8, 8, n)
        a = rng.uniform(-1, 1, n) * 10.0 ** e
        b = rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-8, 8, n)
        assert _same(_eval(4, a, b), a / b).all(), "f64 division"
        s = np.abs(a)
        assert _same(_eval(5, s), np.sqrt(s)).all(), "f64 sqrt"


def test_f32_division_and_conversion():
    rng = np.random.default_rng(11)
    n = 1 << 24
    for _ in range(6):
        a = rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-40, 40, n)
        b = rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-30, 30, n)
        want = (a.astype(np.float32) / b.astype(np.float32)).astype(np.float64)
        assert _same(_eval(6, a, b), want).all(), "f32 division"
        assert _same(_eval(7, a), a.astype(np.float32).astype(np.float64)).all(), "f64 -> f32"
