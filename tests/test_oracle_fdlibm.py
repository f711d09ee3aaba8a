"""The oracle's fdlibm restatement against netlib fdlibm 5.3, bit for bit.

Always: the committed fixture (tests/golden/fdlibm_vectors.npz, expected values from the reference's vendored
fdlibm sources compiled unchanged). Where the reference build is present (oracle/_ref/libfdlibm_ref.so): the
reference's full sample recipe (tests/test_portable_math.cpp:40-105, 400,000 sin/cos and pow samples)."""
import ctypes as C
import os

import numpy as np
import pytest

import pyoracle
from conftest import GOLDEN


def _bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def _same(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return (_bits(a) == _bits(b)) | (np.isnan(a) & np.isnan(b))


def test_fixture_vectors():
    v = np.load(os.path.join(GOLDEN, "fdlibm_vectors.npz"))
    sin = np.array([pyoracle.fdl_sin(float(x)) for x in v["trig_x"]])
    cos = np.array([pyoracle.fdl_cos(float(x)) for x in v["trig_x"]])
    pw = np.array([pyoracle.fdl_pow(float(x), float(y)) for x, y in zip(v["pow_x"], v["pow_y"])])
    sc = np.array([pyoracle.fdl_scalbn(float(x), int(n)) for x, n in zip(v["scalbn_x"], v["scalbn_n"])])
    assert _same(sin, v["sin"]).all()
    assert _same(cos, v["cos"]).all()
    assert _same(pw, v["pow"]).all()
    assert _same(sc, v["scalbn"]).all()


def test_known_values():
    """Closed forms fdlibm must meet: sin 0 = 0 with its sign, cos 0 = 1, pow special cases, exact squares."""
    assert _bits(pyoracle.fdl_sin(-0.0)) == _bits(-0.0)
    assert pyoracle.fdl_cos(0.0) == 1.0
    assert pyoracle.fdl_pow(2.0, 10.0) == 1024.0
    assert pyoracle.fdl_pow(0.25, 1.5) == 0.125
    assert pyoracle.fdl_pow(4.0, 0.5) == 2.0
    assert pyoracle.fdl_pow(-2.0, 3.0) == -8.0
    assert np.isnan(pyoracle.fdl_pow(-2.0, 0.5))
    assert pyoracle.fdl_pow(7.0, 0.0) == 1.0
    assert pyoracle.fdl_scalbn(1.0, -1074) == 5e-324
    x = 0.7
    assert abs(pyoracle.fdl_sin(x) ** 2 + pyoracle.fdl_cos(x) ** 2 - 1.0) < 2e-16


REF_LIB = os.path.join(pyoracle.REF_DIR, "libfdlibm_ref.so")


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference fdlibm build absent (make -C oracle ref)")
def test_full_recipe_against_netlib_build():
    import make_fdlibm_vectors as M  # tests/golden
    L = C.CDLL(REF_LIB)
    for n in ("fdlibm_sin", "fdlibm_cos"):
        getattr(L, n).restype = C.c_double
        getattr(L, n).argtypes = [C.c_double]
    L.fdlibm_pow.restype = C.c_double
    L.fdlibm_pow.argtypes = [C.c_double, C.c_double]
    trig, px, py, sx, sn = M.samples(100000)
    s_ref = np.array([L.fdlibm_sin(float(x)) for x in trig])
    c_ref = np.array([L.fdlibm_cos(float(x)) for x in trig])
    p_ref = np.array([L.fdlibm_pow(float(x), float(y)) for x, y in zip(px, py)])
    assert _same([pyoracle.fdl_sin(float(x)) for x in trig], s_ref).all()
    assert _same([pyoracle.fdl_cos(float(x)) for x in trig], c_ref).all()
    assert _same([pyoracle.fdl_pow(float(x), float(y)) for x, y in zip(px, py)], p_ref).all()


@pytest.fixture(autouse=True, scope="module")
def _golden_on_path():
    import sys
    sys.path.insert(0, GOLDEN)
    yield
    sys.path.remove(GOLDEN)
