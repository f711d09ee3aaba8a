"""The C-ABI library loads, exports every entry point include/sinksim_hip.h declares, and rejects bad arguments on
the host before any device work (no compute call here: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT
from sinksim_amd import _capi as A


def _declared():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            with open(os.path.join(ROOT, "include", h)) as f:
                text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
            names += re.findall(r"\b(ss_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_exports_every_declared_entry_point():
    lib = A.load_library()
    names = _declared()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(A.SIGNATURES), "the ctypes table and the header disagree"


def test_invalid_arguments_are_rejected_on_the_host(titanic):
    lib = A.load_library()
    ship, sim = titanic
    s, m = A.ShipStruct(ship), A.SimStruct(sim)
    h = C.c_void_p()
    opt = A.SsBatchOptions(0, A.SS_NUMERICS_PORTABLE, 0, 0, 60.0, 0.0, 3600.0, 0)
    assert lib.ss_batch_create(C.byref(s.s), C.byref(m.s), C.byref(opt), C.byref(h)) == A.SS_INVALID_ARGUMENT
    assert b"count must be at least 1" in lib.ss_last_error()
    opt.count, opt.numerics = 4, 0
    assert lib.ss_batch_create(C.byref(s.s), C.byref(m.s), C.byref(opt), C.byref(h)) == A.SS_INVALID_ARGUMENT
    assert b"CPU-only" in lib.ss_last_error()
    opt.numerics = A.SS_NUMERICS_PORTABLE
    assert lib.ss_batch_create(None, C.byref(m.s), C.byref(opt), C.byref(h)) == A.SS_INVALID_ARGUMENT
    assert b"null ship" in lib.ss_last_error()
    bad = A.SimStruct(sim)
    bad.s.nnode = 63
    assert lib.ss_batch_create(C.byref(s.s), C.byref(bad.s), C.byref(opt), C.byref(h)) == A.SS_INVALID_ARGUMENT
    assert b"disagree on the node count" in lib.ss_last_error()
    assert not h.value
    # null handles are refused, and destroying a null handle is a no-op
    assert lib.ss_batch_run(None, None) == A.SS_INVALID_ARGUMENT
    lib.ss_batch_destroy(None)


def test_python_errors_map_the_status_classes():
    with pytest.raises(ValueError):
        A.check(A.SS_INVALID_ARGUMENT)
    with pytest.raises(IndexError):
        A.check(A.SS_OUT_OF_RANGE)
    with pytest.raises(A.EngineError):
        A.check(A.SS_LOGIC)
