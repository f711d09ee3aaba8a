"""The allocation-free stepping entry point (ss_batch_step_noalloc) against the allocating one (ss_batch_step), through
the C ABI, bit for bit: two calls in a row on the same device scratch array, the second with fewer instances and
steps, each from golden states with per-instance connection tables (kind-scaled areas, some doors opened, a late
breach activation), in both numerics. The device buffers of inputs and outputs are the caller's, allocated here with
the HIP runtime the library itself uses."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceBuffer:
    """A caller-owned device buffer (hipMalloc through the library's own HIP runtime), filled from / read into numpy."""

    hip = None

    def __init__(self, host: np.ndarray):
        if DeviceBuffer.hip is None:
            DeviceBuffer.hip = C.CDLL("libamdhip64.so.7")
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        h = np.ascontiguousarray(host)
        assert self.hip.hipMemcpy(self.ptr, h.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), H2D) == 0

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _tables(sim, count, seed):
    rng = np.random.default_rng(seed)
    nc = sim.connections
    en = np.tile(np.asarray(sim.cEn, np.uint8), (count, 1))
    area = np.tile(np.asarray(sim.cArea, np.float64), (count, 1))
    tOn = np.tile(np.asarray(sim.cTOn, np.float64), (count, 1))
    kinds = np.asarray(sim.cKind)
    for i in range(count):
        for k in np.unique(kinds):
            area[i, kinds == k] *= np.exp(rng.uniform(np.log(0.5), np.log(2.0)))
        opened = rng.random(nc) < 0.2
        en[i, opened] = 1
        tOn[i, rng.integers(0, nc)] = 600.0
    return en, area, tOn


def _reference_step(ship, sim, recs, en, area, tOn, steps, mixed):
    from sinksim_amd.batch import Batch, BatchOptions
    nn = ship.nodes
    b = Batch(ship, sim, BatchOptions(count=len(recs), mixed_precision=mixed, traced_instances=0, t_max=5 * 3600.0))
    for i, r in enumerate(recs):
        b.restore(i, r[:7], *[r[10 + k * nn:10 + (k + 1) * nn] for k in range(5)])
    b.set_connections(en, area, tOn)
    b.step(steps)
    out = b.state_records()
    b.close()
    return out


@pytest.mark.parametrize("mixed", [False, True])
def test_noalloc_equals_step_on_two_calls(titanic, goldens, mixed):
    from sinksim_amd import _capi as A
    from sinksim_amd.batch import Batch, BatchOptions
    ship, sim = titanic
    gold = goldens["portable32" if mixed else "portable"]
    lib = A.load_library()
    # the handle whose model (and numerics, launch shape) the allocation-free calls use
    h = Batch(ship, sim, BatchOptions(count=1, mixed_precision=mixed, traced_instances=0, t_max=5 * 3600.0))
    for call, (count, steps, first) in enumerate(((48, 30, 0), (13, 7, 100))):
        recs = np.concatenate([gold.dense[first:first + count // 2], gold.samples[first // 10:first // 10 + count - count // 2]])
        en, area, tOn = _tables(sim, count, seed=call + (7 if mixed else 0))
        want = _reference_step(ship, sim, recs, en, area, tOn, steps, mixed)
        rin, ten, tar, tto = (DeviceBuffer(np.ascontiguousarray(x)) for x in (recs, en, area, tOn))
        rout = DeviceBuffer(np.full_like(recs, np.nan))
        print(f"noalloc call {call}: {count} instances x {steps} steps, mixed={mixed}", flush=True)
        st = lib.ss_batch_step_noalloc(h._h, count, steps, rin.ptr, ten.ptr, tar.ptr, tto.ptr, rout.ptr, None)
        A.check(st)
        assert DeviceBuffer.hip.hipDeviceSynchronize() == 0
        got = rout.numpy()
        bad = np.argwhere(got != want)
        assert bad.size == 0, f"call {call}: first difference at instance {bad[0][0]}, field {bad[0][1]}"
        # the input records are untouched
        assert np.array_equal(rin.numpy(), recs)
        if call == 0:
            # a count beyond the scratch array's capacity is refused on the host, before any device work
            st = lib.ss_batch_step_noalloc(h._h, 1 << 20, 1, rin.ptr, ten.ptr, tar.ptr, tto.ptr, rout.ptr, None)
            assert st == A.SS_INVALID_ARGUMENT and b"bytes of scratch" in lib.ss_last_error()
        for buf in (rin, ten, tar, tto, rout):
            buf.free()
    h.close()
