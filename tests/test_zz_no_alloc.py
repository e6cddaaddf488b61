"""The allocation-free CVA entry point (eqhip_unilateral_cva_from_exposure_batch_noalloc) against the
allocating one (eqhip_unilateral_cva_from_exposure_batch) on the same profiles: two calls in a row, the
second smaller (the library's fixed weight array keeps the first call's contents), on the context's
stream and on a caller stream, and the capacity limit and host memory refused before anything is
launched. The test, as the caller, owns its device buffers: it allocates them through the HIP runtime the
library itself runs on (torch would bring a second runtime of its own into the process)."""
import ctypes

import numpy as np
import pytest

from engine_quant_amd import eqhip

pytestmark = pytest.mark.gpu

MODEL = (0.10, 0.03, 0.01, 0.02)
MAX_DATES = 16384
H2D, D2H = 1, 2


def _hip():
    eqhip.lib()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert len(paths) == 1, paths
    hip = ctypes.CDLL(paths.pop())
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


class DeviceArray:
    def __init__(self, hip, n):
        self.hip, self.n, self.ptr = hip, n, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(8, 8 * n)) == 0

    def put(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        assert x.size == self.n and self.hip.hipMemcpy(self.ptr, x.ctypes.data, 8 * self.n, H2D) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, D2H) == 0
        return out

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        self.hip.hipFree(self.ptr)


def test_noalloc_cva_equals_the_allocating_entry_point(ctx):
    hip = _hip()
    rng = np.random.default_rng(3)
    bufs = []
    for n_profiles, n_dates, hazard, recovery in ((300, 360, 0.02, 0.4), (17, 40, 0.035, 0.25)):
        print(f"no-alloc CVA: {n_profiles} profiles x {n_dates} dates", flush=True)
        times = np.arange(1, n_dates + 1) / 12.0
        ee = rng.uniform(0.0, 1e6, (n_profiles, n_dates))
        ref = ctx.unilateral_cva_batch(*MODEL, times, ee, hazard, recovery)
        d_t, d_e, d_c = DeviceArray(hip, n_dates), DeviceArray(hip, ee.size), DeviceArray(hip, n_profiles)
        bufs += [d_t, d_e, d_c]
        d_t.put(times)
        d_e.put(ee)
        d_c.put(np.full(n_profiles, np.nan))
        ctx.unilateral_cva_batch_device(*MODEL, d_t.addr, d_e.addr, n_profiles, n_dates, hazard, recovery, d_c.addr)
        assert np.array_equal(d_c.get(), ref)
    # on a caller stream the call only enqueues; the caller orders and waits
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    d_c.put(np.full(n_profiles, np.nan))
    ctx.unilateral_cva_batch_device(*MODEL, d_t.addr, d_e.addr, n_profiles, n_dates, hazard, recovery, d_c.addr,
                                    stream=stream.value)
    assert hip.hipStreamSynchronize(stream) == 0
    assert np.array_equal(d_c.get(), ref)
    hip.hipStreamDestroy(stream)
    for b in bufs:
        b.free()


def test_noalloc_cva_refuses_what_it_cannot_hold(ctx):
    hip = _hip()
    d_t, d_e, d_c = DeviceArray(hip, MAX_DATES + 1), DeviceArray(hip, MAX_DATES + 1), DeviceArray(hip, 1)
    with pytest.raises(eqhip.EqhipError) as e:
        ctx.unilateral_cva_batch_device(*MODEL, d_t.addr, d_e.addr, 1, MAX_DATES + 1, 0.02, 0.4, d_c.addr)
    assert e.value.code == eqhip.E_CAPACITY
    host = np.zeros(4)  # host memory is refused, never handed to a kernel
    with pytest.raises(eqhip.EqhipError) as e:
        ctx.unilateral_cva_batch_device(*MODEL, host.ctypes.data, d_e.addr, 1, 4, 0.02, 0.4, d_c.addr)
    assert e.value.code == eqhip.E_INVALID
    for b in (d_t, d_e, d_c):
        b.free()
