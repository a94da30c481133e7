"""mrlbm_reduce_checked_noalloc (the allocation-free entry point) against
mrlbm_reduce_checked through the C ABI: the same sums bit for bit, on two
calls in a row, the second on a smaller lattice (with non-finite nodes
planted in its moments), and a host pointer refused. The output buffer is
the caller's device memory, allocated here with hipMalloc."""
import ctypes

import numpy as np
import pytest

import mrlbm
from mrlbm import _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class DeviceBuffer:
    """Three doubles of device memory (the caller's output buffer)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(24)) == 0
        assert self.hip.hipMemset(self.ptr, 0xFF, ctypes.c_size_t(24)) == 0  # NaN pattern: must be overwritten

    def read(self):
        out = np.empty(3)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(24), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def noalloc(lat, buf):
    L = _lib.load()
    _lib.check(L.mrlbm_reduce_checked_noalloc(lat._h, buf.ptr))
    _lib.check(L.mrlbm_synchronize(lat._h))
    return buf.read()


@pytest.mark.parametrize("storage", [mrlbm.STORE_FP32, mrlbm.STORE_FP16])
def test_noalloc_reduce_equals_reduce_checked(storage):
    buf = DeviceBuffer()
    try:
        say(f"no-alloc reduce: TGV 64^3 storage {storage}, 50 steps")
        with mrlbm.Lattice((64, 64, 64), q=19, tau=0.53, storage=storage) as lat:
            lat.init_taylor_green(0.05)
            lat.step(50)
            ref = lat.reduce_checked()
            got = noalloc(lat, buf)
        say("reduce_checked", ref, "noalloc", got)
        assert got[0] == ref[0] and got[1] == ref[1] and got[2] == ref[2] == 0

        say("no-alloc reduce: 16x16x8 with 3 non-finite nodes, second call")
        with mrlbm.Lattice((16, 16, 8), q=19, tau=0.6, storage=storage) as lat:
            lat.init_taylor_green(0.02)
            lat.step(10)
            m = lat.moments()
            m[1, 5] = np.nan
            m[0, 77] = np.inf
            m[3, 300] = np.nan
            lat.write_moments(m)
            ref = lat.reduce_checked()
            got = noalloc(lat, buf)
        say("reduce_checked", ref, "noalloc", got)
        assert ref[2] == 3 and got[2] == 3
        assert np.isnan(got[0]) == np.isnan(ref[0]) and np.isnan(got[1]) == np.isnan(ref[1])
        assert (got[0] == ref[0] or np.isnan(ref[0])) and (got[1] == ref[1] or np.isnan(ref[1]))
    finally:
        buf.free()


def test_noalloc_reduce_refuses_host_output():
    with mrlbm.Lattice((16, 16, 16), q=19, tau=0.6) as lat:
        lat.init_uniform()
        host = (ctypes.c_double * 3)()
        rc = _lib.load().mrlbm_reduce_checked_noalloc(lat._h, ctypes.cast(host, ctypes.c_void_p))
        assert rc == -1  # MRLBM_EINVAL
        assert b"device memory" in _lib.load().mrlbm_last_error()
        assert lat.reduce_checked()[2] == 0  # the context still works
