"""qx_rfft3_dev_noalloc (the rfft3 drop-in on device arrays with its work
grid in the library's fixed device array) against qx_rfft3 on the same
inputs: two calls in a row, the second on a smaller grid, in both directions
(r2c x1/N, then c2r of that half grid).  Both run the same kernels with the
same twiddles, so the results must agree bit for bit.  The device arrays come
straight from the HIP runtime (hipMalloc / hipMemcpy through ctypes)."""
import ctypes as C

import numpy as np
import pytest

from qxmd_amd import lib, pw


class _Hip:
    """The few HIP runtime calls the test needs."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, dst, a):
        assert self.h.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # H2D

    def get(self, src, a):
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2) == 0  # D2H


def _dev_rfft3(hip, key, x, y, nd):
    n = int(np.prod(nd))
    dx, dy = hip.alloc(8 * n), hip.alloc(8 * n)
    try:
        hip.put(dx, np.ascontiguousarray(x, np.float64))
        hip.put(dy, np.ascontiguousarray(y, np.float64))
        ierr = C.c_int(-1)
        rc = lib.load().qx_rfft3_dev_noalloc(int(key), dx, dy, *[int(v) for v in nd], None,
                                             C.byref(ierr))
        assert rc == 0 and ierr.value == 0, lib.load().qx_last_error()
        ox, oy = np.empty(n), np.empty(n)
        hip.get(dx, ox)
        hip.get(dy, oy)
        return ox, oy
    finally:
        hip.h.hipFree(dx)
        hip.h.hipFree(dy)


@pytest.mark.gpu
def test_rfft3_noalloc_matches_rfft3():
    hip = _Hip()
    rng = np.random.default_rng(20261203)
    for nd in [(60, 60, 60), (30, 24, 20)]:  # second call: a smaller grid
        N = int(np.prod(nd))
        print("rfft3 noalloc", nd, flush=True)
        x = rng.standard_normal(N)
        fx, fy = _dev_rfft3(hip, 1, x, np.zeros(N), nd)
        rx, ry = pw.rfft3(1, x, np.zeros(N), nd)
        assert np.array_equal(fx, rx) and np.array_equal(fy, ry), nd
        bx, by = _dev_rfft3(hip, 2, rx, ry, nd)
        ox, oy = pw.rfft3(2, rx, ry, nd)
        assert np.array_equal(bx, ox) and np.array_equal(by, oy), nd
        assert np.abs(bx - x).max() < 1e-12 * np.abs(x).max()


def test_rfft3_noalloc_rejects_oversized_grid():
    L = lib.load()
    mp, ma = C.c_long(0), C.c_int(0)
    assert L.qx_noalloc_limits(C.byref(mp), C.byref(ma)) == 0
    assert mp.value == 128 ** 3
    ierr = C.c_int(0)
    # refused by the size check before any device call (runs without a GPU)
    assert L.qx_rfft3_dev_noalloc(1, None, None, 130, 128, 128, None, C.byref(ierr)) == -2
    assert ierr.value == 1
    assert L.qx_rfft3_dev_noalloc(1, None, None, 4096, 1, 1, None, C.byref(ierr)) == -2
