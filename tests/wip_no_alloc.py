"""qx_rfft3_dev_noalloc (the rfft3 drop-in on device arrays with its work
grid in the library's fixed device array) against qx_rfft3 on the same
inputs: two calls in a row, the second on a smaller grid, in both directions
(r2c x1/N, then c2r of that half grid).  Both run the same kernels with the
same twiddles, so the results must agree bit for bit."""
import ctypes as C

import numpy as np
import pytest

from qxmd_amd import lib, pw


def _dev_rfft3(key, x, y, nd):
    import torch
    dx = torch.from_numpy(np.ascontiguousarray(x, np.float64)).cuda()
    dy = torch.from_numpy(np.ascontiguousarray(y, np.float64)).cuda()
    torch.cuda.synchronize()
    ierr = C.c_int(-1)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.load().qx_rfft3_dev_noalloc(int(key), C.c_void_p(dx.data_ptr()),
                                         C.c_void_p(dy.data_ptr()), *[int(v) for v in nd],
                                         stream, C.byref(ierr))
    assert rc == 0 and ierr.value == 0, lib.load().qx_last_error()
    return dx.cpu().numpy(), dy.cpu().numpy()


@pytest.mark.gpu
def test_rfft3_noalloc_matches_rfft3():
    rng = np.random.default_rng(20261203)
    for nd in [(60, 60, 60), (30, 24, 20)]:  # second call: a smaller grid
        N = int(np.prod(nd))
        print("rfft3 noalloc", nd, flush=True)
        x = rng.standard_normal(N)
        fx, fy = _dev_rfft3(1, x, np.zeros(N), nd)
        rx, ry = pw.rfft3(1, x, np.zeros(N), nd)
        assert np.array_equal(fx, rx) and np.array_equal(fy, ry), nd
        bx, by = _dev_rfft3(2, rx, ry, nd)
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
