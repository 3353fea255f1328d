"""uhip_spmv_dot_noalloc: y = A x and x.y with the partial sums in the library's
file-scope scratch (no allocation), against uhip_spmv_dot through the C ABI,
bit for bit: two calls in a row, the second on a smaller matrix, for every plan
format and both precisions; oversized plans are rejected before any launch.
"""
import ctypes

import numpy as np
import pytest

import utopia_amd as ua

pytestmark = pytest.mark.gpu


def _pair(ctx, A, dtype, fmt, seed):
    D = ua.DeviceCsr(ctx, A, dtype=dtype, fmt=fmt)
    x = ua.gen_uniform(seed, A.shape[0]).astype(dtype)
    xd = ua.DeviceArray.from_numpy(ctx, x)
    ys = [ua.DeviceArray(ctx, A.shape[0], dtype) for _ in range(2)]
    dots = ua.DeviceArray(ctx, 2)
    L = ua.lib()
    ua.check(L.uhip_spmv_dot(ctx.h, D.ref, ctypes.c_void_p(xd.ptr), ctypes.c_void_p(ys[0].ptr),
                             ctypes.c_void_p(dots.ptr)), "uhip_spmv_dot")
    ua.check(L.uhip_spmv_dot_noalloc(ctx.h, D.ref, ctypes.c_void_p(xd.ptr), ctypes.c_void_p(ys[1].ptr),
                                     ctypes.c_void_p(dots.ptr + 8)), "uhip_spmv_dot_noalloc")
    d = dots.numpy()
    return ys[0].numpy(), ys[1].numpy(), d[0], d[1]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("fmt", [ua.FMT_AUTO, ua.FMT_ROWBLOCK, ua.FMT_SELL64, ua.FMT_DIA])
def test_spmv_dot_noalloc_equals_spmv_dot(ctx, dtype, fmt):
    big = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 64, 64, 64, seed=4)
    small = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 20, 18, 16, seed=5)
    for A, seed in ((big, 11), (small, 12)):  # second call: a smaller input
        y0, y1, d0, d1 = _pair(ctx, A, dtype, fmt, seed)
        assert np.array_equal(y0, y1)
        assert d0 == d1
        x = ua.gen_uniform(seed, A.shape[0]).astype(dtype).astype(np.float64)
        ref = float(np.dot(x, A.to_scipy() @ x))
        assert abs(d0 - ref) <= (1e-5 if dtype == np.float32 else 1e-12) * abs(ref)


def test_spmv_dot_noalloc_rejects_oversized_plan(ctx):
    A = ua.gen_stencil(ua.STENCIL_2D5, 8, 8)
    D = ua.DeviceCsr(ctx, A, fmt=ua.FMT_ROWBLOCK)
    m = ua.uhip_csr.from_buffer_copy(D.m)
    m.nblocks = (1 << 18) + 1  # checked before anything is launched
    m.nparts = m.nblocks
    xd, yd, dd = ua.DeviceArray(ctx, 64), ua.DeviceArray(ctx, 64), ua.DeviceArray(ctx, 1)
    st = ua.lib().uhip_spmv_dot_noalloc(ctx.h, ctypes.byref(m), ctypes.c_void_p(xd.ptr), ctypes.c_void_p(yd.ptr),
                                        ctypes.c_void_p(dd.ptr))
    assert st == -1  # UHIP_ERR_INVALID
    assert b"scratch holds" in ua.lib().uhip_last_error()
