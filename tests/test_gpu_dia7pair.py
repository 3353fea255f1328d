"""The fused Chebyshev smoothing pairs on plane-walking 7-point operators
(uhip_cheb_pre_residual, uhip_cheb_residual_post) against the two separate launches they
replace, whose results are bit-exact against the oracle (test_gpu_kernels.py): every output bit
for bit, the post-smoothing dot within its rounding bound. Shapes cover one chunk per plane (2048
rows FP32, two of 1024 FP64), y halos of 64, 128 and 256 rows and of 32 and 96 (not a multiple of
the 64-lane wave: the halo rows' x neighbours are loaded, not moved across lanes), the benchmark's
65,536-row planes, and plane counts that leave a partial group of planes (32 planes per
workgroup)."""
import ctypes

import numpy as np
import pytest

import utopia_amd as ua
from utopia_amd._abi import lib, uhip_cheb_step

pytestmark = pytest.mark.gpu

P = lambda a: ctypes.c_void_p(a.ptr)  # noqa: E731
UNSUPPORTED = -5

SHAPES = [(64, 32, 13), (128, 16, 9), (64, 64, 40), (64, 32, 70), (256, 256, 21), (96, 64, 9), (32, 64, 11)]
COEF = dict(a_prev=0.41, beta=0.073, a_cur=0.29)


def _vecs(ctx, n, dt, seeds):
    return [ua.DeviceArray.from_numpy(ctx, ua.gen_uniform(s, n, -1.0, 1.0).astype(dt)) for s in seeds]


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pre_residual_equals_step_then_residual(ctx, dims, dt):
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, *dims, seed=11)
    n = A.shape[0]
    D = ua.DeviceCsr(ctx, A, dtype=dt)
    assert D.m.dia_sym == 2 and D.m.dia_nan == 1
    assert lib().uhip_cheb_pair_supported(D.ref) == 1
    code = ua.F64 if dt is np.float64 else ua.F32
    (b,) = _vecs(ctx, n, dt, [3])
    c1, r1, c2, r2 = (ua.DeviceArray(ctx, n, dt) for _ in range(4))
    # separate: the last step from a zero guess (pprev = b), then r_R = b - A c
    st = uhip_cheb_step(pprev=b.ptr, p_dtype=code, rprev=None, xin=None, xout=c1.ptr, x_dtype=code, rout=None,
                        pout=None, last=1, dot_with=None, dot_d=None, **COEF)
    ua.check(lib().uhip_cheb_step_apply(ctx.h, D.ref, ctypes.byref(st)))
    ua.check(lib().uhip_spmv_residual(ctx.h, D.ref, P(c1), P(b), code, P(r1)))
    ua.check(lib().uhip_cheb_pre_residual(ctx.h, D.ref, P(b), COEF["a_prev"], COEF["beta"], COEF["a_cur"], P(c2),
                                          P(r2)))
    assert np.array_equal(c2.numpy(), c1.numpy())
    assert np.array_equal(r2.numpy(), r1.numpy())


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("dot", [False, True])
def test_residual_post_equals_residual_then_step(ctx, dims, dt, dot):
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, *dims, seed=12)
    n = A.shape[0]
    D = ua.DeviceCsr(ctx, A, dtype=dt)
    code = ua.F64 if dt is np.float64 else ua.F32
    x, b = _vecs(ctx, n, dt, [4, 5])
    (dw,) = _vecs(ctx, n, np.float64, [6])
    pa, z1, z2 = (ua.DeviceArray(ctx, n, dt) for _ in range(3))
    d1, d2 = ua.DeviceArray(ctx, 1), ua.DeviceArray(ctx, 1)
    ua.check(lib().uhip_spmv_residual(ctx.h, D.ref, P(x), P(b), code, P(pa)))
    st = uhip_cheb_step(pprev=pa.ptr, p_dtype=code, rprev=None, xin=x.ptr, xout=z1.ptr, x_dtype=code, rout=None,
                        pout=None, last=1, dot_with=dw.ptr if dot else None, dot_d=d1.ptr if dot else None, **COEF)
    ua.check(lib().uhip_cheb_step_apply(ctx.h, D.ref, ctypes.byref(st)))
    ua.check(lib().uhip_cheb_residual_post(ctx.h, D.ref, P(x), P(b), COEF["a_prev"], COEF["beta"], COEF["a_cur"],
                                           P(z2), P(dw) if dot else None, P(d2) if dot else None))
    z = z2.numpy()
    assert np.array_equal(z, z1.numpy())
    if dot:
        w = dw.numpy()
        ref = float(np.dot(z.astype(np.float64), w))
        bound = 2 * n * np.finfo(np.float64).eps * float(np.dot(np.abs(z.astype(np.float64)), np.abs(w)))
        assert abs(d2.numpy()[0] - ref) <= bound
        assert abs(d2.numpy()[0] - d1.numpy()[0]) <= bound


def test_pair_workgroups_follow_the_grid(ctx):
    """Planes per workgroup shrink from 32 (the 256^3 grid: 256 groups in FP32) to 8 so that a
    pair launch keeps 256 workgroups where it can: 256 x 256 x 32 (8 planes each: 128 groups in
    FP32, 256 in FP64) and a small grid (8 planes each: 2 and 4 groups)."""
    for dims, want32, want64 in (((256, 256, 32), 128, 256), ((64, 32, 16), 2, 4)):
        A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, *dims, seed=1)
        assert lib().uhip_cheb_pair_workgroups(ua.DeviceCsr(ctx, A, dtype=np.float32).ref) == want32
        assert lib().uhip_cheb_pair_workgroups(ua.DeviceCsr(ctx, A).ref) == want64
    R = ua.gen_stencil(ua.STENCIL_2D5, 64, 64)
    assert lib().uhip_cheb_pair_workgroups(ua.DeviceCsr(ctx, R).ref) == 0


def test_unsupported_operators_are_reported(ctx):
    """Planes of 256 and 1024 rows (no whole FP32 chunk), a 2D operator, overlapping vectors."""
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 16, 16, 13, seed=1)
    D = ua.DeviceCsr(ctx, A, dtype=np.float32)
    assert D.m.dia_sym == 2 and lib().uhip_cheb_pair_supported(D.ref) == 0
    n = A.shape[0]
    b, c, r = (ua.DeviceArray(ctx, n, np.float32) for _ in range(3))
    st = lib().uhip_cheb_pre_residual(ctx.h, D.ref, P(b), 0.1, 0.1, 0.1, P(c), P(r))
    assert st == UNSUPPORTED and b"chunks per plane" in lib().uhip_last_error()
    A1 = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 32, 32, 5, seed=1)  # 1024-row planes: FP64 only
    assert lib().uhip_cheb_pair_supported(ua.DeviceCsr(ctx, A1, dtype=np.float32).ref) == 0
    assert lib().uhip_cheb_pair_supported(ua.DeviceCsr(ctx, A1).ref) == 1
    R = ua.gen_stencil(ua.STENCIL_2D5, 64, 64)
    assert lib().uhip_cheb_pair_supported(ua.DeviceCsr(ctx, R).ref) == 0
    # overlapping vectors are refused
    A2 = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 64, 32, 4, seed=1)
    D2 = ua.DeviceCsr(ctx, A2, dtype=np.float32)
    m = A2.shape[0]
    u, v = ua.DeviceArray(ctx, m, np.float32), ua.DeviceArray(ctx, m, np.float32)
    assert lib().uhip_cheb_pre_residual(ctx.h, D2.ref, P(u), 0.1, 0.1, 0.1, P(u), P(v)) == -1
    assert lib().uhip_cheb_residual_post(ctx.h, D2.ref, P(u), P(v), 0.1, 0.1, 0.1, P(u), None, None) == -1


@pytest.mark.parametrize("kind", ["mg", "mg_mixed"])
def test_multigrid_cycles_fused_equal_unfused(ctx, kind):
    """Multigrid as the solver (no dot in its cycles): fused and unfused smoothing give the same
    iterates bit for bit; FP64 levels and, through the mixed CG, FP32 levels (fused_smoothing 2:
    pairs also on a grid too small to give every CU a workgroup)."""
    n = 64
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n // 2, seed=5)
    b = ua.gen_uniform(105, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n // 2, 3)
    out = []
    for fused in (0, 2):
        ua.srand(1)
        if kind == "mg":
            s = ua.Multigrid(ctx, atol=0, rtol=0, stol=0, max_it=3, pre_smoothing_steps=2, post_smoothing_steps=2,
                             fused_smoothing=fused)
            s.set_transfer_operators(interps)
            out.append(s.solve(A, b))
        else:
            cg = ua.ConjugateGradient(ctx, "multigrid_mixed", atol=1e-10 * np.linalg.norm(b), rtol=0, stol=0,
                                      max_it=200, pre_smoothing_steps=2, post_smoothing_steps=2,
                                      fused_smoothing=fused)
            cg.set_transfer_operators(interps)
            x = cg.solve(A, b)
            out.append((x, cg.status["iterates"]))
    if kind == "mg":
        assert np.array_equal(out[0], out[1])
    else:
        # the CG's z.r is summed in another order by the fused post-smoothing: the iterates agree
        # to rounding, the iteration count exactly
        (x0, i0), (x1, i1) = out
        assert i0 == i1
        assert np.linalg.norm(x1 - x0) <= 1e-12 * np.linalg.norm(x0)
