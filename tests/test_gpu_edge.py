"""Edge cases of the HIP path: empty vectors, 1x1 and diagonal systems, a zero
right-hand side, and bound-pinned QPs, each against the oracle or a closed form."""
import ctypes

import numpy as np
import pytest

import oracle as orc
import utopia_amd as ua

pytestmark = pytest.mark.gpu


def P(d):
    return ctypes.c_void_p(d.ptr)


def test_blas1_on_empty_vectors(ctx):
    h, l = ctx.h, ua.lib()
    x, y = ua.DeviceArray(ctx, 0), ua.DeviceArray(ctx, 0)
    s = ctx.scalars
    ua.check(l.uhip_vec_dot_f64(h, 0, P(x), P(y), ctypes.c_void_p(s)))
    assert ctx.read_scalars(s, 1)[0] == 0.0
    ua.check(l.uhip_vec_norm2_f64(h, 0, P(x), ctypes.c_void_p(s)))
    assert ctx.read_scalars(s, 1)[0] == 0.0
    ua.check(l.uhip_vec_reduce_f64(h, 0, P(x), 0, ctypes.c_void_p(s)))
    assert ctx.read_scalars(s, 1)[0] == 0.0
    ua.check(l.uhip_vec_axpy_f64(h, 0, 2.0, P(x), P(y)))
    ua.check(l.uhip_vec_set_f64(h, 0, 1.0, P(y)))


def test_one_by_one_system(ctx):
    A = ua.Csr(np.array([0, 1], np.int32), np.array([0], np.int32), np.array([4.0]), (1, 1))
    D = ua.DeviceCsr(ctx, A)
    assert D.format == ua.FMT_DIA and D.m.ndiag == 1 and D.m.dia_sym == 0
    cg = ua.ConjugateGradient(ctx, "jacobi", atol=1e-14, rtol=0, stol=0, max_it=10)
    x = cg.solve(A, np.array([2.0]))
    assert x[0] == 0.5 and cg.status["iterates"] <= 1


def test_diagonal_operator_single_offset(ctx):
    n = 1000
    d = 1.0 + ua.gen_uniform(3, n, 0, 1)
    A = ua.Csr(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d, (n, n))
    D = ua.DeviceCsr(ctx, A)
    assert D.format == ua.FMT_DIA and D.m.ndiag == 1
    x = ua.gen_uniform(4, n)
    xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray(ctx, n)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), d * x)
    b = ua.gen_uniform(5, n)
    # Jacobi-preconditioned CG is exact after one step on a diagonal operator
    cg = ua.ConjugateGradient(ctx, "jacobi", atol=1e-12, rtol=0, stol=0, max_it=10)
    xs = cg.solve(A, b)
    xo, so = orc.cg_solve(A, b, precond=1, atol=1e-12, rtol=0, stol=0, max_it=10)
    assert cg.status["iterates"] == so["iterates"] <= 2
    assert np.array_equal(xs, xo)


def test_zero_rhs_returns_zero_without_a_step(ctx):
    """b = 0: the unpreconditioned loop stops on rho == 0 and the preconditioned one on
    the initial residual norm (utopia_ConjugateGradient_impl.hpp:115-140,219-235)."""
    A = ua.gen_stencil(ua.STENCIL_2D5, 20, 20)
    b = np.zeros(400)
    for pc in (None, "jacobi"):
        cg = ua.ConjugateGradient(ctx, pc, atol=1e-10, rtol=0, stol=0, max_it=50)
        x = cg.solve(A, b)
        xo, so = orc.cg_solve(A, b, precond=0 if pc is None else 1, atol=1e-10, rtol=0, stol=0, max_it=50)
        assert np.array_equal(x, np.zeros(400)) and np.array_equal(xo, np.zeros(400))
        assert cg.status["iterates"] == so["iterates"] <= 1


def test_qp_with_coinciding_bounds(ctx):
    """lb = ub. ProjectedGaussSeidel clamps every correction to [lb - x, ub - x] = {0}
    and returns the bound. MPRGP's chopped gradient treats a component sitting on both
    bounds as on the lower one (min(g, 0)), so its proportioning steps leave the box;
    the oracle's MPRGP does the same, and the GPU run follows it."""
    n = 24
    A = ua.gen_stencil(ua.STENCIL_2D9, n, n, shift=1e-3)
    b = ua.gen_uniform(9, n * n)
    c = 0.1 * ua.gen_uniform(10, n * n)
    q = ua.ProjectedGaussSeidel(ctx, atol=1e-9, rtol=0, max_it=1000)
    q.update(A)
    q.set_box_constraints(c, c)
    assert np.array_equal(q.apply(b), c)
    m = ua.MPRGP(ctx, atol=1e-9, rtol=0, max_it=1000)
    m.update(A)
    m.set_box_constraints(c, c)
    ua.srand(1)
    x = m.apply(b)
    ua.srand(1)
    xo, so = orc.mprgp_solve(A, b, lb=c, ub=c, atol=1e-9, rtol=0, max_it=1000)
    assert m.status["iterates"] == so["iterates"]
    assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)


def test_mprgp_fused_cg_step_state_matches_separate_passes(ctx):
    """uhip_mprgp_cg_step_state_f64 = uhip_mprgp_cg_step_f64 then uhip_mprgp_state_f64,
    bit for bit: the iterate, the gradients and the four dots."""
    n, a = 70001, 0.37
    rng = np.random.default_rng(5)
    x0 = rng.uniform(-0.2, 0.3, n)
    lb = np.where(rng.random(n) < 0.3, x0, -0.1)  # some components on the bound
    ub = np.full(n, 0.25)
    x0 = np.clip(x0, lb, ub)
    g0, p, Ap = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    L = ua.lib()
    d = {k: ua.DeviceArray.from_numpy(ctx, v) for k, v in (("lb", lb), ("ub", ub), ("p", p), ("Ap", Ap))}
    s = ctx.scalars
    out = []
    for fused in (False, True):
        x, g = ua.DeviceArray.from_numpy(ctx, x0), ua.DeviceArray.from_numpy(ctx, g0)
        fi, beta, gp = ua.DeviceArray(ctx, n), ua.DeviceArray(ctx, n), ua.DeviceArray(ctx, n)
        if fused:
            ua.check(L.uhip_mprgp_cg_step_state_f64(ctx.h, n, P(x), P(g), P(d["p"]), P(d["Ap"]), P(d["lb"]),
                                                    P(d["ub"]), P(fi), P(beta), P(gp), ctypes.c_double(a),
                                                    ctypes.c_void_p(s + 8 * 40)))
            dots = ctx.read_scalars(s + 8 * 40, 4)
        else:
            ua.check(L.uhip_mprgp_cg_step_f64(ctx.h, n, P(x), P(g), P(d["p"]), P(d["Ap"]), P(d["lb"]), P(d["ub"]),
                                              P(fi), ctypes.c_double(a), ctypes.c_void_p(s), 43))
            ua.check(L.uhip_mprgp_state_f64(ctx.h, n, P(x), P(g), P(d["lb"]), P(d["ub"]), P(fi), P(beta), P(gp),
                                            ctypes.c_void_p(s + 8 * 40)))
            dots = ctx.read_scalars(s + 8 * 40, 4)
        out.append([x.numpy(), g.numpy(), fi.numpy(), beta.numpy(), gp.numpy(), dots])
    for u, v in zip(*out):
        assert np.array_equal(u, v)
    assert np.allclose(out[0][0], x0 - a * p, rtol=0, atol=1e-15)
