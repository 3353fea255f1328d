"""Multi-RHS SpMM, segmented reductions and the batched MG-PCG on the GPU.

The batched solver has no reference analogue; it is pinned to the unbatched
solver (ConjugateGradient + Multigrid, itself pinned to the oracle) system by
system and column by column: iteration counts within +-1, FP64 solutions
within 1e-10 (relative), mixed-precision solutions within the FP64 CG's
tolerance (relative difference <= 1e-7 at rtol 1e-10). SpMM columns are bit-identical to the oracle's sequential CSR
row sums.
"""
import ctypes

import numpy as np
import pytest

import oracle as orc
import utopia_amd as ua
from utopia_amd import batched as bt
from utopia_amd._abi import lib

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _p(a):
    return ctypes.c_void_p(a.ptr)


@pytest.mark.parametrize("k", [1, 2, 4, 8, 16])
def test_spmm_columns_bit_exact(ctx, k):
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 14, 11, 9, seed=4)
    n = A.shape[0]
    X = ua.gen_uniform(40 + k, n * k).reshape(n, k)
    dA = ua.DeviceCsr(ctx, A)
    dX = ua.DeviceArray.from_numpy(ctx, X)
    dY = ua.DeviceArray(ctx, n * k)
    ua.check(lib().uhip_spmm(ctx.h, dA.ref, k, _p(dX), ua.F64, _p(dY), ua.F64, 0, None, 0), "spmm")
    Y = dY.numpy().reshape(n, k)
    for j in range(k):
        assert np.array_equal(Y[:, j], orc.spmv(A, np.ascontiguousarray(X[:, j]))), f"column {j}"
    # residual and accumulate forms
    B = ua.gen_uniform(90, n * k).reshape(n, k)
    dB = ua.DeviceArray.from_numpy(ctx, B)
    ua.check(lib().uhip_spmm(ctx.h, dA.ref, k, _p(dX), ua.F64, _p(dY), ua.F64, 2, _p(dB), ua.F64), "spmm r")
    R = dY.numpy().reshape(n, k)
    for j in range(k):
        assert np.array_equal(R[:, j], B[:, j] - orc.spmv(A, np.ascontiguousarray(X[:, j])))
    dY.upload(B.reshape(-1))
    ua.check(lib().uhip_spmm(ctx.h, dA.ref, k, _p(dX), ua.F64, _p(dY), ua.F64, 1, None, 0), "spmm add")
    S = dY.numpy().reshape(n, k)
    for j in range(k):
        assert np.array_equal(S[:, j], B[:, j] + orc.spmv(A, np.ascontiguousarray(X[:, j])))


def test_spmm_fp32_matches_spmv_fp32(ctx):
    A = ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, 50, 37, seed=2)
    n, k = A.shape[0], 8
    X = ua.gen_uniform(7, n * k).astype(np.float32).reshape(n, k)
    dA = ua.DeviceCsr(ctx, A, dtype=np.float32)
    dX = ua.DeviceArray.from_numpy(ctx, X)
    dY = ua.DeviceArray(ctx, n * k, np.float32)
    ua.check(lib().uhip_spmm(ctx.h, dA.ref, k, _p(dX), ua.F32, _p(dY), ua.F32, 0, None, 0), "spmm f32")
    Y = dY.numpy().reshape(n, k)
    for j in range(k):
        assert np.array_equal(Y[:, j], orc.spmv_f32(A, np.ascontiguousarray(X[:, j])))


def test_segdot_per_system(ctx):
    p = bt.make_batch([7, 19, 3, 30], seed=1, levels=2)
    k = 4
    n = p.n
    x = ua.gen_uniform(3, n * k).reshape(n, k)
    y = ua.gen_uniform(4, n * k).reshape(n, k)
    # chunks of at most 64 rows, never across systems
    r0, r1, c0 = [], [], [0]
    for s in range(p.nseg):
        for a in range(p.seg[s], p.seg[s + 1], 64):
            r0.append(a)
            r1.append(min(a + 64, p.seg[s + 1]))
        c0.append(len(r0))
    d = {name: ua.DeviceArray.from_numpy(ctx, np.asarray(v, np.int32)) for name, v in
         (("r0", r0), ("r1", r1), ("c0", c0))}
    dx, dy = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray.from_numpy(ctx, y)
    out = ua.DeviceArray(ctx, p.nseg * k)
    ua.check(lib().uhip_segdot(ctx.h, k, len(r0), _p(d["r0"]), _p(d["r1"]), p.nseg, _p(d["c0"]), _p(dx), _p(dy),
                               ua.F64, _p(out), 0), "segdot")
    got = out.numpy().reshape(p.nseg, k)
    ref = np.array([[x[p.seg[s]:p.seg[s + 1], j] @ y[p.seg[s]:p.seg[s + 1], j] for j in range(k)]
                    for s in range(p.nseg)])
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-13)


def test_pattern_gradient_closed_form(ctx):
    A = ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, 9, 8, seed=5)
    n, k = A.shape[0], 2
    lam = ua.gen_normal(1, n * k).reshape(n, k)
    x = ua.gen_normal(2, n * k).reshape(n, k)
    dA = ua.DeviceCsr(ctx, A)
    dl, dx = ua.DeviceArray.from_numpy(ctx, lam), ua.DeviceArray.from_numpy(ctx, x)
    g = ua.DeviceArray(ctx, A.nnz)
    ua.check(lib().uhip_pattern_grad(ctx.h, dA.ref, k, _p(dl), _p(dx), _p(g)), "pattern_grad")
    rows = np.repeat(np.arange(n), np.diff(A.rowptr))
    ref = -(lam[rows] * x[A.col]).sum(axis=1)
    assert np.allclose(g.numpy(), ref, rtol=1e-14, atol=1e-14)


def _unbatched(ctx, p, s, b, mixed, eigs, rtol):
    """The unbatched CG-MG solve of system s with the batched solver's eigenvalue bounds."""
    A = p.systems[s]
    cg = ua.ConjugateGradient(ctx, "multigrid_mixed" if mixed else "multigrid",
                              atol=rtol * np.linalg.norm(b), rtol=0, stol=0, max_it=500,
                              must_generate_masks=0, refine=0, pre_smoothing_steps=2, post_smoothing_steps=2)
    cg.set_transfer_operators(p.interps[s])
    cg.update(A)
    for l in range(1, p.levels):
        cg.set_level_eigs(l, eigs[l][s], 0.06 * eigs[l][s])
    x = cg.apply(b)
    return x, cg.status["iterates"] + 1


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_matches_unbatched_per_system(ctx, mixed):
    sides = [20, 33, 17, 40, 24]
    p = bt.make_batch(sides, seed=3, levels=3)
    k = 4
    rtol = 1e-10
    B = ua.gen_uniform(77, p.n * k).reshape(p.n, k)
    ua.srand(1)
    bs = bt.BatchedSolver(ctx, k, mixed=mixed, rtol=rtol, max_it=500, pre_smoothing_steps=2,
                          post_smoothing_steps=2)
    bs.setup(p)
    X = bs.solve(B)
    its, flags = bs.convergence()
    assert (flags == 1).all(), flags
    eigs = {l: bs.eig_max(l) for l in range(1, p.levels)}
    worst = 0.0
    for s in range(p.nseg):
        a, e = p.seg[s], p.seg[s + 1]
        for j in range(k):
            b = np.ascontiguousarray(B[a:e, j])
            x, it = _unbatched(ctx, p, s, b, mixed, eigs, rtol)
            assert abs(int(its[s, j]) - it) <= 1, (s, j, its[s, j], it)
            worst = max(worst, rel(X[a:e, j], x))
            r = np.linalg.norm(b - p.systems[s].to_scipy() @ X[a:e, j])
            assert r <= rtol * np.linalg.norm(b) * 1.0001
    print(f"batched ({'mixed' if mixed else 'fp64'}) vs unbatched: worst rel.err {worst:.2e}, its {its.ravel()}")
    # FP64: same arithmetic up to reduction order. Mixed: both runs meet |r| <= 1e-10 |b|;
    # FP32 rounding in the V-cycle may differ by an ulp (coarse inverse summation order),
    # so the bound is the solve tolerance times the conditioning headroom.
    assert worst <= (1e-7 if mixed else 1e-10)


def test_batched_fp64_against_oracle_one_system(ctx):
    """One system of the batch against the oracle's CG-MG (same eigenvalue bounds)."""
    p = bt.make_batch([28, 21], seed=8, levels=3)
    k = 2
    B = ua.gen_uniform(5, p.n * k).reshape(p.n, k)
    bs = bt.BatchedSolver(ctx, k, mixed=False, rtol=1e-10, max_it=500, pre_smoothing_steps=2,
                          post_smoothing_steps=2).setup(p)
    X = bs.solve(B)
    its, _ = bs.convergence()
    s, j = 1, 1
    a, e = p.seg[s], p.seg[s + 1]
    b = np.ascontiguousarray(B[a:e, j])
    mg = orc.MG(p.interps[s], pre=2, post=2, masks=False)
    mg.update(p.systems[s])
    for l in range(1, p.levels):
        em = bs.eig_max(l)[s]
        mg.set_eigs(l, em, 0.06 * em)
    xo, so = mg.cg_solve(b, atol=1e-10 * np.linalg.norm(b), max_it=500)
    assert abs(int(its[s, j]) - (so["iterates"] + 1)) <= 1
    assert rel(X[a:e, j], xo) <= 1e-10


def test_graph_replays_match_eager(ctx):
    """Graph replays (1, 3 and the default 16 iterations per replay, the kernels gated on the
    active-pair count) equal the eager loop bit for bit, with the same iteration count, on
    repeated solves with the host doing eager work between them."""
    p = bt.make_batch([26, 40, 18], seed=4, levels=3)
    k = 8
    B = ua.gen_uniform(11, p.n * k).reshape(p.n, k)
    eager = bt.BatchedSolver(ctx, k, mixed=True, rtol=1e-8, use_graph=0).setup(p)
    Xe = eager.solve(B)
    ite, _ = eager.convergence()
    assert eager.info("graph_replays") == 0
    n_it = eager.info("iterations")
    assert n_it == ite.max()
    for gi in (1, 3, None):
        kw = {} if gi is None else {"graph_iterations": gi}
        g = bt.BatchedSolver(ctx, k, mixed=True, rtol=1e-8, use_graph=1, **kw)
        g.setup(p)
        for l in range(1, p.levels):  # identical Chebyshev bounds
            for s in range(p.nseg):
                g.set_level_eig(l, s, eager.eig_max(l)[s])
        for rep in range(3):
            Xg = g.solve(B)
            itg, fl = g.convergence()
            assert g.info("graph_replays") >= 1
            assert g.info("iterations") == n_it
            assert g.info("graph_replays") == -(-n_it // (gi or 16))
            assert np.array_equal(Xg, Xe), (gi, rep)
            assert np.array_equal(itg, ite)
            assert (fl == 1).all()
            # eager work between replays
            ua.DeviceCsr(ctx, p.A).spmv(ua.DeviceArray.from_numpy(ctx, Xg[:, 0].copy()), ua.DeviceArray(ctx, p.n))


def test_graph_recaptured_after_workspace_grows(ctx):
    """A captured graph holds the reduction workspace's address: after the workspace is
    reallocated (here by a reservation) the next solve captures again, and its result
    equals the eager solve bit for bit."""
    p = bt.make_batch([22, 31, 15], seed=9, levels=3)
    k = 4
    B = ua.gen_uniform(12, p.n * k).reshape(p.n, k)
    eager = bt.BatchedSolver(ctx, k, mixed=True, rtol=1e-8, use_graph=0).setup(p)
    Xe = eager.solve(B)
    g = bt.BatchedSolver(ctx, k, mixed=True, rtol=1e-8, use_graph=1, graph_iterations=4).setup(p)
    for l in range(1, p.levels):
        for s in range(p.nseg):
            g.set_level_eig(l, s, eager.eig_max(l)[s])
    assert np.array_equal(g.solve(B), Xe)
    assert g.info("graph_captures") == 1
    assert np.array_equal(g.solve(B), Xe)
    assert g.info("graph_captures") == 1
    g.set_param("workspace_reserve", 1 << 22)  # reallocates the workspace
    assert np.array_equal(g.solve(B), Xe)
    assert g.info("graph_captures") == 2


def test_batched_uneven_levels_config_shape(ctx):
    """Sides drawn log-uniform in [64, 512] (8 systems here), 8 RHS, mixed V-cycle, rtol 1e-8."""
    sides = bt.batch_sides(21, 8)
    p = bt.make_batch(sides, seed=100)
    k = 8
    B = ua.gen_uniform(5, p.n * k).reshape(p.n, k)
    bs = bt.BatchedSolver(ctx, k, mixed=True, rtol=1e-8, max_it=300).setup(p)
    X = bs.solve(B)
    its, fl = bs.convergence()
    assert (fl == 1).all()
    Asp = p.A.to_scipy()
    R = B - Asp @ X
    for s in range(p.nseg):
        a, e = p.seg[s], p.seg[s + 1]
        assert (np.linalg.norm(R[a:e], axis=0) <= 1e-8 * np.linalg.norm(B[a:e], axis=0) * 1.0001).all()
    print(f"sides {sides} levels {p.levels} iterations {its.min()}..{its.max()}")


def test_batched_rejects_bad_setup(ctx):
    p = bt.make_batch([10, 12], seed=1, levels=2)
    bs = bt.BatchedSolver(ctx, 2, mixed=False)
    bad = bt.BatchProblem(p.sides, p.levels, p.systems, p.interps, p.A, p.seg.copy(), p.batch_interps,
                          [c.copy() for c in p.coarse_segs])
    bad.coarse_segs[0][-1] += 1
    with pytest.raises(ua.HipError):
        bs.setup(bad)
    with pytest.raises(ua.HipError):
        bt.BatchedSolver(ctx, 3)


def test_spmm_autotuner_and_autotuned_batched_solve(ctx):
    """uhip_csr_autotune_spmm keeps one of the two kernels and the product stays exact
    (thread-per-row: column order; the tree-ordered wavefront kernel within rounding);
    a batched solve with autotuned kernels converges like the heuristic one."""
    A = ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, 60, 50, seed=2)
    k = 8
    X = ua.gen_uniform(3, A.shape[1] * k).reshape(-1, k)
    D = ua.DeviceCsr(ctx, A)
    choice, ms = ctypes.c_int32(), ctypes.c_float()
    ua.check(ua.lib().uhip_csr_autotune_spmm(ctx.h, D.ref, k, 5, ctypes.byref(choice), ctypes.byref(ms)))
    assert choice.value in (1, 2) and D.m.spmm_kernel == choice.value and ms.value > 0
    Xd, Yd = ua.DeviceArray.from_numpy(ctx, X.ravel()), ua.DeviceArray(ctx, A.shape[0] * k)
    ua.check(ua.lib().uhip_spmm(ctx.h, D.ref, k, ctypes.c_void_p(Xd.ptr), ua.F64, ctypes.c_void_p(Yd.ptr), ua.F64, 0,
                                None, 0))
    Y = Yd.numpy().reshape(-1, k)
    ref = np.stack([orc.spmv(A, np.ascontiguousarray(X[:, j])) for j in range(k)], 1)
    if choice.value == 1:
        assert np.array_equal(Y, ref)
    else:
        assert np.abs(Y - ref).max() <= 1e-14 * np.abs(ref).max()
    p = bt.make_batch([20, 33, 17, 40, 24], seed=3, levels=3)
    B = ua.gen_uniform(77, p.n * 4).reshape(p.n, 4)
    out = []
    for at in (0, 3):
        ua.srand(1)
        bs = bt.BatchedSolver(ctx, 4, mixed=True, rtol=1e-10, max_it=500, pre_smoothing_steps=2,
                              post_smoothing_steps=2, autotune=at)
        bs.setup(p)
        out.append((bs.solve(B), bs.convergence()))
    (X0, (i0, f0)), (X1, (i1, f1)) = out
    assert (f0 == 1).all() and (f1 == 1).all()
    assert np.abs(i0.astype(int) - i1.astype(int)).max() <= 1
    assert rel(X1, X0) <= 1e-8
