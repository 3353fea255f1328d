"""Parity at the benchmark workloads' full sizes, GPU against the oracle on the same seeded inputs.

  config 2  MG-PCG FP64, 3D 7-point Poisson 128^3, 5 levels, V(2,2) Chebyshev, rtol 1e-10:
            iterations +-1, solution rel. error <= 1e-10 (PetscLinearSolverTest.cpp:400-477 solves
            the same kind of system with the same solver stack)
  config 3  MG-PCG 3D 7-point variable coefficient 256^3, FP64 and mixed forms, rtol 1e-10:
            iterations +-1, FP64 rel. error <= 1e-10, mixed within the FP64 bar too (its outer
            CG is FP64)
  config 4  MPRGP, 1024^2 9-point + 1e-3 I, l = -0.05 - 0.15 u, projected gradient 1e-9: KKT
            certificate of the GPU solution and active sets identical up to ties
  config 5  the whole batch, 64 systems x 8 right-hand sides, FP32 V-cycle in FP64 CG, rtol 1e-8:
            every pair's residual, sampled systems against the oracle's FP64 MG-PCG, and the
            adjoint gradient against the transposed-solve gradient and central differences

  config 3's 128^3 Galerkin level (P^T A P of the 256^3 operator, 2.1M rows, 27 offsets) in the
            plans the solve gives it: y = A x bit for bit against the oracle's CSR sums in FP64
            (the near-symmetric layout: upper half and lower bit-pattern deltas) and FP32

Error bounds of the rtol-1e-8 batch solves are a-posteriori: for SPD A, |x - x*| <= |r| / lmin(A).
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as orc
import utopia_amd as ua
from utopia_amd import batched as bt

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _mg_pcg_vs_oracle(ctx, kind, n, levels, seed, form):
    A = ua.gen_stencil(kind, n, n, n, seed=seed)
    b = ua.gen_uniform(seed + 100, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, levels)
    atol = 1e-10 * np.linalg.norm(b)
    ua.srand(1)
    cg = ua.ConjugateGradient(ctx, form, atol=atol, rtol=0, stol=0, max_it=1000, pre_smoothing_steps=2,
                              post_smoothing_steps=2)
    cg.set_transfer_operators(interps)
    x = cg.solve(A, b)
    its = cg.status["iterates"]
    del cg
    ua.srand(1)
    mg = orc.MG(interps, pre=2, post=2)
    mg.update(A)
    xo, so = mg.cg_solve(b, atol=atol, max_it=1000)
    print(f"{form} {n}^3 kind {kind}: gpu it={its} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}", flush=True)
    assert abs(its - so["iterates"]) <= 1
    assert np.linalg.norm(b - A.to_scipy() @ x) <= 1.01 * atol
    assert rel(x, xo) <= 1e-10
    return its


def test_config2_poisson_128_fp64_matches_oracle(ctx):
    _mg_pcg_vs_oracle(ctx, ua.STENCIL_3D7, 128, 5, 3, "multigrid")


@pytest.mark.timeout(600)
def test_config3_varcoef_256_fp64_and_mixed_match_oracle(ctx):
    """The benchmark's own problem (seed 3): 136 iterations in the oracle."""
    n, levels, seed = 256, 5, 3
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=seed)
    b = ua.gen_uniform(seed + 100, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, levels)
    atol = 1e-10 * np.linalg.norm(b)
    xs, its = {}, {}
    for form in ("multigrid", "multigrid_mixed"):
        ua.srand(1)
        cg = ua.ConjugateGradient(ctx, form, atol=atol, rtol=0, stol=0, max_it=1000, pre_smoothing_steps=2,
                                  post_smoothing_steps=2)
        cg.set_transfer_operators(interps)
        xs[form] = cg.solve(A, b)
        its[form] = cg.status["iterates"]
        del cg
    ua.srand(1)
    mg = orc.MG(interps, pre=2, post=2)
    mg.update(A)
    xo, so = mg.cg_solve(b, atol=atol, max_it=1000)
    del mg
    Asp = A.to_scipy()
    for form, x in xs.items():
        print(f"{form} 256^3: gpu it={its[form]} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}", flush=True)
        assert abs(its[form] - so["iterates"]) <= 1
        assert np.linalg.norm(b - Asp @ x) <= 1.01 * atol
        assert rel(x, xo) <= 1e-10
    assert so["iterates"] + 1 == 136


@pytest.mark.timeout(300)
def test_config4_mprgp_1024_active_set_matches_oracle(ctx):
    """MPRGP's proportioning / expansion tests compare quantities equal up to rounding, so a
    one-ulp difference changes its path: the iteration count is reported, not compared (the
    oracle's own count moves 10-20 % under a one-ulp change of b); solution and active set are."""
    n, seed = 1024, 3
    A = ua.gen_stencil(ua.STENCIL_2D9, n, n, shift=1e-3)
    b = ua.gen_uniform(seed, n * n)
    lb = -0.05 - 0.15 * ua.gen_uniform(seed + 1, n * n, 0, 1)
    q = ua.MPRGP(ctx, atol=1e-9, rtol=0, max_it=1000000)
    q.update(A)
    q.set_box_constraints(lb, None)
    x = q.apply(b)
    xo, so = orc.mprgp_solve(A, b, lb=lb, atol=1e-9, rtol=0, max_it=1000000)
    g = A.to_scipy() @ x - b
    act = np.abs(x - lb) < 1e-12
    assert np.all(x >= lb)
    assert np.all(g[act] >= -1e-8) and np.abs(g[~act]).max() <= 1e-8  # KKT of the GPU solution
    act_o = np.abs(xo - lb) < 1e-12
    ties = np.abs(x - lb) < 1e-9
    print(f"mprgp 1024^2: gpu it={q.status['iterates']} oracle it={so['iterates']} active {act.sum()} / "
          f"{act_o.sum()} rel.err={rel(x, xo):.2e}", flush=True)
    assert np.all((act == act_o) | ties)
    assert rel(x, xo) <= 1e-7


def _lmin(A):
    """Smallest eigenvalue of an SPD matrix (shift-invert Lanczos)."""
    return float(spla.eigsh(A.tocsc(), k=1, sigma=0, which="LM", return_eigenvectors=False)[0])


@pytest.mark.timeout(600)
def test_config5_full_batch_mixed_with_adjoint(ctx):
    seed, k, rtol = 3, 8, 1e-8
    sides = bt.batch_sides(seed, 64)
    p = bt.make_batch(sides, seed=seed * 1000)
    B = ua.gen_uniform(seed + 7, p.n * k).reshape(p.n, k)
    Xs = ua.gen_normal(seed + 8, p.n * k).reshape(p.n, k)
    ua.srand(1)
    bs = bt.BatchedSolver(ctx, k, mixed=True, rtol=rtol, max_it=500, pre_smoothing_steps=2,
                          post_smoothing_steps=2).setup(p)
    Bd = ua.DeviceArray.from_numpy(ctx, B.reshape(-1))
    Xd = ua.DeviceArray(ctx, p.n * k)
    bs.solve_device(Bd, Xd)
    its, flags = bs.convergence()
    replays, n_it = bs.info("graph_replays"), bs.info("iterations")
    X = Xd.numpy().reshape(-1, k)
    # adjoint: G = X - X*, lambda = A^-T G (A symmetric), dL/dA_ij = -sum_c lambda_ic X_jc
    G = X - Xs
    Gd = ua.DeviceArray.from_numpy(ctx, G.reshape(-1))
    Ld = ua.DeviceArray(ctx, p.n * k)
    bs.solve_device(Gd, Ld)
    its_adj, flags_adj = bs.convergence()
    grad = ua.DeviceArray(ctx, p.A.nnz)
    bs.pattern_grad_device(Ld, Xd, grad)
    lam, gA = Ld.numpy().reshape(-1, k), grad.numpy()
    print(f"config 5: {p.nseg} systems, {p.n} rows, iterations {its.min()}..{its.max()}, adjoint "
          f"{its_adj.min()}..{its_adj.max()}, {replays:.0f} graph replays for {n_it:.0f} iterations", flush=True)
    assert (flags == 1).all() and (flags_adj == 1).all()
    assert replays <= n_it / 8
    Asp = p.A.to_scipy()
    R, RL = B - Asp @ X, G - Asp @ lam
    for s in range(p.nseg):
        a, e = p.seg[s], p.seg[s + 1]
        assert (np.linalg.norm(R[a:e], axis=0) <= rtol * np.linalg.norm(B[a:e], axis=0) * 1.0001).all(), s
        assert (np.linalg.norm(RL[a:e], axis=0) <= rtol * np.linalg.norm(G[a:e], axis=0) * 1.0001).all(), s

    rows = np.repeat(np.arange(p.n), np.diff(p.A.rowptr))
    order = np.argsort(sides)
    for s in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        a, e = p.seg[s], p.seg[s + 1]
        As = p.systems[s].to_scipy().tocsc()
        lmin = _lmin(As)
        lu = spla.splu(As)
        Xd_ = lu.solve(np.ascontiguousarray(B[a:e]))
        lam_d = lu.solve(np.ascontiguousarray(Xd_ - Xs[a:e]))
        # forward: a-posteriori bound per column
        for j in range(k):
            assert np.linalg.norm(X[a:e, j] - Xd_[:, j]) <= 1.01 * np.linalg.norm(R[a:e, j]) / lmin, (s, j)
        # adjoint: lambda - lambda_d = A^-1 (r_lambda + (X - X_d))
        dX = np.linalg.norm(X[a:e] - Xd_, axis=0)
        bound_l = (np.linalg.norm(RL[a:e], axis=0) + dX) / lmin
        assert (np.linalg.norm(lam[a:e] - lam_d, axis=0) <= 1.01 * bound_l).all(), s
        # pattern gradient of system s against the transposed-solve gradient; each row and each
        # column holds at most 5 entries, so |dg| <= sqrt(5) (|dlam|_F max|X| + max|lam| |dX|_F)
        ent = slice(p.A.rowptr[a], p.A.rowptr[e])
        r_, c_ = rows[ent], p.A.col[ent]
        ref = -(lam_d[r_ - a] * Xd_[c_ - a]).sum(axis=1)
        bound_g = np.sqrt(5) * (np.linalg.norm(bound_l) * np.abs(X[a:e]).max()
                                + np.abs(lam_d).max() * np.linalg.norm(dX)) * 1.01
        err_g = np.linalg.norm(gA[ent] - ref)
        # the oracle's FP64 MG-PCG on this system's first and last right-hand sides, with the
        # batch's Chebyshev bounds: iterations +-1, solutions within the a-posteriori bound
        mg = orc.MG(p.interps[s], pre=2, post=2, masks=False)
        mg.update(p.systems[s])
        for l in range(1, p.levels):
            em = bs.eig_max(l)[s]
            mg.set_eigs(l, em, 0.06 * em)
        for j in (0, k - 1):
            bj = np.ascontiguousarray(B[a:e, j])
            xo, so = mg.cg_solve(bj, atol=rtol * np.linalg.norm(bj), max_it=500)
            ro = np.linalg.norm(bj - As @ xo)
            assert abs(int(its[s, j]) - (so["iterates"] + 1)) <= 1, (s, j, its[s, j], so["iterates"])
            assert np.linalg.norm(X[a:e, j] - xo) <= 1.01 * (np.linalg.norm(R[a:e, j]) + ro) / lmin
        print(f"system {s} (side {sides[s]}): lmin {lmin:.3e}, gradient rel.err {err_g / np.linalg.norm(ref):.2e} "
              f"(bound {bound_g / np.linalg.norm(ref):.2e}), oracle it {so['iterates'] + 1} gpu {its[s, k - 1]}",
              flush=True)
        assert err_g <= bound_g

    # central differences (h = 1e-6) of L = 1/2 |X - X*|^2 restricted to the smallest system pin
    # the transposed-solve gradient (to 1e-6), which bounds the GPU's above; a diagonal and an
    # off-diagonal entry of A and one entry of b. The direct solves behind L are refined with
    # residuals in extended precision: plain FP64 solves carry errors of kappa * eps, which the
    # difference quotient divides by 2h (a 1e-4 relative error at this size).
    s = int(order[0])
    a, e = p.seg[s], p.seg[s + 1]
    As = p.systems[s].to_scipy().tocsr()
    Bs_, Xss = np.ascontiguousarray(B[a:e]), Xs[a:e]
    lu = spla.splu(As.tocsc())
    Xd_ = lu.solve(Bs_)
    lam_d = lu.solve(np.ascontiguousarray(Xd_ - Xss))

    def loss(M, Bm):
        lu_m, Ml = spla.splu(M.tocsc()), M.astype(np.longdouble)
        Xm = lu_m.solve(Bm).astype(np.longdouble)
        for _ in range(3):
            Xm = Xm + lu_m.solve(np.asarray(Bm.astype(np.longdouble) - Ml @ Xm, np.float64))
        return 0.5 * np.sum((Xm - Xss.astype(np.longdouble)) ** 2)

    h = 1e-6
    base = p.A.rowptr[a]
    for le in (0, 2):  # first row of the system: its diagonal and its +side neighbour
        i, j = 0, int(p.systems[s].col[le])
        M1, M2 = As.copy(), As.copy()
        M1[i, j] += h
        M2[i, j] -= h
        fd = float((loss(M1, Bs_) - loss(M2, Bs_)) / (M1[i, j] - M2[i, j]))
        ref = -(lam_d[i] * Xd_[j]).sum()
        print(f"dL/dA[{i},{j}]: gpu {gA[base + le]:.10e} transposed solve {ref:.10e} fd {fd:.10e}", flush=True)
        assert abs(ref - fd) <= 1e-6 * abs(fd)
    B1, B2 = Bs_.copy(), Bs_.copy()
    B1[3, 2] += h
    B2[3, 2] -= h
    fd = float((loss(As, B1) - loss(As, B2)) / (B1[3, 2] - B2[3, 2]))
    print(f"dL/db[3,2]: gpu {lam[a + 3, 2]:.10e} transposed solve {lam_d[3, 2]:.10e} fd {fd:.10e}", flush=True)
    assert abs(lam_d[3, 2] - fd) <= 1e-6 * abs(fd)


def test_config3_galerkin_level_bit_exact(ctx):
    n = 256
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=3)
    Pm = ua.mg_hierarchy(n, n, n, 2)[0]
    L = ua.lib()
    DA = ua.DeviceCsr(ctx, A, fmt=ua.FMT_ROWBLOCK)
    DP = ua.DeviceCsr(ctx, Pm, fmt=ua.FMT_ROWBLOCK)
    raw = ua.uhip_csr()
    ua.check(L.uhip_csr_ptap(ctx.h, DA.ref, DP.ref, ctypes.byref(raw)))
    C = ua.csr_from_raw(ctx, raw)
    ua.check(L.uhip_csr_free(ctx.h, ctypes.byref(raw)))
    del DA, DP
    m = C.shape[0]
    assert m == 128 ** 3
    x = ua.gen_uniform(7, m)
    D = ua.DeviceCsr(ctx, C)
    assert D.format == ua.FMT_DIA and D.m.dia_w27 == 1 and D.m.dia_sym == 1 and D.m.dia_delta
    xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray(ctx, m)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), orc.spmv(C, x))
    del D
    x32 = x.astype(np.float32)
    D32 = ua.DeviceCsr(ctx, C, dtype=np.float32)
    assert D32.format == ua.FMT_DIA and D32.m.dia_w27 == 1 and D32.m.dia_sym == 1
    xd32, yd32 = ua.DeviceArray.from_numpy(ctx, x32), ua.DeviceArray(ctx, m, np.float32)
    D32.spmv(xd32, yd32)
    assert np.array_equal(yd32.numpy(), orc.spmv_f32(C, x32))
