"""Pins of the CPU oracle (oracle/oracle.c) by checks that do not restate it:
closed forms, brute force, invariants, library routines and the reference's
own fixtures (utopia/data/laplace/matrices_for_petsc, copied to tests/golden).

  oracle function        pinned by
  or_spmv / _f32         scipy product (exact on integer data), Laplacian eigenvector image
  or_dot                 math.fsum with the gamma_n error bound
  or_transpose           scipy transpose (exact)
  or_spgemm / or_ptap    brute-force Boolean product (pattern), dense product (values),
                         Galerkin closed form on the reference Q1 hierarchy
  or_cholesky(_solve)    numpy.linalg.cholesky / solve
  or_cg_solve            Jacobi-PCG on diag-4 operator == CG on A/4 bit for bit (power-of-two
                         scaling invariant); finite termination; scipy spsolve
  or_power_method        Rayleigh bound lambda <= lambda_max (closed form), determinism
  or_chebyshev_smooth    residual = scaled Chebyshev polynomial T_nu on diagonal operators
  or_mg_*                masks of the reference fixture (Dirichlet rows), convergence of
                         MG and CG-MG on the reference fixture (reference test assertions)
  or_mprgp / or_pgs      KKT certificate, brute-force active-set enumeration on tiny QPs
"""
import itertools
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as orc
import utopia_amd as ua
from utopia_amd import petsc_io


def lap1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])


def test_spmv_matches_scipy_exactly_on_integers():
    rng = np.random.default_rng(0)
    M = sp.random(300, 200, density=0.05, random_state=1, format="csr")
    M.data = rng.integers(-5, 6, M.nnz).astype(float)
    A = ua.Csr.from_scipy(M)
    x = rng.integers(-9, 10, 200).astype(float)
    assert np.array_equal(orc.spmv(A, x), M @ x)
    assert np.array_equal(orc.spmv_f32(A, x), (M @ x).astype(np.float32))


def test_spmv_laplacian_eigenvector():
    n = 40
    A = ua.gen_stencil(ua.STENCIL_2D5, n, n)
    i, j = 3, 7
    s = np.arange(1, n + 1) * np.pi / (n + 1)
    v = np.outer(np.sin(j * s), np.sin(i * s)).ravel()
    lam = 4 - 2 * np.cos(i * np.pi / (n + 1)) - 2 * np.cos(j * np.pi / (n + 1))
    assert np.abs(orc.spmv(A, v) - lam * v).max() < 1e-13


def test_dot_within_gamma_bound():
    x, y = ua.gen_uniform(1, 100_000), ua.gen_uniform(2, 100_000)
    exact = math.fsum(x * y)
    n = x.size
    gamma = n * np.finfo(float).eps / (1 - n * np.finfo(float).eps)
    assert abs(orc.dot(x, y) - exact) <= gamma * np.abs(x * y).sum()


def test_transpose_exact():
    M = sp.random(70, 90, density=0.1, random_state=3, format="csr")
    rp, col, val, shape = orc.transpose(ua.Csr.from_scipy(M))
    T = M.T.tocsr()
    T.sort_indices()
    assert shape == (90, 70)
    assert np.array_equal(rp, T.indptr) and np.array_equal(col, T.indices) and np.array_equal(val, T.data)


def _bool_pattern(C):
    rp, col, _, shape = C
    D = np.zeros(shape, bool)
    for i in range(shape[0]):
        D[i, col[rp[i]:rp[i + 1]]] = True
    return D


def test_ptap_pattern_brute_force_and_values():
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 6, 6, 4, seed=5)
    P = ua.gen_prolongation(6, 6, 4)
    C = orc.ptap(A, P)
    Ad, Pd = A.to_scipy().toarray(), P.to_scipy().toarray()
    bool_ref = ((Pd.T != 0).astype(int) @ (Ad != 0).astype(int) @ (Pd != 0).astype(int)) > 0
    assert np.array_equal(_bool_pattern(C), bool_ref)
    rp, col, val, shape = C
    Cd = sp.csr_matrix((val, col, rp), shape=shape).toarray()
    ref = Pd.T @ Ad @ Pd
    assert np.abs(Cd - ref).max() <= 1e-14 * np.abs(ref).max()
    for i in range(shape[0]):  # sorted, unique columns
        assert np.all(np.diff(col[rp[i]:rp[i + 1]]) > 0)


def test_galerkin_property_on_reference_q1_hierarchy(golden):
    d = golden / "utopia_data"
    A = petsc_io.read_mat(d / "f_A")
    I3 = petsc_io.read_mat(d / "I_3")
    rp, col, val, shape = orc.ptap(A, I3)
    C = sp.csr_matrix((val, col, rp), shape=shape)
    # interior coarse rows reproduce the Q1 stiffness stencil (8/3 centre, -1/3 neighbours)
    n = int(round(math.sqrt(shape[0])))
    assert n * n == shape[0]
    checked = 0
    for y in range(2, n - 2):
        for x in range(2, n - 2):
            i = x + n * y
            row = C.getrow(i)
            got = dict(zip(row.indices, row.data))
            assert abs(got[i] - 8 / 3) < 1e-13
            for dy, dx in itertools.product((-1, 0, 1), repeat=2):
                if dx or dy:
                    assert abs(got[(x + dx) + n * (y + dy)] + 1 / 3) < 1e-13
            assert sum(abs(v) > 1e-13 for v in row.data) == 9
            checked += 1
    assert checked > 100


def test_seven_point_galerkin_stencil_is_translation_invariant():
    n = 16
    A = ua.gen_stencil(ua.STENCIL_3D7, n, n, n)
    P = ua.gen_prolongation(n, n, n)
    rp, col, val, shape = orc.ptap(A, P)
    m = n // 2
    stencils = set()
    for z in range(2, m - 2):
        for y in range(2, m - 2):
            for x in range(2, m - 2):
                i = x + m * (y + m * z)
                offs = tuple(zip((col[rp[i]:rp[i + 1]] - i).tolist(), np.round(val[rp[i]:rp[i + 1]], 12).tolist()))
                stencils.add(offs)
    assert len(stencils) == 1 and len(next(iter(stencils))) == 27


def test_cholesky_matches_numpy():
    rng = np.random.default_rng(4)
    M = rng.standard_normal((50, 50))
    A = M @ M.T + 50 * np.eye(50)
    Lo, info = orc.cholesky(A)
    assert info == 0
    assert np.abs(Lo - np.linalg.cholesky(A)).max() < 1e-12
    b = rng.standard_normal(50)
    assert np.abs(orc.cholesky_solve(Lo, b) - np.linalg.solve(A, b)).max() < 1e-12
    _, info = orc.cholesky(np.diag([1.0, -1.0, 2.0]))
    assert info == 2


def test_jacobi_pcg_equals_cg_on_scaled_operator_bitwise():
    # diag(A) = 4: InvDiag-PCG on (A, b) and plain CG on (A/4, b/4) have the same iterates exactly
    A = ua.gen_stencil(ua.STENCIL_2D5, 48, 40)
    b = ua.gen_uniform(3, A.shape[0])
    atol = 1e-8 * np.linalg.norm(b)
    x1, s1 = orc.cg_solve(A, b, precond=True, atol=atol, rtol=0, stol=0, max_it=5000)
    A4 = ua.Csr(A.rowptr, A.col, A.val / 4, A.shape)
    x2, s2 = orc.cg_solve(A4, b / 4, precond=False, atol=atol / 4, rtol=0, stol=0, max_it=5000)
    assert s1["converged"] and s2["converged"]
    # the unpreconditioned loop counts from 1 (utopia_ConjugateGradient_impl.hpp:125)
    assert s2["iterates"] == s1["iterates"] + 1
    assert np.array_equal(x1, x2)


def test_cg_finite_termination_and_scipy_solution():
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((8, 8)))
    A = Q @ np.diag(np.arange(1, 9.0)) @ Q.T
    C = ua.Csr.from_scipy(sp.csr_matrix(A))
    b = rng.standard_normal(8)
    x, st = orc.cg_solve(C, b, precond=False, atol=1e-12, rtol=0, stol=0, max_it=100)
    assert st["iterates"] <= 9 and np.abs(A @ x - b).max() < 1e-10
    A2 = ua.gen_stencil(ua.STENCIL_2D5, 30, 30)
    b2 = ua.gen_uniform(9, 900)
    x2, st2 = orc.cg_solve(A2, b2, precond=True, atol=1e-12 * np.linalg.norm(b2), rtol=0, stol=0)
    assert np.abs(x2 - spla.spsolve(A2.to_scipy().tocsc(), b2)).max() < 1e-9


def test_power_method_rayleigh_bound_and_determinism():
    n = 12
    A = ua.gen_stencil(ua.STENCIL_3D7, n, n, n)
    lam_max = 3 * (2 + 2 * np.cos(np.pi / (n + 1)))
    orc.srand(1)
    l1 = orc.power_method(A)
    orc.srand(1)
    l2 = orc.power_method(A)
    assert l1 == l2
    assert 0.85 * lam_max <= l1 <= lam_max * (1 + 1e-14)


@pytest.mark.parametrize("nu", [1, 2, 3, 5])
def test_chebyshev_residual_polynomial(nu):
    d = np.linspace(0.3, 9.5, 40)
    A = ua.Csr.from_scipy(sp.diags(d).tocsr())
    b = ua.gen_uniform(2, 40)
    emax, emin = 10.0, 0.6
    x = orc.chebyshev_smooth(A, b, np.zeros(40), nu, emax, emin)
    avg, half = (emax + emin) / 2, (emax - emin) / 2
    T = np.polynomial.chebyshev.Chebyshev.basis(nu)
    poly = T((avg - d) / half) / T(avg / half)
    # residual b - A x_nu = P_nu(A) b with the scaled Chebyshev polynomial
    assert np.abs((b - d * x) - poly * b).max() < 1e-13


def test_mg_masks_on_reference_fixture(golden):
    d = golden / "utopia_data"
    A = petsc_io.read_mat(d / "f_A")
    Is = [petsc_io.read_mat(d / f) for f in ("I_1", "I_2", "I_3")]
    mg = orc.MG(Is, pre=3, post=3, masks=True, fix_semidefinite=True)
    orc.srand(1)
    assert mg.update(A) == 0
    m = mg.mask(3)
    rows_offdiag = np.array([np.any((A.col[A.rowptr[i]:A.rowptr[i + 1]] != i)
                                    & (np.abs(A.val[A.rowptr[i]:A.rowptr[i + 1]]) > 1e-10))
                             for i in range(A.shape[0])])
    assert np.array_equal(m, rows_offdiag.astype(float))
    # coarse mask: 0 where any fine Dirichlet row feeds the coarse row through R
    R = Is[2].to_scipy().T.tocsr()
    feeds = (abs(R) > 0).astype(float) @ (1 - m)
    assert np.array_equal(mg.mask(2), (feeds == 0).astype(float))


def test_cg_mg_on_reference_fixture(golden):
    # PetscLinearSolverTest::petsc_cg_mg asserts approxeq(A*x, rhs, 1e-6)
    d = golden / "utopia_data"
    A = petsc_io.read_mat(d / "f_A")
    rhs = petsc_io.read_vec(d / "f_rhs")
    Is = [petsc_io.read_mat(d / f) for f in ("I_1", "I_2", "I_3")]
    mg = orc.MG(Is, pre=3, post=3, masks=True, fix_semidefinite=True)
    orc.srand(1)
    mg.update(A)
    x, st = mg.cg_solve(rhs, x0=rhs, atol=1e-15, max_it=200)
    assert np.abs(A.to_scipy() @ x - rhs).max() < 1e-6
    x2 = mg.apply(rhs, max_it=40, atol=1e-13, rtol=1e-13)
    assert np.abs(A.to_scipy() @ x2 - rhs).max() < 1e-6


def _qp_bruteforce(Ad, b, lb):
    n = b.size
    best = None
    for act in itertools.product([False, True], repeat=n):
        act = np.array(act)
        x = lb.copy()
        free = ~act
        if free.any():
            rhs = b[free] - Ad[np.ix_(free, act)] @ lb[act]
            x[free] = np.linalg.solve(Ad[np.ix_(free, free)], rhs)
        g = Ad @ x - b
        if np.all(x >= lb - 1e-12) and np.all(g[act] >= -1e-12) and np.all(np.abs(g[free]) < 1e-10):
            best = x
    return best


def test_mprgp_and_pgs_against_bruteforce_qp():
    A = ua.gen_stencil(ua.STENCIL_2D9, 3, 3, shift=1e-3)
    Ad = A.to_scipy().toarray()
    b = ua.gen_uniform(7, 9)
    lb = -0.05 - 0.15 * ua.gen_uniform(8, 9, 0, 1)
    x_ref = _qp_bruteforce(Ad, b, lb)
    assert x_ref is not None
    x, st = orc.mprgp_solve(A, b, lb=lb, atol=1e-12, rtol=0, max_it=1000)
    assert st["converged"] and np.abs(x - x_ref).max() < 1e-10
    x2, st2 = orc.pgs_solve(A, b, lb=lb, stol=1e-14, max_it=10000)
    assert np.abs(x2 - x_ref).max() < 1e-10


def test_mprgp_kkt_certificate():
    n = 40
    A = ua.gen_stencil(ua.STENCIL_2D9, n, n, shift=1e-3)
    b = ua.gen_uniform(1, n * n)
    lb = -0.05 - 0.15 * ua.gen_uniform(2, n * n, 0, 1)
    x, st = orc.mprgp_solve(A, b, lb=lb, atol=1e-9, rtol=0, max_it=20000)
    assert st["converged"]
    g = A.to_scipy() @ x - b
    act = np.abs(x - lb) < 1e-14
    assert np.all(x >= lb)
    assert np.all(g[act] >= -1e-9) and np.abs(g[~act]).max() < 1e-8
    assert act.sum() > 0 and (~act).sum() > 0


def _reference_gs_local_sweeps(A, r, dinv, order_fwd, symmetric=True):
    """GaussSeidel::local_sweeps (solvers/smoothers/utopia_GaussSeidel.hpp:133-183) in
    plain Python floats, visiting rows in order_fwd forward and reversed backward."""
    n = A.shape[0]
    c = [0.0] * n
    passes = [order_fwd, list(reversed(order_fwd))] if symmetric else [order_fwd]
    for order in passes:
        for i in order:
            s_ = float(r[i])
            for e in range(A.rowptr[i], A.rowptr[i + 1]):
                j = int(A.col[e])
                if j != i:
                    s_ -= float(A.val[e]) * c[j]
            c[i] = float(dinv[i]) * s_
    return np.array(c)


def test_gauss_seidel_lexicographic_equals_reference_sweep():
    """With one colour per row (color[i] = i) the multicolour sweep is the
    reference's lexicographic symmetric Gauss-Seidel, bit for bit."""
    A = ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, 7, 5, seed=4)
    n = A.shape[0]
    b = ua.gen_uniform(3, n)
    d = A.to_scipy().diagonal()
    dinv = 1.0 / d
    x = orc.gs_smooth(A, dinv, np.arange(n), n, b, sweeps=1)
    ref = _reference_gs_local_sweeps(A, b, dinv, list(range(n)))  # x0 = 0: r = b, x = c
    assert np.array_equal(x, ref)


def test_gauss_seidel_red_black_equals_colour_ordered_sweep():
    """Red-black colouring of the 5-point grid: the sweep equals the reference
    sweep over the rows listed colour by colour (order inside a colour is free)."""
    nx, ny = 6, 7
    A = ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, nx, ny, seed=8)
    n = A.shape[0]
    color = np.array([(i % nx + i // nx) % 2 for i in range(n)], np.int32)
    b = ua.gen_uniform(5, n)
    dinv = 1.0 / A.to_scipy().diagonal()
    x = orc.gs_smooth(A, dinv, color, 2, b, sweeps=1)
    order = [i for i in range(n) if color[i] == 0] + [i for i in range(n) if color[i] == 1]
    # backward pass: colour 1 then colour 0 (rows of one colour are independent)
    ref = _reference_gs_local_sweeps(A, b, dinv, order)
    assert np.array_equal(x, ref)
    # two sweeps reduce the residual of the SPD system
    x2 = orc.gs_smooth(A, dinv, color, 2, b, sweeps=2)
    Asp = A.to_scipy()
    assert np.linalg.norm(b - Asp @ x2) < np.linalg.norm(b - Asp @ x) < np.linalg.norm(b)


def test_gauss_seidel_on_diagonal_matrix_is_jacobi():
    n = 9
    d = 1.0 + np.arange(n)
    A = ua.Csr(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d, (n, n))
    b = ua.gen_uniform(1, n)
    x = orc.gs_smooth(A, 1.0 / d, np.zeros(n, np.int32), 1, b, sweeps=1)
    assert np.array_equal(x, (1.0 / d) * b)


def test_mprgp_iterations_are_rounding_sensitive():
    """At 256^2 the MPRGP iteration count is not a rounding-stable quantity:
    perturbing b by one ulp changes the oracle's count by more than 5 %, while
    the solution (and its active set) stays put. Iteration counts of MPRGP runs
    with different summation orders are therefore compared with a band, the
    solution and the active set exactly (up to ties)."""
    n = 256
    A = ua.gen_stencil(ua.STENCIL_2D9, n, n, shift=1e-3)
    b = ua.gen_uniform(4, n * n)
    lb = -0.05 - 0.15 * ua.gen_uniform(5, n * n, 0, 1)
    its, xs = [], []
    for eps in (0.0, 1e-16):
        bb = b * (1 + eps * np.sign(np.sin(np.arange(n * n))))
        ua.srand(1)
        x, st = orc.mprgp_solve(A, bb, lb=lb, atol=1e-9, rtol=0, max_it=100000)
        its.append(st["iterates"])
        xs.append(x)
    assert abs(its[1] - its[0]) > 0.05 * its[0]
    assert np.linalg.norm(xs[1] - xs[0]) <= 1e-7 * np.linalg.norm(xs[0])
    tie = np.abs(xs[0] - lb) < 1e-9
    assert np.all(((np.abs(xs[0] - lb) < 1e-12) == (np.abs(xs[1] - lb) < 1e-12)) | tie)


def _reference_qp_problem(n=20):
    """QPSolverTestProblem (tests/test_problems/utopia_QPSolverTestProblem.hpp:28-100):
    1D Laplacian with identity boundary rows (explicit zeros kept), rhs 1 with
    zero ends, upper bound 30, max_it 40 n."""
    rp, col, val = [0], [], []
    for i in range(n):
        if i == 0:
            col += [0, 1]
            val += [1.0, 0.0]
        elif i == n - 1:
            col += [n - 2, n - 1]
            val += [0.0, 1.0]
        else:
            col += [i - 1, i, i + 1]
            val += [-1.0, 2.0, -1.0]
        rp.append(len(col))
    A = ua.Csr(np.array(rp, np.int32), np.array(col, np.int32), np.array(val), (n, n))
    b = np.ones(n)
    b[0] = b[-1] = 0.0
    return A, b, np.full(n, 30.0)


def test_projected_gradient_solves_reference_qp_problem():
    A, b, ub = _reference_qp_problem()
    x, st = orc.pg_solve(A, b, ub=ub, max_it=40 * 20)
    assert st["reason"] == 4  # converged on the step norm before max_it
    r = b - A.to_scipy() @ x
    act = np.abs(x - ub) < 1e-12
    assert act.sum() == 4 and np.all(x <= ub)
    assert np.abs(r[~act]).max() < 1e-10 and r[act].min() > 0  # variational inequality


def test_projected_gradient_first_step_and_steepest_descent_limit():
    A = ua.gen_stencil(ua.STENCIL_2D5, 9, 7)
    b = ua.gen_uniform(4, 63)
    x, _ = orc.pg_solve(A, b, max_it=0, stol=0)  # one step with alpha = 1 (utopia_ProjectedGradient.hpp:69)
    assert np.array_equal(x, b - A.to_scipy() @ np.zeros(63))
    x, _ = orc.pg_solve(A, b, max_it=4000, stol=0)
    xs = np.linalg.solve(A.to_scipy().toarray(), b)
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)


def test_projected_cg_first_step_is_steepest_descent_and_far_bounds_are_inert():
    A = ua.gen_stencil(ua.STENCIL_2D5, 12, 10)
    n = A.shape[0]
    b = ua.gen_uniform(3, n)
    x1, _ = orc.pcg_solve(A, b, max_it=0, stol=0)
    Asp = A.to_scipy()
    a = (b @ b) / (b @ (Asp @ b))
    assert np.allclose(x1, a * b, rtol=1e-14, atol=0)
    xu, _ = orc.pcg_solve(A, b, max_it=200, stol=0)
    xb, _ = orc.pcg_solve(A, b, lb=np.full(n, -1e10), ub=np.full(n, 1e10), max_it=200, stol=0)
    assert np.array_equal(xu, xb)
    xs = np.linalg.solve(Asp.toarray(), b)
    assert np.linalg.norm(xu - xs) <= 1e-7 * np.linalg.norm(xs)


def test_projected_cg_is_rounding_sensitive():
    """On the reference's QP test problem the ProjectedConjugateGradient iteration
    runs to max_it without meeting stol, and one ulp of b moves its result by
    more than 0.1 % after 800 iterations but by < 1e-12 after 10."""
    A, b, ub = _reference_qp_problem()
    bb = b * (1 + 1e-16 * np.sign(np.sin(np.arange(20) + 1)))
    x0, s0 = orc.pcg_solve(A, b, ub=ub, max_it=800)
    x1, _ = orc.pcg_solve(A, bb, ub=ub, max_it=800)
    assert s0["reason"] == -5
    assert np.linalg.norm(x1 - x0) > 1e-3 * np.linalg.norm(x0)
    x0, _ = orc.pcg_solve(A, b, ub=ub, max_it=10, stol=0)
    x1, _ = orc.pcg_solve(A, bb, ub=ub, max_it=10, stol=0)
    assert np.linalg.norm(x1 - x0) < 1e-12 * np.linalg.norm(x0)


def test_jones_plassmann_colouring_is_proper_and_small():
    for A in (ua.gen_stencil(ua.STENCIL_2D5, 17, 11), ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 9, 8, 7, seed=1),
              ua.gen_stencil(ua.STENCIL_2D9, 13, 12)):
        color, nc = orc.jp_color(A)
        G = A.to_scipy().tocoo()
        off = G.row != G.col
        assert not np.any(color[G.row[off]] == color[G.col[off]])
        deg = np.diff(A.rowptr).max() - 1
        assert color.min() == 0 and color.max() == nc - 1 and nc <= deg + 1


def _monotone_mg_problem(n_levels=6, n_coarse=50):
    """QPSolverTest::monotone_mg_test (tests/utopia_QPSolverTest.cpp:718-777) on
    MultiLevelTestProblem1D<Poisson1D> (tests/test_problems/large_scale_benchmark/
    utopia_MultilevelTestProblem1D.hpp:56-193, utopia_Poisson1D.hpp:101-290, problem
    type 2): dofs (n_coarse + 1) 2^l + ..., H = tridiag(-1, 2, -1) / h with identity
    boundary rows (explicit zeros kept), x0 = 0, rhs = g = H x0 - f = -10 h inside,
    interpolations with rows 2j -> (j, 1) and 2j+1 -> (j, .5), (j+1, .5), the finest with
    its two boundary entries zeroed (remove_bc), bounds [-0.8, 200].
    Returns A, b, interps (coarse to fine), lb, ub."""
    dofs = [n_coarse + 1]
    for _ in range(1, n_levels):
        dofs.append((dofs[-1] - 1) * 2 + 1)
    interps = []
    for l in range(n_levels - 1):
        nc, nf = dofs[l], dofs[l + 1]
        rp, col, val = [0], [], []
        for k in range(nf):
            j = k // 2
            if k % 2 == 0:
                col.append(j)
                val.append(1.0)
            elif j + 1 < nc:
                col += [j, j + 1]
                val += [0.5, 0.5]
            rp.append(len(col))
        val = np.array(val)
        if l == n_levels - 2:
            val[rp[0]] = 0.0            # I(0, 0)
            val[rp[nf - 1]] = 0.0       # I(last_h, last_H)
        interps.append(ua.Csr(np.array(rp, np.int32), np.array(col, np.int32), val, (nf, nc)))
    n = dofs[-1]
    h = 1.0 / (n - 1)
    rp, col, val = [0], [], []
    for i in range(n):
        if i == 0:
            col += [0, 1]
            val += [1.0, 0.0]
        elif i == n - 1:
            col += [n - 2, n - 1]
            val += [0.0, 1.0]
        else:
            col += [i - 1, i, i + 1]
            val += [-1.0 / h, 2.0 / h, -1.0 / h]
        rp.append(len(col))
    A = ua.Csr(np.array(rp, np.int32), np.array(col, np.int32), np.array(val), (n, n))
    b = np.full(n, -10.0 * h)
    b[0] = b[-1] = 0.0
    return A, b, interps, np.full(n, -0.8), np.full(n, 200.0)


def _obstacle_1d_bruteforce(Ad, b, lb):
    """The unique minimiser of 1/2 x'Ax - b'x over x >= lb for a tridiagonal M-matrix:
    its contact set is an interval (discrete maximum principle), so every interval
    [a, e) is tried and the one meeting the KKT conditions is returned."""
    n = b.size
    for a in range(n + 1):
        for e in range(a, n + 1):
            act = np.zeros(n, bool)
            act[a:e] = True
            x = lb.copy()
            free = ~act
            x[free] = np.linalg.solve(Ad[np.ix_(free, free)], b[free] - Ad[np.ix_(free, act)] @ lb[act])
            g = Ad @ x - b
            if np.all(x >= lb - 1e-12) and np.all(g[act] >= -1e-10) and np.all(np.abs(g[free]) < 1e-10):
                return x, act
    return None, None


def test_monotone_mg_solves_reference_obstacle_problem():
    A, b, Is, lb, ub = _monotone_mg_problem(3, 10)
    Ad = A.to_scipy().toarray()
    x_ref, act = _obstacle_1d_bruteforce(Ad, b, lb)
    assert x_ref is not None and 0 < act.sum() < b.size - 2
    for levels in (2, 3):
        x, st = orc.mmg_solve(A, b, Is[3 - levels:], lb=lb, ub=ub, pre=5, post=5, active_set_tol=1e-15, atol=1e-12,
                              rtol=0, stol=0, max_it=40)
        assert st["converged"] and st["iterates"] < 40
        assert np.abs(x - x_ref).max() < 1e-10


def _colour_sgs(A, x, b, steps):
    """Symmetric Gauss-Seidel in the Jones-Plassmann colour order as two sparse
    triangular solves of the colour-permuted system: forward (D + L) c1 = r,
    backward (D + U) c = r - L c1; x += c."""
    color, _ = orc.jp_color(A)
    perm = np.argsort(color, kind="stable")
    M = A.to_scipy()[perm][:, perm].tocsr()
    D = sp.diags(M.diagonal())
    L, U = sp.tril(M, -1), sp.triu(M, 1)
    for _ in range(steps):
        r = (b - A.to_scipy() @ x)[perm]
        c1 = spla.spsolve_triangular((D + L).tocsr(), r, lower=True)
        c = spla.spsolve_triangular((D + U).tocsr(), r - L @ c1, lower=False)
        x = x.copy()
        x[perm] += c
    return x


def test_monotone_mg_unconstrained_step_is_two_grid_cycle():
    """Without bounds one MonotoneMultigrid step is pre+1 SGS steps, the exact
    Galerkin coarse correction I (I'AI)^-1 I'r, and post+1 SGS steps."""
    A, b, Is, _, _ = _monotone_mg_problem(3, 10)
    x0 = np.sin(np.arange(b.size))
    x, _ = orc.mmg_solve(A, b, Is[-1:], x0=x0, pre=2, post=1, atol=0, rtol=0, stol=0, max_it=2)
    S = A.to_scipy()
    I = Is[-1].to_scipy().toarray()
    y = _colour_sgs(A, x0, b, 3)
    AH = I.T @ S.toarray() @ I
    y = y + I @ np.linalg.solve(AH, I.T @ (b - S @ y))
    y = _colour_sgs(A, y, b, 2)
    assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max()


def test_monotone_mg_coarse_correction_is_truncated_on_the_active_set():
    """pre = post = 0: one step is the coarse correction through the finest
    interpolation with the rows of active nodes zeroed; active nodes do not move."""
    A, b, Is, lb, ub = _monotone_mg_problem(3, 10)
    n = b.size
    x0 = 0.1 * np.cos(np.arange(n))
    act = (np.arange(n) % 3 == 1) | ((np.arange(n) > 15) & (np.arange(n) < 24))
    x0[act] = lb[act]
    x, _ = orc.mmg_solve(A, b, Is[-1:], lb=lb, ub=ub, x0=x0, pre=0, post=0, atol=0, rtol=0, stol=0, max_it=2)
    S = A.to_scipy().toarray()
    It = Is[-1].to_scipy().toarray()
    It[act] = 0.0
    AH = It.T @ S @ It
    dead = np.abs(np.diag(AH)) < 1e-12
    assert dead.any()  # coarse nodes whose children are all active
    AH[dead, dead] += 1.0
    y = x0 + It @ np.linalg.solve(AH, It.T @ (b - S @ x0))
    assert np.array_equal(x[act], x0[act])
    assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max()


def test_monotone_mg_coarse_space_multigrid_tends_to_exact_coarse_solve():
    A, b, Is, lb, ub = _monotone_mg_problem(5, 6)
    kw = dict(lb=lb, ub=ub, pre=1, post=1, atol=0, rtol=0, stol=0, max_it=2)
    x2, _ = orc.mmg_solve(A, b, Is[-1:], **kw)
    x5, _ = orc.mmg_solve(A, b, Is, coarse_space_it=80, **kw)
    x5_1, _ = orc.mmg_solve(A, b, Is, coarse_space_it=1, **kw)
    assert np.abs(x5 - x2).max() <= 1e-8 * np.abs(x2).max()
    assert np.abs(x5_1 - x2).max() > 1e-6 * np.abs(x2).max()


def _amg_reference_problem(n=100):
    """AMGTest::test_1D (tests/utopia_AMGTest.cpp:69-121): assemble_laplacian_1D with
    bc = true (identity rows 0 and n-1 holding one value,
    tests/test_problems/utopia_assemble_laplacian_1D.hpp:11-44), b = 2 with zero ends."""
    rp, col, val = [0], [], []
    for i in range(n):
        if i in (0, n - 1):
            col.append(i)
            val.append(1.0)
        else:
            col += [i - 1, i, i + 1]
            val += [-1.0, 2.0, -1.0]
        rp.append(len(col))
    A = ua.Csr(np.array(rp, np.int32), np.array(col, np.int32), np.array(val), (n, n))
    b = np.full(n, 2.0)
    b[0] = b[-1] = 0.0
    return A, b


def test_agglomerate_pairs_the_1d_laplacian_and_galerkin_is_the_laplacian_again():
    """On the 1D Laplacian each open row claims its right neighbour (its left one is taken),
    the single-value boundary rows join no aggregate, and P^T A P of the pairs is again
    tridiag(-1, 2, -1)."""
    A, _ = _amg_reference_problem(100)
    rp, col, val, shape = orc.agglomerate(A)
    assert shape == (100, 49)
    assert rp[1] == 0 and rp[100] - rp[99] == 0  # rows 0 and 99: no entry
    assert np.array_equal(col, (np.arange(1, 99) - 1) // 2) and np.all(val == 1.0)
    Ps, As = orc.amg_transfers(A, 2)
    Ac = As[0].to_scipy().toarray()
    m = 49
    assert np.array_equal(Ac, 2 * np.eye(m) - np.eye(m, k=1) - np.eye(m, k=-1))


def test_agglomerate_hand_computed_five_point_3x3():
    """5-point Laplacian on a 3x3 grid (all couplings strong): row 0 claims 1 and 3,
    row 2 claims 5, row 4 claims 7; rows 6 and 8 find every neighbour taken, their largest
    entry is their diagonal, so they stay open and end as singletons."""
    A = ua.gen_stencil(ua.STENCIL_2D5, 3, 3)
    assert np.array_equal(A.to_scipy().diagonal(), np.full(9, 4.0))
    rp, col, val, shape = orc.agglomerate(A)
    assert shape == (9, 5)
    assert np.array_equal(col, [0, 0, 1, 0, 2, 1, 3, 2, 4]) and np.array_equal(np.diff(rp), np.ones(9))
    # bmax caps the claims: with bmax = 1 row 0 takes only 1, row 2 takes 5, row 3 takes 4,
    # row 6 takes 7 and row 8 is left over
    rp1, col1, _, shape1 = orc.agglomerate(A, bmax=1)
    assert np.array_equal(col1, [0, 0, 1, 2, 2, 1, 3, 3, 4]) and shape1 == (9, 5)
    # weight: with weight 1 nothing is strong (|a_ij| = max|a_ij|), every row a singleton
    rpw, colw, _, shapew = orc.agglomerate(A, weight=1.0)
    assert shapew == (9, 9) and np.array_equal(colw, np.arange(9))


def test_cg_amg_solves_reference_amg_problem():
    """AMGTest asserts |b - A x|_1 < 1e-5 after AlgebraicMultigrid (6 levels) on this
    problem; here with GaussSeidel smoothing inside CG."""
    A, b = _amg_reference_problem(100)
    Ps, As = orc.amg_transfers(A, 6)
    assert [P.shape for P in Ps] == [(7, 4), (13, 7), (25, 13), (49, 25), (100, 49)]
    mg = orc.MG(Ps, smoother=2, pre=3, post=3, masks=False)
    mg.update(A)
    for l in range(1, 6):
        mg.set_colors(l, *orc.jp_color(As[l]))
    x, st = mg.cg_solve(b, atol=1e-9, max_it=200)
    assert st["converged"] and np.abs(b - A.to_scipy() @ x).sum() < 1e-5


def test_cg_solves_reference_hexa_fem_system(golden):
    """PetscTest::petsc_matlab_connection (tests/petsc/utopia_PetscTest.cpp:396-421):
    K x = M rhs on the 10x10x10 hexahedral stiffness / mass matrices, |M rhs - K x| < 1e-6
    (the reference solves it with BiCGStab; CG converges on it too)."""
    d = golden / "utopia_data"
    K = petsc_io.read_mat(d / "K_hexa_10x10x10_3D")
    M = petsc_io.read_mat(d / "M_hexa_10x10x10_3D")
    rhs = petsc_io.read_vec(d / "RHS_10x10x10_hexa_3D")
    assert K.shape == M.shape == (1331, 1331)
    b = orc.spmv(M, rhs)
    assert np.abs(b - M.to_scipy() @ rhs).max() <= 1e-14 * (abs(M.to_scipy()) @ np.abs(rhs)).max()
    for pc in (0, 1):
        x, st = orc.cg_solve(K, b, precond=pc, atol=1e-10, rtol=0, stol=0, max_it=2000)
        assert st["converged"] and np.linalg.norm(b - K.to_scipy() @ x) < 1e-6


# ---- BlockAgglomerate (solvers/multilevel/agglomerate/utopia_BlockAgglomerate.hpp) ----

def _block_parents(P, B):
    """Aggregate of each block row, read from a block interpolation."""
    rp, col, val, shape = P
    assert np.array_equal(rp, np.arange(len(rp))) and np.all(val == 1.0)
    assert np.array_equal(col % B, np.arange(len(col)) % B)
    return col[::B] // B


def _scalar_parents(S):
    rp, col, val, shape = orc.agglomerate(S)
    assert np.array_equal(rp, np.arange(len(rp))) and np.all(val == 1.0)  # every row aggregated, a_max > 0
    return col


@pytest.mark.parametrize("B", [2, 3, 4])
def test_block_agglomerate_of_kron_is_scalar_agglomerate(B):
    """On S (x) M the block graph is S's graph and every block weight is S_IJ M_cc, so the
    block pass must reproduce the (pinned) scalar pass on S, expanded to B sub-rows."""
    S = ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, 13, 11, seed=2)
    M = np.diag(np.arange(1.0, B + 1)) + 0.25 * (np.ones((B, B)) - np.eye(B))
    A = ua.Csr.from_scipy(sp.kron(S.to_scipy(), sp.csr_matrix(M)).tocsr())
    P = orc.agglomerate_blocks(A, B)
    assert P[3] == (A.shape[0], (max(_scalar_parents(S)) + 1) * B)
    assert np.array_equal(_block_parents(P, B), _scalar_parents(S))


def test_block_agglomerate_weight_component():
    """Two components with different couplings: the pass follows the component asked for
    (default B - 1), and the zero-weight blocks of the other component's pattern change
    nothing. A pass reading the wrong component fails: the two scalar results differ."""
    S0 = ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, 12, 10, seed=1).to_scipy()
    S1 = ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, 12, 10, seed=7).to_scipy()
    E0, E1 = sp.csr_matrix(np.diag([1.0, 0.0])), sp.csr_matrix(np.diag([0.0, 1.0]))
    A = ua.Csr.from_scipy((sp.kron(S0, E0) + sp.kron(S1, E1)).tocsr())
    p0, p1 = _scalar_parents(ua.Csr.from_scipy(S0)), _scalar_parents(ua.Csr.from_scipy(S1))
    assert not np.array_equal(p0, p1)
    assert np.array_equal(_block_parents(orc.agglomerate_blocks(A, 2, component=0), 2), p0)
    assert np.array_equal(_block_parents(orc.agglomerate_blocks(A, 2), 2), p1)


def test_block_agglomerate_hand_example():
    """Three block nodes, weights (component 1) w00=10, w01=w10=-1, w12=w21=-0.1, wii=10.
    a_max excludes the diagonal block (utopia_BlockAgglomerate.hpp:55-63): node 0 opens
    aggregate 0 and takes node 1 (|-1| > 1/3); node 2 (a_max 0.1) finds node 1 taken and
    itself the strongest block, so it stays open and becomes aggregate 1. Counting the
    diagonal in a_max would leave all three nodes singletons."""
    W = np.array([[10.0, -1.0, 0.0], [-1.0, 10.0, -0.1], [0.0, -0.1, 10.0]])
    other = np.array([[3.0, 2.0, 0.0], [2.0, 3.0, 2.0], [0.0, 2.0, 3.0]])  # component 0: ignored
    A = ua.Csr.from_scipy(sp.csr_matrix(np.kron(other, np.diag([1.0, 0.0])) + np.kron(W, np.diag([0.0, 1.0]))))
    P = orc.agglomerate_blocks(A, 2)
    assert P[3] == (6, 4)
    assert list(_block_parents(P, 2)) == [0, 0, 1]
