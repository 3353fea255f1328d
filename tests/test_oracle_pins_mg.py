"""Pins of the oracle's Multigrid cycle, PowerMethod schedule and Jacobi /
PointJacobi smoothers by closed forms evaluated in dense numpy (no step of the
oracle's own recurrences is restated here):

  oracle function              pinned by
  or_mg_apply (V / W cycle)    the two-/three-level cycle composed from closed-form
                               smoothers (Chebyshev: residual polynomial T_nu via the
                               Chebyshev recurrence; Jacobi: I - G^nu with G the
                               iteration matrix), masked Galerkin restriction, exact
                               coarse solve, on the reference's Q1 hierarchy fixture
  or_mg_apply (line search)    the same with c += alpha c_I, alpha = c_I.r_R / c_I.A c_I
  or_mg_apply (FULL_CYCLE)     one standard cycle, max_it forced to 1 (every standard
                               cycle restarts from c = 0)
  or_power_method              lambda_k = |d^k v0| / |d^(k-1) v0| on diag(d), v0 the glibc
                               rand() draws, stopping index from |lambda_k - lambda_(k-1)| < tol
  or_jacobi_smooth             x_nu = x* + G^nu (x0 - x*), G = I - D^+ A (pseudo-inverse 0)
  or_point_jacobi_smooth       x_nu = sum_(k<nu) G^k c + G^nu x0, G = -D'^-1 (A - D),
                               D' = D with zeros replaced by ones (check_indef)

Each pin is shown to fail on a plausible mistake: the planted variant of the
closed form differs from the oracle by orders of magnitude more than the pin's
tolerance.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
import utopia_amd as ua
from utopia_amd import petsc_io

RAND_MAX = 2147483647  # glibc


def _libc_rand(n):
    libc = ctypes.CDLL(None)
    return np.array([libc.rand() for _ in range(n)], dtype=np.float64)


def cheb_residual_poly(A, v, nu, emax, emin):
    """P_nu(A) v, P_nu(z) = T_nu((avg - z)/half) / T_nu(avg/half), by the Chebyshev
    polynomial recurrence T_(k+1) = 2 X T_k - T_(k-1)."""
    if nu == 0:
        return v.copy()
    avg, half = (emax + emin) / 2, (emax - emin) / 2

    def X(w):
        return (avg * w - A @ w) / half

    t0, t1 = v, X(v)
    for _ in range(1, nu):
        t0, t1 = t1, 2 * X(t1) - t0
    return t1 / np.polynomial.chebyshev.Chebyshev.basis(nu)(avg / half)


def affine_iterate(G, c, x0, nu):
    """x_nu of x_(k+1) = G x_k + c: sum_(k<nu) G^k c + G^nu x0."""
    x = np.linalg.matrix_power(G, nu) @ x0
    for k in range(nu):
        x = x + np.linalg.matrix_power(G, k) @ c
    return x


def smooth_model(kind, A, b, x0, nu, eigs=None):
    n = A.shape[0]
    if kind == "cheb":
        # Chebyshev3level::solve updates x before its first convergence check
        # (utopia_Chebyshev3level.hpp:137-160): sweeps(0) is one step
        nu = max(nu, 1)
        e = b - A @ x0
        return x0 + np.linalg.solve(A, e - cheb_residual_poly(A, e, nu, *eigs))
    d = np.diag(A).copy()
    if kind == "jacobi":
        dinv = np.where(np.abs(d) > 0, 1.0 / np.where(d == 0, 1, d), 0.0)
        return affine_iterate(np.eye(n) - dinv[:, None] * A, dinv * b, x0, nu)
    if kind == "point":
        dinv = np.where(d != 0, 1.0 / np.where(d == 0, 1, d), 1.0)
        return affine_iterate(-dinv[:, None] * (A - np.diag(d)), dinv * b, x0, nu)
    if nu == 0:  # Jacobi / PointJacobi: for (it < sweeps) sweep
        return x0.copy()
    raise ValueError(kind)


class DenseHierarchy:
    """Galerkin levels P^T A P (dense), masks, transfers: level 0 coarsest."""

    def __init__(self, A, interps, fix_semidefinite=True, masks=True):
        L = len(interps) + 1
        self.P = [I.to_scipy().toarray() for I in interps]
        self.R = [P.T.copy() for P in self.P]
        self.A = [None] * L
        self.A[L - 1] = A.to_scipy().toarray()
        for l in range(L - 1, 0, -1):
            Ac = self.P[l - 1].T @ self.A[l] @ self.P[l - 1]
            if fix_semidefinite:
                i = np.where(np.abs(np.diag(Ac)) < 1e-12)[0]
                Ac[i, i] += 1.0
            self.A[l - 1] = Ac
        self.mask = [None] * L
        if masks:
            Af = self.A[L - 1]
            off = np.abs(Af - np.diag(np.diag(Af))) > np.finfo(float).eps * 1e6
            self.mask[L - 1] = off.any(axis=1).astype(float)
            for l in range(L - 1, 0, -1):
                feeds = (np.abs(self.R[l - 1]) > np.finfo(float).eps * 1e6).astype(float) @ (1 - self.mask[l])
                self.mask[l - 1] = (feeds == 0).astype(float)


def model_cycle(H, l, r, kind, pre, post, mg_type=1, eigs=None, line_search=False, plant=None):
    if l == 0:
        return np.linalg.solve(H.A[0], r)
    A = H.A[l]
    c = np.zeros_like(r)
    for _ in range(mg_type):
        c = smooth_model(kind, A, r, c, pre, eigs and eigs[l])
        rR = r - A @ c
        src = r if plant == "restrict_r" else rR
        rc = H.R[l - 1] @ src
        if H.mask[l - 1] is not None and plant != "no_mask":
            rc = H.mask[l - 1] * rc
        cc = model_cycle(H, l - 1, rc, kind, pre, post, mg_type, eigs, line_search, plant)
        cI = H.P[l - 1] @ cc
        if line_search:
            alpha = (cI @ rR) / ((A @ cI) @ cI)
            c = c + (cI if alpha <= 0 else alpha * cI)
        else:
            c = c + cI
        if plant != "no_post":
            c = smooth_model(kind, A, r, c, post, eigs and eigs[l])
    return c


def _fixture(golden, nlev):
    d = golden / "utopia_data"
    A = petsc_io.read_mat(d / "f_A")
    rhs = petsc_io.read_vec(d / "f_rhs")
    Is = [petsc_io.read_mat(d / f) for f in ("I_1", "I_2", "I_3")][3 - (nlev - 1):]
    return A, rhs, Is


def _oracle_mg(A, Is, kind, pre, post, mg_type, H, **kw):
    smoother = {"cheb": 0, "jacobi": 1, "point": 3}[kind]
    mg = orc.MG(Is, smoother=smoother, pre=pre, post=post, mg_type=mg_type, masks=True, fix_semidefinite=True, **kw)
    orc.srand(1)
    assert mg.update(A) == 0
    eigs = None
    if kind == "cheb":
        eigs = {}
        for l in range(1, len(Is) + 1):
            lam = np.abs(np.linalg.eigvals(H.A[l])).max()
            eigs[l] = (1.2 * lam, 0.06 * 1.2 * lam)
            mg.set_eigs(l, *eigs[l])
    return mg, eigs


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nlev,kind,pre,post,mg_type", [
    (2, "cheb", 2, 2, 1), (3, "cheb", 2, 2, 1), (3, "cheb", 3, 1, 2), (3, "jacobi", 2, 2, 1),
    (3, "jacobi", 1, 3, 2), (3, "point", 2, 2, 1), (2, "point", 3, 3, 1), (3, "cheb", 0, 2, 1),
    (3, "jacobi", 0, 2, 1),
])
def test_mg_cycle_equals_dense_composition(golden, nlev, kind, pre, post, mg_type):
    A, rhs, Is = _fixture(golden, nlev)
    H = DenseHierarchy(A, Is)
    mg, eigs = _oracle_mg(A, Is, kind, pre, post, mg_type, H)
    for l in range(nlev):  # the oracle's Galerkin levels and masks are the dense ones
        rp, col, val, shape = mg.level(l)
        Al = sp.csr_matrix((val, col, rp), shape=shape).toarray()
        assert np.abs(Al - H.A[l]).max() <= 1e-14 * np.abs(H.A[l]).max()
        assert np.array_equal(mg.mask(l), H.mask[l])
    x = mg.apply(rhs, max_it=1)
    ref = model_cycle(H, nlev - 1, rhs, kind, pre, post, mg_type, eigs)
    assert _rel(x, ref) < 1e-11
    # planted mistakes in the cycle composition are far outside that tolerance
    for plant in ("no_post", "restrict_r", "no_mask"):
        if plant == "no_post" and post == 0:
            continue
        if plant == "restrict_r" and pre == 0 and kind != "cheb":  # c = 0, so r_R = r
            continue
        bad = model_cycle(H, nlev - 1, rhs, kind, pre, post, mg_type, eigs, plant=plant)
        assert _rel(bad, ref) > 1e-6, plant
    if mg_type == 1:
        w = model_cycle(H, nlev - 1, rhs, kind, pre, post, 2, eigs)
        assert _rel(w, ref) > 1e-6


@pytest.mark.parametrize("kind", ["cheb", "jacobi"])
def test_mg_line_search_cycle_equals_dense_composition(golden, kind):
    A, rhs, Is = _fixture(golden, 3)
    H = DenseHierarchy(A, Is)
    mg, eigs = _oracle_mg(A, Is, kind, 2, 2, 1, H, use_line_search=True)
    x = mg.apply(rhs, max_it=1)
    ref = model_cycle(H, 2, rhs, kind, 2, 2, 1, eigs, line_search=True)
    assert _rel(x, ref) < 1e-11
    assert _rel(model_cycle(H, 2, rhs, kind, 2, 2, 1, eigs), ref) > 1e-6


def test_mg_full_cycle_is_one_standard_cycle(golden):
    # Multigrid::full_cycle restricts, solves and interpolates, but every standard_cycle it
    # runs starts from c = 0 (utopia_Multigrid.hpp:264-268) and FULL_CYCLE sets max_it(1)
    A, rhs, Is = _fixture(golden, 3)
    H = DenseHierarchy(A, Is)
    full, eigs = _oracle_mg(A, Is, "cheb", 2, 2, 1, H, cycle_type=3, v_cycle_repetition=2)
    std, _ = _oracle_mg(A, Is, "cheb", 2, 2, 1, H)
    xf = full.apply(rhs, max_it=7)
    assert np.array_equal(xf, std.apply(rhs, max_it=1))
    assert _rel(xf, model_cycle(H, 2, rhs, "cheb", 2, 2, 1, eigs)) < 1e-11
    # a full cycle whose intermediate cycles kept their interpolated guess would differ
    assert _rel(std.apply(rhs, max_it=2), xf) > 1e-6


@pytest.mark.parametrize("tol,max_it", [(1e-2, 30), (1e-4, 30), (1e-9, 6), (1e-9, 11)])
def test_power_method_schedule_on_diagonal(tol, max_it):
    n = 64
    d = 1.0 + 2.0 * np.linspace(0, 1, n) ** 0.5  # slow convergence: lambda_k moves at every k
    A = ua.Csr.from_scipy(sp.diags(d).tocsr())
    orc.srand(5)
    v0 = _libc_rand(n) / RAND_MAX + 1
    lam = [0.0] + [np.linalg.norm(d ** k * v0) / np.linalg.norm(d ** (k - 1) * v0) for k in range(1, 60)]
    K = next(k for k in range(1, 60) if abs(lam[k - 1] - lam[k]) < tol or k > max_it)
    orc.srand(5)
    got = orc.power_method(A, tol, max_it)
    assert abs(got - lam[K]) <= 1e-14 * lam[K]
    # one iteration more or less, or a start from ones, is visible far above that
    for k in (K - 1, K + 1):
        assert abs(lam[k] - lam[K]) > 1e-9 * lam[K]
    ones = [np.linalg.norm(d ** K) / np.linalg.norm(d ** (K - 1))]
    assert abs(ones[0] - lam[K]) > 1e-9 * lam[K]


def _small_spd(n=12, seed=3):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    return M @ M.T + n * np.eye(n)


@pytest.mark.parametrize("nu", [1, 2, 5])
def test_jacobi_and_point_jacobi_closed_form(nu):
    Ad = _small_spd()
    n = Ad.shape[0]
    A = ua.Csr.from_scipy(sp.csr_matrix(Ad))
    b = ua.gen_uniform(4, n)
    x0 = ua.gen_uniform(5, n)
    xs = np.linalg.solve(Ad, b)
    G = np.eye(n) - Ad / np.diag(Ad)[:, None]
    ref = xs + np.linalg.matrix_power(G, nu) @ (x0 - xs)
    xj = orc.jacobi_smooth(A, b, x0, nu)
    xp = orc.jacobi_smooth(A, b, x0, nu, point=True)
    assert _rel(xj, ref) < 1e-12 and _rel(xp, ref) < 1e-12
    # one sweep fewer is far off
    assert _rel(xs + np.linalg.matrix_power(G, nu - 1) @ (x0 - xs), ref) > 1e-6


def test_jacobi_and_point_jacobi_zero_diagonal_rows():
    # row 2 has no diagonal entry: Jacobi's pseudo-inverse (tol 0) leaves x_2 as it is,
    # PointJacobi's check_indef sets d_inv = 1, so x_2 = b_2 - sum_j a_2j x_j
    Ad = _small_spd(8, 7)
    Ad[2, 2] = 0.0
    S = sp.csr_matrix(Ad)
    S.eliminate_zeros()
    A = ua.Csr.from_scipy(S)
    n = 8
    b, x0 = ua.gen_uniform(8, n), ua.gen_uniform(9, n)
    d = np.diag(Ad)
    for nu in (1, 3):
        dj = np.where(d != 0, 1 / np.where(d == 0, 1, d), 0.0)
        refj = affine_iterate(np.eye(n) - dj[:, None] * Ad, dj * b, x0, nu)
        dp = np.where(d != 0, 1 / np.where(d == 0, 1, d), 1.0)
        refp = affine_iterate(-dp[:, None] * (Ad - np.diag(d)), dp * b, x0, nu)
        assert _rel(orc.jacobi_smooth(A, b, x0, nu), refj) < 1e-12
        assert _rel(orc.jacobi_smooth(A, b, x0, nu, point=True), refp) < 1e-12
        # PointJacobi without check_indef (d_inv 0 on that row) would zero x_2
        dz = np.where(d != 0, 1 / np.where(d == 0, 1, d), 0.0)
        bad = affine_iterate(-dz[:, None] * (Ad - np.diag(d)), dz * b, x0, nu)
        assert _rel(bad, refp) > 1e-6
    assert orc.jacobi_smooth(A, b, x0, 1)[2] == x0[2]


def test_block_mg_fixture_reference_assertion(golden):
    """PetscLinearSolverTest::test_block_mg (tests/petsc/utopia_PetscLinearSolverTest.cpp:343-383)
    on the reference's data/mg fixture: Multigrid with symmetric Gauss-Seidel smoothing (PETSc's
    SOR with omega 1; multicolour order here), pre/post 3, fix_semidefinite_operators, 20
    cycles; the reference asserts approxeq(rhs, A x, 1e-6)."""
    d = golden / "utopia_data"
    A = petsc_io.read_mat(d / "mg_A.bin")
    I = petsc_io.read_mat(d / "mg_I.bin")
    rhs = petsc_io.read_vec(d / "mg_rhs.bin")
    assert A.shape == (410, 410) and I.shape == (410, 126)
    mg = orc.MG([I], smoother=2, pre=3, post=3, masks=True, fix_semidefinite=True)
    orc.srand(1)
    assert mg.update(A) == 0
    for l in range(2):
        rp, col, val, shape = mg.level(l)
        color, nc = orc.jp_color(ua.Csr(rp, col, val, shape))
        mg.set_colors(l, color, nc)
    x = mg.apply(rhs, max_it=20, atol=1e-15, rtol=1e-15)
    assert np.abs(rhs - A.to_scipy() @ x).max() <= 1e-6
