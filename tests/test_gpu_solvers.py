"""Solver-level parity on the GPU (through the solver C ABI) against the CPU
oracle on the same seeded inputs.

Bars (the benchmark workloads): iteration counts within +-1 of the oracle, FP64
solution relative error <= 1e-10 for MG-PCG at rtol 1e-10 (1e-6 for the
Jacobi-PCG run at rtol 1e-8, whose ~400 iterations amplify dot-product
rounding), mixed precision within 1e-8 of the FP64 solution, QP active sets
identical up to ties at 1e-12.
"""
import numpy as np
import scipy.sparse as sp
import pytest

import oracle as orc
import utopia_amd as ua
from utopia_amd import petsc_io

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_jacobi_pcg_poisson_256(ctx):
    """2D 5-point 256x256, Jacobi-PCG, rtol 1e-8 (atol = 1e-8 |b|), x0 = 0."""
    A = ua.gen_stencil(ua.STENCIL_2D5, 256, 256)
    b = ua.gen_uniform(1, A.shape[0])
    atol = 1e-8 * np.linalg.norm(b)
    cg = ua.ConjugateGradient(ctx, "jacobi", atol=atol, rtol=0, stol=0, max_it=5000)
    x = cg.solve(A, b)
    st = cg.status
    xo, so = orc.cg_solve(A, b, precond=True, atol=atol, rtol=0, stol=0, max_it=5000)
    print(f"jacobi-pcg 256^2: gpu it={st['iterates']} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}")
    assert abs(st["iterates"] - so["iterates"]) <= 1
    assert np.linalg.norm(b - A.to_scipy() @ x) <= 1.01 * atol
    assert rel(x, xo) <= 1e-6
    assert st["reason"] == ua.CONVERGED_FNORM_ABS


def test_unpreconditioned_cg(ctx):
    A = ua.gen_stencil(ua.STENCIL_2D5, 40, 30)
    b = ua.gen_uniform(2, A.shape[0])
    atol = 1e-10 * np.linalg.norm(b)
    cg = ua.ConjugateGradient(ctx, None, atol=atol, rtol=0, stol=0, max_it=2000)
    x = cg.solve(A, b)
    xo, so = orc.cg_solve(A, b, precond=False, atol=atol, rtol=0, stol=0, max_it=2000)
    assert abs(cg.status["iterates"] - so["iterates"]) <= 1
    assert rel(x, xo) <= 1e-9


def _mg_pcg(ctx, kind, n, levels, seed, pre=2, post=2, mixed=False, refine=None):
    A = ua.gen_stencil(kind, n, n, n, seed=seed)
    b = ua.gen_uniform(seed + 100, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, levels)
    atol = 1e-10 * np.linalg.norm(b)
    ua.srand(1)
    cg = ua.ConjugateGradient(ctx, "multigrid_mixed" if mixed else "multigrid", atol=atol, rtol=0, stol=0,
                              max_it=500, pre_smoothing_steps=pre, post_smoothing_steps=post)
    if refine is not None:
        cg.set_param("refine", refine)
    cg.set_transfer_operators(interps)
    x = cg.solve(A, b)
    ua.srand(1)
    mg = orc.MG(interps, pre=pre, post=post)
    mg.update(A)
    xo, so = mg.cg_solve(b, atol=atol, max_it=500)
    return A, b, atol, cg, x, mg, xo, so


@pytest.mark.parametrize("kind,n,levels", [(ua.STENCIL_3D7, 32, 4), (ua.STENCIL_3D7_VARCOEF, 32, 4),
                                           (ua.STENCIL_3D7, 24, 3)])
def test_mg_pcg_fp64_matches_oracle(ctx, kind, n, levels):
    A, b, atol, cg, x, mg, xo, so = _mg_pcg(ctx, kind, n, levels, seed=3)
    st = cg.status
    print(f"mg-pcg {n}^3 kind={kind}: gpu it={st['iterates']} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}")
    assert abs(st["iterates"] - so["iterates"]) <= 1
    assert np.linalg.norm(b - A.to_scipy() @ x) <= 1.01 * atol
    assert rel(x, xo) <= 1e-10


def test_mg_hierarchy_masks_and_eigs_match_oracle(ctx):
    n = 16
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=5)
    interps = ua.mg_hierarchy(n, n, n, 3)
    ua.srand(1)
    s = ua.Multigrid(ctx, pre_smoothing_steps=2, post_smoothing_steps=2, max_it=1)
    s.set_transfer_operators(interps)
    s.update(A)
    ua.srand(1)
    mg = orc.MG(interps, pre=2, post=2)
    mg.update(A)
    for l in range(3):
        G = s.level(l)
        rp, col, val, shape = mg.level(l)
        assert np.array_equal(G.rowptr, rp) and np.array_equal(G.col, col) and np.array_equal(G.val, val)
        assert np.array_equal(s.mask(l), mg.mask(l))
    for l in (1, 2):
        emax, emin = mg.eigs(l)
        assert abs(s.info(f"eig_max_{l}") - emax) <= 1e-10 * emax
        assert abs(s.info(f"eig_min_{l}") - emin) <= 1e-10 * emin


@pytest.mark.parametrize("pre,post,mg_type", [(2, 2, 1), (3, 3, 1), (1, 2, 2)])
def test_mg_cycles_match_oracle_with_pinned_eigs(ctx, pre, post, mg_type):
    n = 16
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=6)
    b = ua.gen_uniform(7, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, 3)
    ua.srand(1)
    s = ua.Multigrid(ctx, pre_smoothing_steps=pre, post_smoothing_steps=post, mg_type=mg_type, max_it=3,
                     atol=0, rtol=0, stol=0)
    s.set_transfer_operators(interps)
    s.update(A)
    ua.srand(1)
    mg = orc.MG(interps, pre=pre, post=post, mg_type=mg_type)
    mg.update(A)
    for l in (1, 2):
        emax, emin = mg.eigs(l)
        s.set_level_eigs(l, emax, emin)
    x = s.apply(b)
    xo = mg.apply(b, max_it=3, atol=0, rtol=0)
    # identical cycles up to the coarse solve (explicit inverse vs triangular solves)
    assert rel(x, xo) <= 1e-12


@pytest.mark.parametrize("pre,post", [(1, 1), (2, 2)])
def test_mg_gauss_seidel_smoother_matches_oracle(ctx, pre, post):
    """Multicolour symmetric GaussSeidel smoothing in the V-cycle: GPU cycles equal
    the oracle's with the same colouring (the GPU's, read back per level)."""
    n = 16
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=6)
    b = ua.gen_uniform(7, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, 3)
    s = ua.Multigrid(ctx, smoother=2, pre_smoothing_steps=pre, post_smoothing_steps=post, max_it=3, atol=0, rtol=0,
                     stol=0)
    s.set_transfer_operators(interps)
    s.update(A)
    mg = orc.MG(interps, smoother=2, pre=pre, post=post)
    mg.update(A)
    for l in (1, 2):
        color, nc = s.colors(l)
        assert 2 <= nc <= 27 and color.min() == 0 and color.max() == nc - 1
        G = s.level(l).to_scipy().tocoo()
        off = G.row != G.col
        assert not np.any(color[G.row[off]] == color[G.col[off]])  # proper colouring
        mg.set_colors(l, color, nc)
    x = s.apply(b)
    xo = mg.apply(b, max_it=3, atol=0, rtol=0)
    assert rel(x, xo) <= 1e-12


def test_cg_mg_gauss_seidel_converges_like_oracle(ctx):
    n = 24
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=2)
    b = ua.gen_uniform(102, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, 3)
    atol = 1e-10 * np.linalg.norm(b)
    cg = ua.ConjugateGradient(ctx, "multigrid", atol=atol, rtol=0, stol=0, max_it=300, smoother=2,
                              pre_smoothing_steps=1, post_smoothing_steps=1)
    cg.set_transfer_operators(interps)
    x = cg.solve(A, b)
    mg = orc.MG(interps, smoother=2, pre=1, post=1)
    mg.update(A)
    for l in (1, 2):
        mg.set_colors(l, *cg.colors(l))
    xo, so = mg.cg_solve(b, atol=atol, max_it=300)
    assert abs(cg.status["iterates"] - so["iterates"]) <= 1
    assert rel(x, xo) <= 1e-10
    assert np.linalg.norm(b - A.to_scipy() @ x) <= 1.0001 * atol


def test_autotuned_mg_pcg_matches_default_plans(ctx):
    """Plans change the partial-sum layout of the fused dots, not the row sums:
    the autotuned solve agrees with the default one to rounding."""
    n = 24
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=4)
    b = ua.gen_uniform(104, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, 3)
    atol = 1e-10 * np.linalg.norm(b)
    xs = []
    for tune in (0, 3):
        ua.srand(1)
        cg = ua.ConjugateGradient(ctx, "multigrid_mixed", atol=atol, rtol=0, stol=0, max_it=300, autotune=tune,
                                  pre_smoothing_steps=2, post_smoothing_steps=2)
        cg.set_transfer_operators(interps)
        xs.append((cg.solve(A, b), cg.status["iterates"]))
    assert abs(xs[0][1] - xs[1][1]) <= 1
    assert rel(xs[1][0], xs[0][0]) <= 1e-10


def test_cg_mg_reference_fixture(ctx, golden):
    """PetscLinearSolverTest::petsc_cg_mg on utopia/data/laplace/matrices_for_petsc."""
    d = golden / "utopia_data"
    A = petsc_io.read_mat(d / "f_A")
    rhs = petsc_io.read_vec(d / "f_rhs")
    Is = [petsc_io.read_mat(d / f) for f in ("I_1", "I_2", "I_3")]
    ua.srand(1)
    cg = ua.ConjugateGradient(ctx, "multigrid", atol=1e-15, rtol=1e-18, stol=1e-18, max_it=200,
                              fix_semidefinite_operators=1, must_generate_masks=1)
    cg.set_transfer_operators(Is)
    x = cg.solve(A, rhs, rhs)
    assert np.abs(A.to_scipy() @ x - rhs).max() < 1e-6
    ua.srand(1)
    mg = orc.MG(Is, masks=True, fix_semidefinite=True)
    mg.update(A)
    xo, so = mg.cg_solve(rhs, x0=rhs, atol=1e-15, max_it=200)
    assert abs(cg.status["iterates"] - so["iterates"]) <= 1


def test_mixed_precision_mg_pcg(ctx):
    n = 32
    A, b, atol, cg, x, mg, xo, so = _mg_pcg(ctx, ua.STENCIL_3D7_VARCOEF, n, 4, seed=9, mixed=True)
    st = cg.status
    err = rel(x, xo)
    total = cg.info("total_iterations")
    print(f"mixed mg-pcg {n}^3: iterations {st['iterates']} (total {total}, refinements "
          f"{cg.info('refinements')}), fp64 oracle {so['iterates']}, rel.err vs fp64 {err:.2e}")
    assert np.linalg.norm(b - A.to_scipy() @ x) <= 1.01 * atol
    assert err <= 1e-8
    assert total <= so["iterates"] + 1 + 5


def test_mixed_preconditioner_is_fp32_accurate(ctx):
    """The FP32 V-cycle agrees with the FP64 one to single precision, not better."""
    n = 32
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=9)
    r = ua.gen_uniform(77, A.shape[0])
    out = {}
    for kind in ("multigrid", "multigrid_mixed"):
        ua.srand(1)
        cg = ua.ConjugateGradient(ctx, kind, pre_smoothing_steps=2, post_smoothing_steps=2)
        cg.set_transfer_operators(ua.mg_hierarchy(n, n, n, 4))
        cg.update(A)
        out[kind] = cg.precondition(r)
    d = rel(out["multigrid_mixed"], out["multigrid"])
    print(f"mixed vs fp64 V-cycle: rel. difference {d:.2e}")
    assert 1e-9 < d < 1e-5


def _qp(n, seed):
    A = ua.gen_stencil(ua.STENCIL_2D9, n, n, shift=1e-3)
    b = ua.gen_uniform(seed, n * n)
    lb = -0.05 - 0.15 * ua.gen_uniform(seed + 1, n * n, 0, 1)
    return A, b, lb


def _kkt(A, b, x, lb, tol):
    g = A.to_scipy() @ x - b
    act = np.abs(x - lb) < 1e-12
    assert np.all(x >= lb)
    assert np.all(g[act] >= -tol)
    assert np.abs(g[~act]).max() <= tol
    return act


def test_mprgp_matches_oracle(ctx):
    A, b, lb = _qp(64, 11)
    q = ua.MPRGP(ctx, atol=1e-9, rtol=0, max_it=100000)
    q.update(A)
    q.set_box_constraints(lb, None)
    ua.srand(1)
    x = q.apply(b)
    ua.srand(1)
    xo, so = orc.mprgp_solve(A, b, lb=lb, atol=1e-9, rtol=0, max_it=100000)
    st = q.status
    print(f"mprgp 64^2: gpu it={st['iterates']} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}")
    act = _kkt(A, b, x, lb, 1e-8)
    act_o = np.abs(xo - lb) < 1e-12
    ties = np.abs(x - lb) < 1e-9
    assert np.all((act == act_o) | ties)
    assert rel(x, xo) <= 1e-8


def test_mprgp_256_active_set_matches_oracle(ctx):
    A, b, lb = _qp(256, 4)
    q = ua.MPRGP(ctx, atol=1e-9, rtol=0, max_it=100000)
    q.update(A)
    q.set_box_constraints(lb, None)
    x = q.apply(b)  # MPRGP's power method starts from ones (utopia_MPRGP.hpp:356)
    xo, so = orc.mprgp_solve(A, b, lb=lb, atol=1e-9, rtol=0, max_it=100000)
    print(f"mprgp 256^2: gpu it={q.status['iterates']} oracle it={so['iterates']}")
    act = _kkt(A, b, x, lb, 1e-8)
    act_o = np.abs(xo - lb) < 1e-12
    ties = np.abs(x - lb) < 1e-9
    assert np.all((act == act_o) | ties)
    # MPRGP's step choices at this size depend on rounding: a 1-ulp change of b moves
    # the oracle's own count by 10-20 % (test_oracle_pins::test_mprgp_iterations_are_rounding_sensitive)
    assert abs(q.status["iterates"] - so["iterates"]) <= 0.25 * so["iterates"]
    assert rel(x, xo) <= 1e-7


def _reference_qp_problem(n=20):
    from test_oracle_pins import _reference_qp_problem as mk
    return mk(n)


def test_projected_gradient_matches_oracle(ctx):
    """ProjectedGradient on the reference's QP test problem (max_it 40 n) and on a
    2D problem with lower bounds: same iterations and solution as the oracle."""
    A, b, ub = _reference_qp_problem()
    q = ua.ProjectedGradient(ctx, max_it=800)
    q.update(A)
    q.set_box_constraints(None, ub)
    x = q.apply(b)
    xo, so = orc.pg_solve(A, b, ub=ub, max_it=800)
    assert q.status["iterates"] == so["iterates"] and q.status["reason"] == so["reason"] == 4
    assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)
    A2, b2, lb2 = _qp(24, 5)
    q2 = ua.ProjectedGradient(ctx, max_it=20000, stol=1e-10)
    q2.update(A2)
    q2.set_box_constraints(lb2, None)
    x2 = q2.apply(b2)
    xo2, so2 = orc.pg_solve(A2, b2, lb=lb2, max_it=20000, stol=1e-10)
    assert abs(q2.status["iterates"] - so2["iterates"]) <= 20
    assert rel(x2, xo2) <= 1e-8


def test_projected_cg_matches_oracle(ctx):
    """ProjectedConjugateGradient: its iterates are rounding-chaotic on the reference
    QP problem (test_oracle_pins::test_projected_cg_is_rounding_sensitive), so early
    iterates are compared tightly and the max_it outcome within that sensitivity."""
    A, b, ub = _reference_qp_problem()
    for max_it, tol in ((10, 1e-11), (800, 5e-2)):
        q = ua.ProjectedConjugateGradient(ctx, max_it=max_it, stol=0 if max_it == 10 else 1e-11)
        q.update(A)
        q.set_box_constraints(None, ub)
        x = q.apply(b)
        xo, so = orc.pcg_solve(A, b, ub=ub, max_it=max_it, stol=0 if max_it == 10 else 1e-11)
        assert q.status["iterates"] == so["iterates"] and q.status["reason"] == so["reason"]
        assert np.linalg.norm(x - xo) <= tol * np.linalg.norm(xo), max_it
        assert np.all(x <= ub)


def test_pgs_converges_to_the_qp_solution(ctx):
    A, b, lb = _qp(24, 21)
    q = ua.ProjectedGaussSeidel(ctx, stol=1e-13, max_it=20000)
    q.update(A)
    q.set_box_constraints(lb, None)
    x = q.apply(b)
    xo, _ = orc.mprgp_solve(A, b, lb=lb, atol=1e-12, rtol=0, max_it=100000)
    assert 2 <= q.info("ncolors") <= 9
    assert rel(x, xo) <= 1e-9
    _kkt(A, b, x, lb, 1e-8)


def test_pgs_unconstrained_is_symmetric_gauss_seidel(ctx):
    A = ua.gen_stencil(ua.STENCIL_2D5, 20, 20)
    b = ua.gen_uniform(4, 400)
    q = ua.ProjectedGaussSeidel(ctx, stol=1e-14, max_it=50000)
    q.update(A)
    x = q.apply(b)
    import scipy.sparse.linalg as spla
    xs = spla.spsolve(A.to_scipy().tocsc(), b)
    assert rel(x, xs) <= 1e-10


def test_error_paths_report_status(ctx):
    # indefinite coarse operator -> Cholesky failure reported as an error code, no crash
    A = ua.gen_stencil(ua.STENCIL_3D7, 8, 8, 8)
    A = ua.Csr(A.rowptr, A.col, -A.val, A.shape)
    s = ua.Multigrid(ctx, max_it=1)
    s.set_transfer_operators(ua.mg_hierarchy(8, 8, 8, 2))
    with pytest.raises(ua.HipError):
        s.update(A)
    # NaN right-hand side -> CG stops on the NaN step length (utopia_ConjugateGradient_impl.hpp:241-244)
    A = ua.gen_stencil(ua.STENCIL_2D5, 10, 10)
    b = ua.gen_uniform(1, 100)
    b[5] = np.nan
    cg = ua.ConjugateGradient(ctx, "jacobi", atol=1e-10, max_it=50)
    cg.solve(A, b)
    assert cg.status["iterates"] == 0


def test_solver_by_registry_name(ctx):
    """QPSolverRegistry / OmniLinearSolver names select the same solvers as the type codes."""
    A, b, lb = _qp(32, 5)
    x1 = ua.MPRGP(ctx, atol=1e-9, rtol=0, max_it=10000)
    x1.update(A)
    x1.set_box_constraints(lb, None)
    ua.srand(1)
    y1 = x1.apply(b)
    q = ua.Solver.by_name(ctx, "mprgp", atol=1e-9, rtol=0, max_it=10000)
    q.update(A)
    q.set_box_constraints(lb, None)
    ua.srand(1)
    y2 = q.apply(b)
    assert np.array_equal(y1, y2)
    cg = ua.Solver.by_name(ctx, "cg_jacobi", backend="hip", atol=1e-10, rtol=0, max_it=500)
    xc = cg.solve(A, b)
    assert np.linalg.norm(b - A.to_scipy() @ xc) <= 1e-10 * 1.01
    for backend, name in (("any", "ssnewton_x"), ("petsc", "cg")):
        with pytest.raises(ua.HipError, match="not found"):
            ua.Solver.by_name(ctx, name, backend=backend)


def _monotone_mg_problem(n_levels=6, n_coarse=50):
    from test_oracle_pins import _monotone_mg_problem as mk
    return mk(n_levels, n_coarse)


@pytest.mark.parametrize("levels", [6, 2])
def test_monotone_multigrid_matches_oracle_on_reference_problem(ctx, levels):
    """QPSolverTest::monotone_mg_test: 1601 dofs, bounds [-0.8, 200], 5/5 projected
    Gauss-Seidel steps, active-set tol 1e-15, max_it 40 (GaussSeidel coarse smoothing,
    dense coarse solve)."""
    A, b, Is, lb, ub = _monotone_mg_problem(6, 50)
    Is = Is[6 - levels:]
    q = ua.MonotoneMultigrid(ctx, pre_smoothing_steps=5, post_smoothing_steps=5, active_set_tol=1e-15, max_it=40,
                             atol=1e-9, rtol=1e-9, stol=1e-11)
    q.set_transfer_operators(Is)
    q.update(A)
    q.set_box_constraints(lb, ub)
    x = q.apply(b)
    xo, so = orc.mmg_solve(A, b, Is, lb=lb, ub=ub, pre=5, post=5, active_set_tol=1e-15, max_it=40)
    st = q.status
    print(f"mmg {levels} levels: gpu it={st['iterates']} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}")
    assert abs(st["iterates"] - so["iterates"]) <= 1
    assert rel(x, xo) <= 1e-9
    act, act_o = np.abs(x - lb) < 1e-12, np.abs(xo - lb) < 1e-12
    assert act.sum() > 0 and np.array_equal(act, act_o)
    assert np.all(x >= lb)
    assert q.info("active_set_changes") >= 1


def test_monotone_multigrid_2d_obstacle_matches_oracle(ctx):
    """2D nine-point obstacle problem on a 3-level hierarchy."""
    n = 64
    A = ua.gen_stencil(ua.STENCIL_2D9, n, n, shift=1e-3)
    b = ua.gen_uniform(21, n * n)
    lb = -0.05 - 0.15 * ua.gen_uniform(22, n * n, 0, 1)
    Is = ua.mg_hierarchy(n, n, 1, 3)
    q = ua.MonotoneMultigrid(ctx, pre_smoothing_steps=3, post_smoothing_steps=3, coarse_space_it=2, max_it=200,
                             atol=1e-10, rtol=0, stol=0)
    q.set_transfer_operators(Is)
    q.update(A)
    q.set_box_constraints(lb, None)
    x = q.apply(b)
    xo, so = orc.mmg_solve(A, b, Is, lb=lb, pre=3, post=3, coarse_space_it=2, atol=1e-10, rtol=0, stol=0, max_it=200)
    st = q.status
    print(f"mmg 2d 64^2: gpu it={st['iterates']} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}")
    assert st["reason"] > 0 and abs(st["iterates"] - so["iterates"]) <= 1
    assert rel(x, xo) <= 1e-8
    _kkt(A, b, x, lb, 1e-8)


def _amg_reference_problem(n=100):
    from test_oracle_pins import _amg_reference_problem as mk
    return mk(n)


def _same_csr(a, rp, col, val):
    return np.array_equal(a.rowptr, rp) and np.array_equal(a.col, col) and np.array_equal(a.val, val)


@pytest.mark.parametrize("case", ["ref1d", "varcoef32"])
def test_amg_hierarchy_matches_oracle_exactly(ctx, case):
    """Agglomerate transfers and the P^T A P operators of every level, bit for bit."""
    if case == "ref1d":
        A, _ = _amg_reference_problem(100)
        L = 6
    else:
        A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 32, 32, 32, seed=3)
        L = 5
    s = ua.AlgebraicMultigrid(ctx, n_levels=L, smoother=2)
    s.update(A)
    Ps, As = orc.amg_transfers(A, L)
    assert int(s.info("n_levels")) == L
    for l in range(L - 1):
        P = s.transfer(l)
        assert P.shape == Ps[l].shape and _same_csr(P, Ps[l].rowptr, Ps[l].col, Ps[l].val)
    for l in range(L):
        G = s.level(l)
        assert _same_csr(G, As[l].rowptr, As[l].col, As[l].val)


def _oracle_amg(A, L, pre, post, colors, block_size=1):
    Ps, As = orc.amg_transfers(A, L, block_size=block_size)
    mg = orc.MG(Ps, smoother=2, pre=pre, post=post, masks=False)
    mg.update(A)
    for l in range(1, L):
        mg.set_colors(l, *colors(l))
    return mg


def test_cg_amg_reference_problem_matches_oracle(ctx):
    """AMGTest::test_1D's problem (|b - A x|_1 < 1e-5) with CG over a 6-level
    AlgebraicMultigrid (GaussSeidel 3/3)."""
    A, b = _amg_reference_problem(100)
    cg = ua.ConjugateGradient(ctx, "amg", n_levels=6, smoother=2, pre_smoothing_steps=3, post_smoothing_steps=3,
                              atol=1e-9, rtol=0, stol=0, max_it=200)
    x = cg.solve(A, b)
    mg = _oracle_amg(A, 6, 3, 3, cg.colors)
    xo, so = mg.cg_solve(b, atol=1e-9, max_it=200)
    print(f"cg-amg 1d: gpu it={cg.status['iterates']} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}")
    assert abs(cg.status["iterates"] - so["iterates"]) <= 1
    assert rel(x, xo) <= 1e-10
    assert np.abs(b - A.to_scipy() @ x).sum() < 1e-5


def test_amg_solver_cycles_match_oracle(ctx):
    A, b = _amg_reference_problem(100)
    s = ua.AlgebraicMultigrid(ctx, n_levels=6, smoother=2, pre_smoothing_steps=3, post_smoothing_steps=3, max_it=20,
                              atol=0, rtol=0, stol=0)
    x = s.solve(A, b)
    mg = _oracle_amg(A, 6, 3, 3, s.colors)
    xo = mg.apply(b, max_it=20, atol=0, rtol=0)
    assert s.status["iterates"] == 20
    assert rel(x, xo) <= 1e-12


def test_cg_amg_varcoef_matches_oracle(ctx):
    n = 32
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=3)
    b = ua.gen_uniform(5, n ** 3)
    atol = 1e-10 * np.linalg.norm(b)
    cg = ua.ConjugateGradient(ctx, "amg", n_levels=5, smoother=2, pre_smoothing_steps=2, post_smoothing_steps=2,
                              atol=atol, rtol=0, stol=0, max_it=500)
    x = cg.solve(A, b)
    mg = _oracle_amg(A, 5, 2, 2, cg.colors)
    xo, so = mg.cg_solve(b, atol=atol, max_it=500)
    print(f"cg-amg 32^3: gpu it={cg.status['iterates']} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}")
    assert abs(cg.status["iterates"] - so["iterates"]) <= 1
    assert rel(x, xo) <= 1e-10
    assert np.linalg.norm(b - A.to_scipy() @ x) <= 1.01 * atol


@pytest.mark.parametrize("pc", [None, "jacobi"])
def test_cg_reference_hexa_fem_system(ctx, golden, pc):
    """The reference's K x = M rhs fixture (petsc_matlab_connection): SpMV bit-exact,
    CG iterations and solution equal to the oracle's, residual below the reference's 1e-6."""
    d = golden / "utopia_data"
    K = petsc_io.read_mat(d / "K_hexa_10x10x10_3D")
    M = petsc_io.read_mat(d / "M_hexa_10x10x10_3D")
    rhs = petsc_io.read_vec(d / "RHS_10x10x10_hexa_3D")
    Md = ua.DeviceCsr(ctx, M)
    xd, yd = ua.DeviceArray.from_numpy(ctx, rhs), ua.DeviceArray(ctx, M.shape[0])
    Md.spmv(xd, yd)
    b = yd.numpy()
    assert np.array_equal(b, orc.spmv(M, rhs))
    cg = ua.ConjugateGradient(ctx, pc, atol=1e-10, rtol=0, stol=0, max_it=2000)
    x = cg.solve(K, b)
    xo, so = orc.cg_solve(K, b, precond=0 if pc is None else 1, atol=1e-10, rtol=0, stol=0, max_it=2000)
    print(f"hexa fem cg pc={pc}: gpu it={cg.status['iterates']} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}")
    assert abs(cg.status["iterates"] - so["iterates"]) <= 1
    assert rel(x, xo) <= 1e-10
    assert np.linalg.norm(b - K.to_scipy() @ x) < 1e-6


def test_level_plan_formats_reported(ctx):
    """format_{A,R,P}_<l> / format32_{A,R,P}_<l> report the SpMV plan of each level: with
    the automatic rules the 7-point fine operator is DIA in both precisions and the
    trilinear transfers are value-indexed."""
    n = 24
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=4)
    cg = ua.ConjugateGradient(ctx, "multigrid_mixed", atol=1e-8, rtol=0, stol=0, max_it=50)
    cg.set_transfer_operators(ua.mg_hierarchy(n, n, n, 3))
    cg.update(A)
    f = int(cg.info("n_levels")) - 1
    assert cg.info(f"format_A_{f}") == ua.FMT_DIA and cg.info(f"format32_A_{f}") == ua.FMT_DIA
    for w in "RP":
        assert cg.info(f"format_{w}_{f}") == ua.FMT_VI and cg.info(f"format32_{w}_{f}") == ua.FMT_VI
    with pytest.raises(ua.HipError):
        cg.info("format_R_0")  # the coarsest level has no transfers
    for bad in ("level_nnz_99", "level_rows_-1", "eig_max_x", "transfer_nnz_7", "format_A_"):
        with pytest.raises(ua.HipError):
            cg.info(bad)  # an error code, not an exception through the C ABI
    plain = ua.ConjugateGradient(ctx, "multigrid", atol=1e-8, rtol=0, stol=0, max_it=50)
    plain.set_transfer_operators(ua.mg_hierarchy(n, n, n, 3))
    plain.update(A)
    with pytest.raises(ua.HipError):
        plain.info(f"format32_A_{f}")  # no FP32 V-cycle


def _vector_problem(nx, ny, seed):
    """Two unknowns per node: S (x) [[2, .5], [.5, 1]] (SPD, coupled within and across nodes)."""
    S = ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, nx, ny, seed=seed).to_scipy()
    return ua.Csr.from_scipy(sp.kron(S, sp.csr_matrix([[2.0, 0.5], [0.5, 1.0]])).tocsr())


def test_block_amg_hierarchy_matches_oracle_exactly(ctx):
    """AlgebraicMultigrid with block_size 2 (BlockAgglomerate on every level): transfers
    and Galerkin operators bit for bit against the oracle's block pass."""
    A = _vector_problem(40, 36, 3)
    L = 4
    s = ua.AlgebraicMultigrid(ctx, n_levels=L, smoother=2, block_size=2)
    s.update(A)
    Ps, As = orc.amg_transfers(A, L, block_size=2)
    for l in range(L - 1):
        P = s.transfer(l)
        assert P.shape == Ps[l].shape and _same_csr(P, Ps[l].rowptr, Ps[l].col, Ps[l].val)
        assert P.shape[1] % 2 == 0
    for l in range(L):
        G = s.level(l)
        assert _same_csr(G, As[l].rowptr, As[l].col, As[l].val)
    odd = ua.gen_stencil(ua.STENCIL_2D5, 5, 3)  # 15 rows: not a multiple of the block size
    with pytest.raises(ua.HipError):
        ua.AlgebraicMultigrid(ctx, n_levels=2, block_size=2).update(odd)


def test_cg_block_amg_matches_oracle(ctx):
    A = _vector_problem(40, 36, 3)
    b = ua.gen_uniform(8, A.shape[0])
    atol = 1e-10 * np.linalg.norm(b)
    cg = ua.ConjugateGradient(ctx, "amg", n_levels=4, smoother=2, block_size=2, pre_smoothing_steps=2,
                              post_smoothing_steps=2, atol=atol, rtol=0, stol=0, max_it=500)
    x = cg.solve(A, b)
    mg = _oracle_amg(A, 4, 2, 2, cg.colors, block_size=2)
    xo, so = mg.cg_solve(b, atol=atol, max_it=500)
    print(f"cg-block-amg: gpu it={cg.status['iterates']} oracle it={so['iterates']} rel.err={rel(x, xo):.2e}")
    assert abs(cg.status["iterates"] - so["iterates"]) <= 1
    assert rel(x, xo) <= 1e-10
    assert np.linalg.norm(b - A.to_scipy() @ x) <= 1.01 * atol
