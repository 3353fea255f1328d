"""Multigrid options on the GPU against the oracle: Jacobi and PointJacobi
smoothing (as Multigrid cycles and as the CG preconditioner), zero
pre-smoothing steps, use_line_search and FULL_CYCLE
(solvers/multilevel/utopia_Multigrid.hpp:252-415,
solvers/smoothers/utopia_Jacobi.hpp:89-140, utopia_PointJacobi.hpp:61-142).

Bars: cycles equal the oracle's to 1e-12 relative (the explicit coarse inverse
vs triangular solves and the fused dots of the line search differ in rounding
only); MG-PCG within +-1 iteration and 1e-10 of the oracle's solution.
"""
import numpy as np
import pytest

import oracle as orc
import utopia_amd as ua
from utopia_amd import petsc_io

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _mg_pair(ctx, n, levels, smoother, pre, post, mg_type=1, seed=6, max_it=3, **opts):
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=seed)
    b = ua.gen_uniform(seed + 1, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, levels)
    ua.srand(1)
    s = ua.Multigrid(ctx, smoother=smoother, pre_smoothing_steps=pre, post_smoothing_steps=post, mg_type=mg_type,
                     max_it=max_it, atol=0, rtol=0, stol=0,
                     **{k: (1 if v is True else v) for k, v in opts.items()})
    s.set_transfer_operators(interps)
    s.update(A)
    ua.srand(1)
    mg = orc.MG(interps, smoother=smoother, pre=pre, post=post, mg_type=mg_type, **opts)
    mg.update(A)
    if smoother == 0:
        for l in range(1, levels):
            s.set_level_eigs(l, *mg.eigs(l))
    return A, b, s, mg


@pytest.mark.parametrize("smoother,pre,post,mg_type", [(1, 2, 2, 1), (1, 1, 3, 2), (3, 2, 2, 1), (3, 3, 1, 2),
                                                       (1, 3, 3, 1)])
def test_mg_jacobi_smoothers_match_oracle(ctx, smoother, pre, post, mg_type):
    A, b, s, mg = _mg_pair(ctx, 16, 3, smoother, pre, post, mg_type)
    x = s.apply(b)
    xo = mg.apply(b, max_it=3, atol=0, rtol=0)
    print(f"mg smoother={smoother} pre={pre} post={post} mg_type={mg_type}: rel.err={rel(x, xo):.2e}")
    assert rel(x, xo) <= 1e-12


@pytest.mark.parametrize("smoother", [0, 1, 3])
def test_mg_zero_pre_smoothing_matches_oracle(ctx, smoother):
    """pre_smoothing_steps = 0: each cycle pass starts from c = 0 (Chebyshev3level still takes
    one step for sweeps(0), Jacobi and PointJacobi none); several cycles in a row."""
    A, b, s, mg = _mg_pair(ctx, 16, 3, smoother, 0, 2, max_it=4)
    x = s.apply(b)
    xo = mg.apply(b, max_it=4, atol=0, rtol=0)
    assert rel(x, xo) <= 1e-12
    # again on the same objects (no state kept from the previous apply)
    assert rel(s.apply(b), xo) <= 1e-12


@pytest.mark.parametrize("smoother,pre", [(0, 0), (1, 0), (1, 2), (3, 2)])
def test_mg_pcg_jacobi_smoothers_and_zero_pre_match_oracle(ctx, smoother, pre):
    n, levels = 24, 3
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=3)
    b = ua.gen_uniform(103, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, levels)
    atol = 1e-10 * np.linalg.norm(b)
    ua.srand(1)
    cg = ua.ConjugateGradient(ctx, "multigrid", atol=atol, rtol=0, stol=0, max_it=800, smoother=smoother,
                              pre_smoothing_steps=pre, post_smoothing_steps=2)
    cg.set_transfer_operators(interps)
    x = cg.solve(A, b)
    ua.srand(1)
    mg = orc.MG(interps, smoother=smoother, pre=pre, post=2)
    mg.update(A)
    xo, so = mg.cg_solve(b, atol=atol, max_it=800)
    st = cg.status
    print(f"cg-mg smoother={smoother} pre={pre}: gpu it={st['iterates']} oracle it={so['iterates']} "
          f"rel.err={rel(x, xo):.2e}")
    assert abs(st["iterates"] - so["iterates"]) <= 1
    assert rel(x, xo) <= 1e-10
    if smoother == 0 and pre == 0:
        # Chebyshev3level takes one step for sweeps(0): one pre- against two post-steps make
        # the V-cycle non-symmetric, and CG stalls in the reference's arithmetic as well
        assert so["iterates"] == 800 and st["iterates"] == 800
    else:
        assert np.linalg.norm(b - A.to_scipy() @ x) <= 1.01 * atol


@pytest.mark.parametrize("smoother", [1, 3])
def test_mixed_precision_jacobi_smoothing(ctx, smoother):
    """FP32 V-cycle with Jacobi / PointJacobi smoothing in the FP64 CG: same iteration count
    as the FP64 preconditioner within +-1, solution within 1e-8 of the oracle."""
    n, levels = 24, 3
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=3)
    b = ua.gen_uniform(103, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, levels)
    atol = 1e-10 * np.linalg.norm(b)
    cg = ua.ConjugateGradient(ctx, "multigrid_mixed", atol=atol, rtol=0, stol=0, max_it=800, smoother=smoother,
                              pre_smoothing_steps=2, post_smoothing_steps=2)
    cg.set_transfer_operators(interps)
    x = cg.solve(A, b)
    mg = orc.MG(interps, smoother=smoother, pre=2, post=2)
    mg.update(A)
    xo, so = mg.cg_solve(b, atol=atol, max_it=800)
    assert abs(cg.status["iterates"] - so["iterates"]) <= 1
    assert np.linalg.norm(b - A.to_scipy() @ x) <= 1.01 * atol
    assert rel(x, xo) <= 1e-8


@pytest.mark.parametrize("smoother", [0, 1])
def test_mg_line_search_matches_oracle(ctx, smoother):
    A, b, s, mg = _mg_pair(ctx, 16, 3, smoother, 2, 2, use_line_search=True)
    x = s.apply(b)
    xo = mg.apply(b, max_it=3, atol=0, rtol=0)
    assert rel(x, xo) <= 1e-12
    # the line search changes the cycle
    _, _, s2, _ = _mg_pair(ctx, 16, 3, smoother, 2, 2)
    assert rel(s2.apply(b), xo) > 1e-8


def test_mg_full_cycle_matches_oracle(ctx):
    A, b, s, mg = _mg_pair(ctx, 16, 4, 0, 2, 2, cycle_type=3, v_cycle_repetition=2, max_it=5)
    x = s.apply(b)
    xo = mg.apply(b, max_it=5, atol=0, rtol=0)
    assert rel(x, xo) <= 1e-12
    assert s.status["iterates"] == 1  # FULL_CYCLE sets max_it(1)


def test_mg_full_cycle_reference_fixture(ctx, golden):
    """PetscLinearSolverTest::petsc_mg (tests/petsc/utopia_PetscLinearSolverTest.cpp:203-268):
    Multigrid with GaussSeidel on the Q1 fixture, then FULL_CYCLE with v_cycle_repetition 2."""
    d = golden / "utopia_data"
    A = petsc_io.read_mat(d / "f_A")
    rhs = petsc_io.read_vec(d / "f_rhs")
    Is = [petsc_io.read_mat(d / f) for f in ("I_2", "I_3")]
    s = ua.Multigrid(ctx, smoother=2, pre_smoothing_steps=3, post_smoothing_steps=3, atol=1e-9, rtol=1e-9)
    s.set_transfer_operators(Is)
    s.update(A)
    x = s.apply(rhs)
    assert np.linalg.norm(A.to_scipy() @ x - rhs) < 1e-6
    s2 = ua.Multigrid(ctx, smoother=2, pre_smoothing_steps=3, post_smoothing_steps=3, cycle_type=3,
                      v_cycle_repetition=2)
    s2.set_transfer_operators(Is)
    s2.update(A)
    xf = s2.apply(rhs)
    s3 = ua.Multigrid(ctx, smoother=2, pre_smoothing_steps=3, post_smoothing_steps=3, max_it=1)
    s3.set_transfer_operators(Is)
    s3.update(A)
    assert rel(xf, s3.apply(rhs)) <= 1e-14


def test_block_mg_fixture_matches_oracle(ctx, golden):
    """PetscLinearSolverTest::test_block_mg (tests/petsc/utopia_PetscLinearSolverTest.cpp:343-383):
    the reference's data/mg system (410 dofs, block size 2), one coarse level, symmetric
    Gauss-Seidel smoothing, pre/post 3, fix_semidefinite_operators, 20 cycles. GPU against the
    oracle with the GPU's colouring, and the reference's assertion approxeq(rhs, A x, 1e-6)."""
    d = golden / "utopia_data"
    A = petsc_io.read_mat(d / "mg_A.bin")
    I = petsc_io.read_mat(d / "mg_I.bin")
    rhs = petsc_io.read_vec(d / "mg_rhs.bin")
    s = ua.Multigrid(ctx, smoother=2, pre_smoothing_steps=3, post_smoothing_steps=3, fix_semidefinite_operators=1,
                     max_it=20, atol=1e-15, rtol=1e-15, stol=1e-15, block_size=2)
    s.set_transfer_operators([I])
    s.update(A)
    x = s.apply(rhs)
    assert np.abs(rhs - A.to_scipy() @ x).max() <= 1e-6
    mg = orc.MG([I], smoother=2, pre=3, post=3, masks=True, fix_semidefinite=True)
    mg.update(A)
    mg.set_colors(1, *s.colors(1))  # level 0, the coarsest, is the dense solve
    xo = mg.apply(rhs, max_it=20, atol=1e-15, rtol=1e-15)
    assert rel(x, xo) <= 1e-10
