"""Adjoint gradients of the batched solve through torch.autograd.

L = 1/2 |X - X*|^2 with X = A^-1 B (block-diagonal SPD A, k columns). The GPU
gradients (adjoint solve + pattern gradient) are checked against
  * central finite differences, h = 1e-6, of L computed with FP64 direct
    solves (scipy), for dL/dB and for stored entries A_ij, and
  * the transposed-solve gradient: lambda = A^-T (X - X*) by a direct solve,
    dL/dA_ij = -sum_c lambda[i, c] X[j, c].
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import utopia_amd as ua
from utopia_amd import batched as bt

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _loss(A_sp, B, Xs):
    X = spla.spsolve(A_sp.tocsc(), B)
    X = X.reshape(B.shape)
    return 0.5 * np.sum((X - Xs) ** 2)


@pytest.fixture(scope="module")
def setup():
    from utopia_amd.autograd import BatchedSolve
    p = bt.make_batch([12, 9, 15], seed=6, levels=2)
    k = 2
    B = ua.gen_uniform(31, p.n * k).reshape(p.n, k)
    Xs = ua.gen_normal(32, p.n * k).reshape(p.n, k)
    op = BatchedSolve(p, k, mixed=False, rtol=1e-13, max_it=500)
    return p, k, B, Xs, op


def _grads(p, k, B, Xs, op):
    dev = torch.device("cuda:0")
    vals = torch.tensor(p.A.val, dtype=torch.float64, device=dev, requires_grad=True)
    Bt = torch.tensor(B, dtype=torch.float64, device=dev, requires_grad=True)
    X = op(vals, Bt)
    L = 0.5 * ((X - torch.tensor(Xs, dtype=torch.float64, device=dev)) ** 2).sum()
    L.backward()
    return X.detach().cpu().numpy(), float(L.detach()), vals.grad.cpu().numpy(), Bt.grad.cpu().numpy()


def test_forward_matches_direct_solve(setup):
    p, k, B, Xs, op = setup
    X, L, _, _ = _grads(p, k, B, Xs, op)
    Xd = spla.spsolve(p.A.to_scipy().tocsc(), B).reshape(B.shape)
    assert np.linalg.norm(X - Xd) <= 1e-11 * np.linalg.norm(Xd)
    assert abs(L - _loss(p.A.to_scipy(), B, Xs)) <= 1e-10 * abs(L)


def test_gradient_matches_transposed_solve(setup):
    p, k, B, Xs, op = setup
    X, _, gA, gB = _grads(p, k, B, Xs, op)
    A = p.A.to_scipy()
    Xd = spla.spsolve(A.tocsc(), B).reshape(B.shape)
    lam = spla.spsolve(A.T.tocsc(), Xd - Xs).reshape(B.shape)
    rows = np.repeat(np.arange(p.n), np.diff(p.A.rowptr))
    ref = -(lam[rows] * Xd[p.A.col]).sum(axis=1)
    assert np.linalg.norm(gB - lam) <= 1e-10 * np.linalg.norm(lam)
    assert np.linalg.norm(gA - ref) <= 1e-10 * np.linalg.norm(ref)


def test_gradient_matches_central_differences(setup):
    p, k, B, Xs, op = setup
    _, _, gA, gB = _grads(p, k, B, Xs, op)
    h = 1e-6
    rng = np.random.default_rng(0)
    A0 = p.A.to_scipy().tocsr()
    for e in rng.choice(p.A.nnz, 12, replace=False):
        Ap, Am = A0.copy(), A0.copy()
        Ap.data[e] += h
        Am.data[e] -= h
        fd = (_loss(Ap, B, Xs) - _loss(Am, B, Xs)) / (2 * h)
        assert abs(fd - gA[e]) <= 1e-6 * max(1.0, abs(fd)), (e, fd, gA[e])
    for i in rng.choice(p.n * k, 8, replace=False):
        Bp, Bm = B.copy(), B.copy()
        Bp.flat[i] += h
        Bm.flat[i] -= h
        fd = (_loss(A0, Bp, Xs) - _loss(A0, Bm, Xs)) / (2 * h)
        assert abs(fd - gB.flat[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, gB.flat[i])


def test_rejects_non_cuda_tensors(setup):
    p, k, B, Xs, op = setup
    with pytest.raises(TypeError):
        op(torch.tensor(p.A.val), torch.tensor(B))
