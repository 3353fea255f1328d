"""Differentiable batched solve: X = A^-1 B as a torch.autograd.Function.

utopia's scripting layer exposes solvers to Python (scripting/utopia.i,
scripting/utopia_c_solve.h, scripting/examples/example_gradient_descent.py)
and its differentiable/ module differentiates expressions only; no reference
function differentiates through a linear solve. The backward pass here is the
adjoint method: for a loss L(X) with G = dL/dX,

    lambda = A^-T G        (adjoint solve; A is SPD, so the forward solver),
    dL/dB  = lambda,
    dL/dA_ij = -sum_c lambda[i, c] * X[j, c]   for (i, j) in pattern(A).

Values live on the device as torch tensors (float64, cuda); every solve and
the pattern gradient run through libutopia_hip (no CPU fallback).
"""
from __future__ import annotations

import torch

from . import Context
from .batched import BatchedSolver, BatchProblem


def _check(t: torch.Tensor, name: str):
    if not t.is_cuda or t.dtype != torch.float64:
        raise TypeError(f"{name}: float64 CUDA tensor expected, got {t.dtype} on {t.device}")


class _SolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, B, op):
        _check(values, "values")
        _check(B, "B")
        B = B.contiguous()
        X = torch.zeros_like(B)
        torch.cuda.synchronize(B.device)  # B, X and values are complete before the solver stream reads them
        op._ensure_setup(values)
        op.solver.solve_device(B, X)  # returns after its stream finished
        ctx.save_for_backward(values, X)
        ctx.op = op
        return X

    @staticmethod
    def backward(ctx, G):
        values, X = ctx.saved_tensors
        op = ctx.op
        G = G.contiguous()
        lam = torch.zeros_like(G)
        torch.cuda.synchronize(G.device)
        # the operator of this forward: another forward on other values may have set up since
        op._ensure_setup(values)
        op.solver.solve_device(G, lam)  # A^T lam = G, A = A^T
        grad_vals = torch.empty_like(values)
        op.solver.pattern_grad_device(lam, X, grad_vals)
        op.ctx.sync()
        return grad_vals, lam, None


class BatchedSolve:
    """Callable X = solve(values, B) over a fixed block-diagonal pattern.

    values: [nnz] float64 CUDA tensor, the finest operator's CSR values in the
    order of problem.A; B: [n, k] float64 CUDA tensor.
    """

    def __init__(self, problem: BatchProblem, k: int, mixed: bool = False, ctx: Context = None, **params):
        self.problem = problem
        self.ctx = ctx if ctx is not None else Context(torch.cuda.current_device())
        self.solver = BatchedSolver(self.ctx, k, mixed=mixed, **params)
        self._values = None  # the values the solver's hierarchy was built from (kept alive: read in place)
        self._version = None
        self.setups = 0

    def _ensure_setup(self, values: torch.Tensor):
        """Build the hierarchy (Galerkin levels, eigenvalue estimates, coarse inverses)
        from `values` unless it was built from these very values, unchanged since."""
        v = self._values
        if (v is not None and v.data_ptr() == values.data_ptr() and v.shape == values.shape
                and self._version == values._version):
            return
        self.solver.setup(self.problem, values=values)
        self._values, self._version = values, values._version
        self.setups += 1

    def __call__(self, values: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
        return _SolveFn.apply(values, B, self)
