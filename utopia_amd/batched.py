"""Batched MG-preconditioned CG over many independent SPD systems.

No reference analogue: utopia solves one system per solver object
(solvers/smoothers/utopia_ConjugateGradient_impl.hpp:195-294 with the
Multigrid preconditioner of solvers/multilevel/utopia_Multigrid.hpp:252-367).
Here ``nseg`` systems with ``k`` right-hand sides each are one block-diagonal
operator; each (system, rhs) pair follows that arithmetic on its own and stops
at its own tolerance. One CG iteration of the batch is a captured hipGraph.

Layout: right-hand sides and solutions are row-major ``[n_total, k]`` device
arrays (entry ``i*k + j`` is row ``i`` of column ``j``); system ``s`` owns
rows ``seg[s]:seg[s+1]``.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import (Context, Csr, DeviceArray, DeviceCsr, STENCIL_2D5_EDGECOEF, _alive, _ptr, check, gen_prolongation,
               gen_stencil, gen_uniform)
from ._abi import lib, uhip_csr


def _block_diag(mats: Sequence[Csr]) -> Csr:
    rows = sum(m.shape[0] for m in mats)
    cols = sum(m.shape[1] for m in mats)
    nnz = sum(m.nnz for m in mats)
    rp = np.zeros(rows + 1, np.int64)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    r0 = c0 = e0 = 0
    for m in mats:
        n, nz = m.shape[0], m.nnz
        rp[r0 + 1:r0 + n + 1] = m.rowptr[1:].astype(np.int64) + e0
        col[e0:e0 + nz] = m.col[:nz] + c0
        val[e0:e0 + nz] = m.val[:nz]
        r0, c0, e0 = r0 + n, c0 + m.shape[1], e0 + nz
    if nnz >= 2**31:
        raise ValueError("batched operator exceeds int32 nnz")
    return Csr(rp.astype(np.int32), col, val, (rows, cols))


def _offsets(sizes: Sequence[int]) -> np.ndarray:
    return np.concatenate([[0], np.cumsum(np.asarray(sizes, np.int64))]).astype(np.int32)


@dataclass
class BatchProblem:
    """Block-diagonal 2D systems with a geometric hierarchy per system.

    sides[s] is the interior grid side of system s; level l (0 = coarsest) has
    side sides[s] // 2**(levels-1-l) (integer halving, vertex-centred).
    """

    sides: List[int]
    levels: int
    systems: List[Csr]                      # per system, finest operator
    interps: List[List[Csr]]                # per system, coarse to fine
    A: Csr = None                           # block-diagonal finest operator
    seg: np.ndarray = None                  # nseg+1 row offsets (finest)
    batch_interps: List[Csr] = field(default_factory=list)   # coarse to fine
    coarse_segs: List[np.ndarray] = field(default_factory=list)  # coarsest first

    @property
    def nseg(self) -> int:
        return len(self.sides)

    @property
    def n(self) -> int:
        return int(self.seg[-1])


def batch_sides(seed: int, nsys: int = 64, lo: int = 64, hi: int = 512) -> List[int]:
    """Grid sides drawn log-uniform in [lo, hi] (rounded to integers)."""
    u = gen_uniform(seed, nsys, math.log(lo), math.log(hi))
    return [int(min(hi, max(lo, round(math.exp(v))))) for v in u]


def levels_for(max_side: int, coarse_side: int = 16) -> int:
    L = 1
    while max_side // 2 ** (L - 1) > coarse_side:
        L += 1
    return L


def make_batch(sides: Sequence[int], seed: int = 0, levels: Optional[int] = None, coarse_side: int = 16,
               kind: int = STENCIL_2D5_EDGECOEF) -> BatchProblem:
    """Systems s = 0..nseg-1: stencil `kind` on sides[s]^2 with seed `seed + s`
    (per-edge coefficients exp(N(0, 0.5)) for the default kind)."""
    sides = [int(s) for s in sides]
    L = levels if levels is not None else levels_for(max(sides), coarse_side)
    if min(sides) // 2 ** (L - 1) < 1:
        raise ValueError(f"{L} levels do not fit a side of {min(sides)}")
    systems = [gen_stencil(kind, n, n, 1, seed + s) for s, n in enumerate(sides)]
    interps = []
    for n in sides:
        dims = [n // 2 ** i for i in range(L)]  # finest first
        interps.append([gen_prolongation(dims[i], dims[i]) for i in reversed(range(L - 1))])
    p = BatchProblem(sides, L, systems, interps)
    p.A = _block_diag(systems)
    p.seg = _offsets([m.shape[0] for m in systems])
    for l in range(L - 1):  # interps[s][l]: level l -> l+1
        p.batch_interps.append(_block_diag([interps[s][l] for s in range(len(sides))]))
        p.coarse_segs.append(_offsets([interps[s][l].shape[1] for s in range(len(sides))]))
    return p


def _as_ptr(a) -> int:
    """Device pointer of a DeviceArray, a torch tensor or an int."""
    if isinstance(a, DeviceArray):
        return a.ptr
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return int(a)


class BatchedSolver:
    """uhip_bsolver: batched FP64 CG with an FP64 or FP32 (mixed) V-cycle."""

    def __init__(self, ctx: Context, k: int, mixed: bool = True, **params):
        self.ctx, self.k, self.mixed = ctx, int(k), bool(mixed)
        h = ctypes.c_void_p()
        check(lib().uhip_bsolver_create(ctx.h, self.k, 1 if mixed else 0, ctypes.byref(h)), "uhip_bsolver_create")
        self.h = h
        self.problem = None
        for key, v in params.items():
            self.set_param(key, v)

    def __del__(self):
        try:
            if _alive() and self.h:
                lib().uhip_bsolver_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_param(self, key: str, value):
        check(lib().uhip_bsolver_set_param(self.h, key.encode(), float(value)), f"set_param {key}")
        return self

    def setup(self, p: BatchProblem, values=None):
        """Upload the block-diagonal hierarchy and build the solver.

        values: optional device array (DeviceArray / torch tensor / pointer) of
        the finest operator's nnz values replacing p.A.val (not copied; it
        must stay alive while the solver is used).
        """
        self.problem = p
        self._A = DeviceCsr(self.ctx, p.A)
        m = uhip_csr.from_buffer_copy(self._A.m)
        if values is not None:
            m.val = _as_ptr(values)
        self._Am = m
        self._P = [DeviceCsr(self.ctx, I) for I in p.batch_interps]
        n = len(self._P)
        ptrs = (ctypes.c_void_p * max(n, 1))(*[ctypes.addressof(d.m) for d in self._P])
        self._cs = [np.ascontiguousarray(c, np.int32) for c in p.coarse_segs]
        csp = (ctypes.c_void_p * max(n, 1))(*[c.ctypes.data for c in self._cs])
        seg = np.ascontiguousarray(p.seg, np.int32)
        check(lib().uhip_bsolver_setup(self.h, p.nseg, _ptr(seg), ctypes.byref(m), n + 1, ptrs, csp),
              "uhip_bsolver_setup")
        return self

    def solve_device(self, B, X) -> int:
        """X (initial guess, overwritten) and B: device [n, k] FP64."""
        it = ctypes.c_int64()
        check(lib().uhip_bsolver_solve(self.h, ctypes.c_void_p(_as_ptr(B)), ctypes.c_void_p(_as_ptr(X)),
                                       ctypes.byref(it)), "uhip_bsolver_solve")
        return it.value

    def solve(self, B: np.ndarray, X: Optional[np.ndarray] = None) -> np.ndarray:
        B = np.ascontiguousarray(B, np.float64).reshape(-1)
        Bd = DeviceArray.from_numpy(self.ctx, B)
        Xd = DeviceArray.from_numpy(self.ctx, np.zeros_like(B) if X is None else np.asarray(X, np.float64))
        self.solve_device(Bd, Xd)
        return Xd.numpy().reshape(-1, self.k)

    def convergence(self):
        """(iterations per (system, rhs) [nseg, k], flags [nseg, k])."""
        m = self.problem.nseg * self.k
        it = np.zeros(m, np.int32)
        fl = np.zeros(m, np.int32)
        check(lib().uhip_bsolver_get_conv(self.h, _ptr(it), _ptr(fl)), "uhip_bsolver_get_conv")
        return it.reshape(-1, self.k), fl.reshape(-1, self.k)

    def pattern_grad_device(self, lam, x, grad):
        check(lib().uhip_bsolver_pattern_grad(self.h, ctypes.c_void_p(_as_ptr(lam)), ctypes.c_void_p(_as_ptr(x)),
                                              ctypes.c_void_p(_as_ptr(grad))), "uhip_bsolver_pattern_grad")

    def info(self, key: str) -> float:
        v = ctypes.c_double()
        check(lib().uhip_bsolver_get_info(self.h, key.encode(), ctypes.byref(v)), f"get_info {key}")
        return v.value

    def eig_max(self, l: int) -> np.ndarray:
        return np.array([self.info(f"eig_max_{l}_{s}") for s in range(self.problem.nseg)])

    def set_level_eig(self, l: int, s: int, eig_max: float):
        check(lib().uhip_bsolver_set_level_eig(self.h, l, s, eig_max), "set_level_eig")
