"""utopia_amd - MI355X (gfx950) backend for utopia's sparse-solver hot path.

Python face of libutopia_hip.so (C ABI in include/utopia_hip.h):

* generators:        ``gen_stencil``, ``gen_prolongation``, ``gen_uniform``, ``gen_normal``
* kernel level:      ``Context``, ``DeviceArray``, ``DeviceCsr`` (SpMV family, BLAS1, RAP, ...)
* solver level:      ``ConjugateGradient``, ``Multigrid``, ``MPRGP``, ``ProjectedGaussSeidel``
                     with utopia's parameter names (``atol``, ``max_it``,
                     ``pre_smoothing_steps`` ...) and update()/apply()/solve().

Every numeric call runs on the GPU through the library; nothing here falls
back to a CPU implementation.
"""
from __future__ import annotations

import atexit
import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from ._abi import LibraryMissing, lib, uhip_cheb_step, uhip_csr  # noqa: F401

# Device objects still alive when the interpreter exits are left to the process
# teardown: no HIP call is made from a finaliser after exit has begun.
_EXITING = [False]
atexit.register(lambda: _EXITING.__setitem__(0, True))


def _alive() -> bool:
    return not _EXITING[0]

F64, F32 = 0, 1
FMT_AUTO, FMT_ROWBLOCK, FMT_SELL64, FMT_DIA, FMT_VI = -1, 0, 1, 2, 3
_NP = {F64: np.float64, F32: np.float32}
_DT = {np.dtype(np.float64): F64, np.dtype(np.float32): F32}

(SOLVER_CG, SOLVER_CG_JACOBI, SOLVER_CG_MG, SOLVER_CG_MG_MIXED, SOLVER_MG, SOLVER_MPRGP, SOLVER_PGS,
 SOLVER_PCG, SOLVER_PG, SOLVER_MMG, SOLVER_AMG, SOLVER_CG_AMG) = range(12)

# ConvergenceReason (solvers/utopia_ConvergenceReason.hpp:84-100)
CONVERGED_FNORM_ABS, CONVERGED_FNORM_RELATIVE, CONVERGED_SNORM_RELATIVE, DIVERGED_MAX_IT = 2, 3, 4, -5


class HipError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = lib().uhip_last_error()
        super().__init__(f"{where} failed ({code}): {msg.decode() if msg else ''}")
        self.code = code


def check(status: int, where: str = "uhip call"):
    if status != 0:
        raise HipError(status, where)


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


# ------------------------------------------------------------------ host CSR

@dataclass
class Csr:
    """Host CSR matrix (int32 indices, float64 values)."""

    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    shape: tuple

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=self.shape)

    @staticmethod
    def from_scipy(m) -> "Csr":
        m = m.tocsr()
        m.sort_indices()
        return Csr(m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data.astype(np.float64), m.shape)


STENCIL_2D5, STENCIL_3D7, STENCIL_3D7_VARCOEF, STENCIL_2D9, STENCIL_2D5_EDGECOEF, STENCIL_3D7_BLOCKPERM = range(6)


def gen_stencil(kind: int, nx: int, ny: int, nz: int = 1, seed: int = 0, shift: float = 0.0) -> Csr:
    """Seeded CSR stencil operator (see uhip_gen_stencil for the kinds)."""
    L = lib()
    nnz = L.uhip_gen_nnz(kind, nx, ny, nz)
    if nnz < 0:
        raise ValueError(f"bad stencil kind {kind}")
    n = nx * ny * nz
    rp = np.zeros(n + 1, np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float64)
    check(L.uhip_gen_stencil(kind, nx, ny, nz, ctypes.c_uint64(seed), shift, _ptr(rp), _ptr(col), _ptr(val)),
          "uhip_gen_stencil")
    assert rp[-1] == nnz
    return Csr(rp, col[:nnz], val[:nnz], (n, n))


def gen_prolongation(nx: int, ny: int, nz: int = 1) -> Csr:
    """Trilinear interpolation from the (nx/2, ny/2, nz/2) grid to (nx, ny, nz)."""
    L = lib()
    nnz = L.uhip_gen_prolongation_nnz(nx, ny, nz)
    n = nx * ny * nz
    cx, cy, cz = (max(1, d // 2) if d > 1 else 1 for d in (nx, ny, nz))
    rp = np.zeros(n + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    check(L.uhip_gen_prolongation(nx, ny, nz, _ptr(rp), _ptr(col), _ptr(val)), "uhip_gen_prolongation")
    return Csr(rp, col, val, (n, cx * cy * cz))


def gen_uniform(seed: int, n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    x = np.zeros(n, np.float64)
    check(lib().uhip_gen_uniform(ctypes.c_uint64(seed), n, lo, hi, _ptr(x)), "uhip_gen_uniform")
    return x


def gen_normal(seed: int, n: int, mu: float = 0.0, sigma: float = 1.0) -> np.ndarray:
    x = np.zeros(n, np.float64)
    check(lib().uhip_gen_normal(ctypes.c_uint64(seed), n, mu, sigma, _ptr(x)), "uhip_gen_normal")
    return x


def mg_hierarchy(nx: int, ny: int, nz: int, levels: int):
    """Interpolation operators coarse->fine for `levels` levels of factor-2 coarsening."""
    dims = [(nx, ny, nz)]
    for _ in range(levels - 1):
        x, y, z = dims[-1]
        dims.append(tuple(d // 2 if d > 1 else 1 for d in (x, y, z)))
    ops = [gen_prolongation(*dims[l]) for l in range(levels - 1)]
    return list(reversed(ops))


# ------------------------------------------------------------------ device

class Context:
    """uhip_ctx: one device, one stream, reduction workspace."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().uhip_ctx_create(device, None, ctypes.byref(h)), "uhip_ctx_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib().uhip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            if _alive():
                self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().uhip_sync(self.h), "uhip_sync")

    @property
    def scalars(self) -> int:
        return lib().uhip_ctx_scalars(self.h)

    def set_autotune(self, reps: int):
        """reps > 0: automatic SpMV plans made through this context are chosen by timing."""
        check(lib().uhip_ctx_set_autotune(self.h, int(reps)), "uhip_ctx_set_autotune")

    def read_scalars(self, addr: int, n: int) -> np.ndarray:
        out = np.zeros(n, np.float64)
        check(lib().uhip_read_scalars(self.h, ctypes.c_void_p(addr), n, _ptr(out)), "uhip_read_scalars")
        return out


class DeviceArray:
    """Device buffer owned by the library allocator."""

    def __init__(self, ctx: Context, n: int, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        p = ctypes.c_void_p()
        check(lib().uhip_malloc(ctx.h, ctypes.byref(p), max(1, self.n) * self.dtype.itemsize), "uhip_malloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, ctx: Context, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        if a.size:
            check(lib().uhip_memcpy_htod(ctx.h, ctypes.c_void_p(d.ptr), _ptr(a), a.nbytes), "htod")
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.n:
            check(lib().uhip_memcpy_dtoh(self.ctx.h, _ptr(out), ctypes.c_void_p(self.ptr), out.nbytes), "dtoh")
        return out

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib().uhip_memcpy_htod(self.ctx.h, ctypes.c_void_p(self.ptr), _ptr(a), a.nbytes), "htod")

    @property
    def code(self) -> int:
        return _DT[self.dtype]

    def free(self):
        if self.ptr:
            lib().uhip_free(self.ctx.h, ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if _alive() and self.ctx.h:
                self.free()
        except Exception:
            pass


class DeviceCsr:
    """uhip_csr on the device with its row-block plan."""

    def __init__(self, ctx: Context, A: Csr, dtype=np.float64, block_nnz: int = 0, fmt: Optional[int] = None):
        """fmt: None (block_nnz > 0: row-block CSR, 0: automatic), or FMT_AUTO / FMT_ROWBLOCK / FMT_SELL64 / FMT_DIA / FMT_VI."""
        self.ctx = ctx
        self.shape = A.shape
        self._rp = DeviceArray.from_numpy(ctx, A.rowptr.astype(np.int32))
        self._col = DeviceArray.from_numpy(ctx, A.col.astype(np.int32))
        self._val = DeviceArray.from_numpy(ctx, A.val.astype(dtype))
        m = uhip_csr()
        m.nrows, m.ncols, m.nnz = A.shape[0], A.shape[1], A.nnz
        m.rowptr, m.col, m.val = self._rp.ptr, self._col.ptr, self._val.ptr
        m.dtype = _DT[np.dtype(dtype)]
        self.m = m
        if fmt is None:
            check(lib().uhip_csr_plan(ctx.h, ctypes.byref(m), block_nnz), "uhip_csr_plan")
        else:
            check(lib().uhip_csr_plan_format(ctx.h, ctypes.byref(m), fmt, block_nnz), "uhip_csr_plan_format")

    @property
    def format(self) -> int:
        return self.m.format

    @property
    def ref(self):
        return ctypes.byref(self.m)

    @property
    def dtype(self):
        return _NP[self.m.dtype]

    def __del__(self):
        try:
            if _alive() and self.ctx.h:
                lib().uhip_csr_plan_destroy(self.ctx.h, ctypes.byref(self.m))
        except Exception:
            pass

    def blocks(self) -> np.ndarray:
        out = np.zeros(self.m.nblocks + 1, np.int32)
        check(lib().uhip_memcpy_dtoh(self.ctx.h, _ptr(out), ctypes.c_void_p(self.m.blocks), out.nbytes), "dtoh")
        return out

    # ---- SpMV family -------------------------------------------------
    def spmv(self, x: DeviceArray, y: DeviceArray):
        check(lib().uhip_spmv(self.ctx.h, self.ref, ctypes.c_void_p(x.ptr), ctypes.c_void_p(y.ptr)), "uhip_spmv")

    def residual(self, x: DeviceArray, b: DeviceArray, r: DeviceArray):
        check(lib().uhip_spmv_residual(self.ctx.h, self.ref, ctypes.c_void_p(x.ptr), ctypes.c_void_p(b.ptr), b.code,
                                       ctypes.c_void_p(r.ptr)), "uhip_spmv_residual")


def csr_from_raw(ctx: Context, m: uhip_csr) -> Csr:
    """Copy a uhip_csr (FP64) produced by the library back to the host."""
    rp = np.zeros(m.nrows + 1, np.int32)
    col = np.zeros(max(m.nnz, 1), np.int32)
    val = np.zeros(max(m.nnz, 1), np.float64)
    L = lib()
    check(L.uhip_memcpy_dtoh(ctx.h, _ptr(rp), ctypes.c_void_p(m.rowptr), rp.nbytes), "dtoh")
    if m.nnz:
        check(L.uhip_memcpy_dtoh(ctx.h, _ptr(col), ctypes.c_void_p(m.col), 4 * m.nnz), "dtoh")
        check(L.uhip_memcpy_dtoh(ctx.h, _ptr(val), ctypes.c_void_p(m.val), 8 * m.nnz), "dtoh")
    return Csr(rp, col[:m.nnz], val[:m.nnz], (m.nrows, m.ncols))


# ------------------------------------------------------------------ solvers

class Solver:
    """A uhip_solver: one of utopia's solvers on the HIP backend."""

    TYPE = None

    def __init__(self, ctx: Context, type_: Optional[int] = None, **params):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().uhip_solver_create(ctx.h, self.TYPE if type_ is None else type_, ctypes.byref(h)),
              "uhip_solver_create")
        self.h = h
        self.n = None
        for k, v in params.items():
            self.set_param(k, v)

    @classmethod
    def by_name(cls, ctx: Context, type_name: str, backend: str = "any", **params) -> "Solver":
        """A solver by its registry name, as QPSolverRegistry::find(backend, type) and
        OmniLinearSolver's "type" select one ("cg", "mprgp", "pgs", "pcg", "pg", "mmg", ...)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().uhip_solver_create_by_name(ctx.h, backend.encode(), type_name.encode(), ctypes.byref(h)),
              f"uhip_solver_create_by_name {backend}/{type_name}")
        self.h = h
        self.n = None
        for k, v in params.items():
            self.set_param(k, v)
        return self

    def __del__(self):
        try:
            if _alive() and getattr(self, "h", None):
                lib().uhip_solver_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_param(self, key: str, value):
        check(lib().uhip_solver_set_param(self.h, key.encode(), float(value)), f"set_param {key}")
        return self

    # utopia-style setters
    def atol(self, v): return self.set_param("atol", v)
    def rtol(self, v): return self.set_param("rtol", v)
    def stol(self, v): return self.set_param("stol", v)
    def max_it(self, v): return self.set_param("max_it", v)
    def verbose(self, v): return self.set_param("verbose", 1 if v else 0)

    def set_transfer_operators(self, ops: Sequence[Csr]):
        """Interpolation operators, coarse to fine (Multigrid::set_transfer_operators)."""
        self._ops = [(np.ascontiguousarray(o.rowptr, np.int32), np.ascontiguousarray(o.col, np.int32),
                      np.ascontiguousarray(o.val, np.float64), o.shape) for o in ops]
        n = len(ops)
        nr = (ctypes.c_int32 * n)(*[o.shape[0] for o in ops])
        nc = (ctypes.c_int32 * n)(*[o.shape[1] for o in ops])
        rps = (ctypes.c_void_p * n)(*[o[0].ctypes.data for o in self._ops])
        cols = (ctypes.c_void_p * n)(*[o[1].ctypes.data for o in self._ops])
        vals = (ctypes.c_void_p * n)(*[o[2].ctypes.data for o in self._ops])
        check(lib().uhip_solver_set_transfer_operators(self.h, n, nr, nc, rps, cols, vals),
              "set_transfer_operators")
        return self

    def update(self, A: Csr):
        self.n = A.shape[0]
        rp = np.ascontiguousarray(A.rowptr, np.int32)
        col = np.ascontiguousarray(A.col, np.int32)
        val = np.ascontiguousarray(A.val, np.float64)
        check(lib().uhip_solver_update_crs(self.h, self.n, _ptr(rp), _ptr(col), _ptr(val)), "update_crs")
        return self

    def set_box_constraints(self, lb: Optional[np.ndarray] = None, ub: Optional[np.ndarray] = None):
        lb_ = None if lb is None else np.ascontiguousarray(lb, np.float64)
        ub_ = None if ub is None else np.ascontiguousarray(ub, np.float64)
        check(lib().uhip_solver_set_box_constraints(self.h, None if lb_ is None else _ptr(lb_),
                                                    None if ub_ is None else _ptr(ub_)), "set_box_constraints")
        self._bounds = (lb_, ub_)
        return self

    def apply(self, b: np.ndarray, x: Optional[np.ndarray] = None) -> np.ndarray:
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(x, np.float64).copy()
        check(lib().uhip_solver_apply(self.h, _ptr(b), _ptr(x)), "apply")
        return x

    def apply_device(self, b: DeviceArray, x: DeviceArray):
        check(lib().uhip_solver_apply_device(self.h, ctypes.c_void_p(b.ptr), ctypes.c_void_p(x.ptr)), "apply_device")

    def precondition(self, r: np.ndarray) -> np.ndarray:
        """z = M r with this CG solver's preconditioner."""
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        check(lib().uhip_solver_precondition(self.h, _ptr(r), _ptr(z)), "precondition")
        return z

    def solve(self, A: Csr, b: np.ndarray, x: Optional[np.ndarray] = None) -> np.ndarray:
        self.update(A)
        return self.apply(b, x)

    @property
    def status(self) -> dict:
        it = ctypes.c_int64()
        g = ctypes.c_double()
        r = ctypes.c_int()
        s = ctypes.c_double()
        check(lib().uhip_solver_get_status(self.h, ctypes.byref(it), ctypes.byref(g), ctypes.byref(r),
                                           ctypes.byref(s)), "get_status")
        return {"iterates": it.value, "gradient_norm": g.value, "reason": r.value, "seconds": s.value}

    def info(self, key: str) -> float:
        v = ctypes.c_double()
        check(lib().uhip_solver_get_info(self.h, key.encode(), ctypes.byref(v)), f"get_info {key}")
        return v.value

    # multigrid introspection
    def level(self, l: int) -> Csr:
        n = int(self.info(f"level_rows_{l}"))
        nnz = int(self.info(f"level_nnz_{l}"))
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1), np.float64)
        check(lib().uhip_solver_get_level(self.h, l, _ptr(rp), _ptr(col), _ptr(val)), "get_level")
        return Csr(rp, col[:nnz], val[:nnz], (n, n))

    def transfer(self, l: int) -> Csr:
        """Interpolation l (coarse to fine) of a multigrid solver."""
        n = int(self.info(f"transfer_rows_{l}"))
        m = int(self.info(f"transfer_cols_{l}"))
        nnz = int(self.info(f"transfer_nnz_{l}"))
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1), np.float64)
        check(lib().uhip_solver_get_transfer(self.h, l, _ptr(rp), _ptr(col), _ptr(val)), "get_transfer")
        return Csr(rp, col[:nnz], val[:nnz], (n, m))

    def mask(self, l: int) -> np.ndarray:
        n = int(self.info(f"level_rows_{l}"))
        out = np.zeros(n, np.float64)
        check(lib().uhip_solver_get_mask(self.h, l, _ptr(out)), "get_mask")
        return out

    def set_level_eigs(self, l: int, eig_max: float, eig_min: float):
        check(lib().uhip_solver_set_level_eigs(self.h, l, eig_max, eig_min), "set_level_eigs")

    def colors(self, l: int):
        """(colour of every row, number of colours) of level l (GaussSeidel smoother)."""
        n = int(self.info(f"level_rows_{l}"))
        out = np.zeros(max(n, 1), np.int32)
        nc = ctypes.c_int32()
        check(lib().uhip_solver_get_colors(self.h, l, _ptr(out), ctypes.byref(nc)), "get_colors")
        return out[:n], nc.value


class ConjugateGradient(Solver):
    """ConjugateGradient<HipCrsMatrix, HipVector, HOMEMADE>.

    preconditioner: None, "jacobi" (InvDiagPreconditioner), "multigrid" (FP64
    V-cycle), "multigrid_mixed" (FP32 V-cycle inside the FP64 CG) or "amg"
    (AlgebraicMultigrid, one cycle).
    """

    _TYPES = {None: SOLVER_CG, "jacobi": SOLVER_CG_JACOBI, "multigrid": SOLVER_CG_MG,
              "multigrid_mixed": SOLVER_CG_MG_MIXED, "amg": SOLVER_CG_AMG}

    def __init__(self, ctx: Context, preconditioner: Optional[str] = None, **params):
        super().__init__(ctx, self._TYPES[preconditioner], **params)


class Multigrid(Solver):
    TYPE = SOLVER_MG


class MPRGP(Solver):
    TYPE = SOLVER_MPRGP


class ProjectedGaussSeidel(Solver):
    TYPE = SOLVER_PGS


class ProjectedConjugateGradient(Solver):
    TYPE = SOLVER_PCG


class ProjectedGradient(Solver):
    TYPE = SOLVER_PG


class AlgebraicMultigrid(Solver):
    """AlgebraicMultigrid: Agglomerate transfers built from the operator at update
    (params n_levels, bmax, weight; block_size 2-4 aggregates B x B blocks with
    BlockAgglomerate, component picks the block weight entry; and the Multigrid
    smoother/cycle parameters)."""
    TYPE = SOLVER_AMG


class MonotoneMultigrid(Solver):
    """MonotoneMultigrid: set_transfer_operators (coarse to fine; the finest is
    truncated on the active set) before update."""
    TYPE = SOLVER_MMG


def srand(seed: int = 1):
    """Seed the C library rand() stream drawn by PowerMethod (default seed is 1)."""
    ctypes.CDLL(None).srand(ctypes.c_uint(seed))


def device_count() -> int:
    c = ctypes.c_int()
    check(lib().uhip_device_count(ctypes.byref(c)), "uhip_device_count")
    return c.value
