"""ctypes binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py as the checker. Nothing in utopia_amd imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


class or_csr(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int32), ("ncols", ctypes.c_int32), ("nnz", ctypes.c_int64),
                ("rowptr", ctypes.c_void_p), ("col", ctypes.c_void_p), ("val", ctypes.c_void_p)]


class or_status(ctypes.Structure):
    _fields_ = [("iterates", ctypes.c_int64), ("gradient_norm", ctypes.c_double), ("reason", ctypes.c_int),
                ("converged", ctypes.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_L = None


def lib():
    global _L
    if _L is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P, D, I, I64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int64
        sig = {
            "or_spmv": (None, [P, P, P]),
            "or_spmv_f32": (None, [ctypes.c_int32, P, P, P, P, P]),
            "or_dot": (D, [I64, P, P]),
            "or_norm2": (D, [I64, P]),
            "or_transpose": (None, [P, P]),
            "or_spgemm": (None, [P, P, P]),
            "or_ptap": (None, [P, P, P]),
            "or_csr_free": (None, [P]),
            "or_cholesky": (I, [ctypes.c_int32, P]),
            "or_cholesky_solve": (None, [ctypes.c_int32, P, P, P]),
            "or_cg_solve": (I, [P, P, P, I, D, D, D, I64, P]),
            "or_mg_create": (P, [I, P, I, I, I, I, I, I]),
            "or_mg_update": (I, [P, P]),
            "or_mg_set_cycle": (None, [P, I, I64, I]),
            "or_mg_apply": (I, [P, P, P, I64, D, D]),
            "or_cg_mg_solve": (I, [P, P, P, D, I64, P]),
            "or_mg_nlevels": (I, [P]),
            "or_mg_level": (P, [P, I]),
            "or_mg_mask": (P, [P, I]),
            "or_mg_eigs": (None, [P, I, P, P]),
            "or_mg_set_eigs": (None, [P, I, D, D]),
            "or_mg_set_colors": (None, [P, I, P, ctypes.c_int32]),
            "or_gs_smooth": (None, [P, P, P, ctypes.c_int32, P, P, I, I]),
            "or_jp_color": (None, [P, ctypes.c_uint64, P, P]),
            "or_mg_destroy": (None, [P]),
            "or_power_method": (D, [P, D, I64]),
            "or_jacobi_smooth": (None, [P, P, P, I]),
            "or_point_jacobi_smooth": (None, [P, P, P, I]),
            "or_chebyshev_smooth": (None, [P, P, P, I, D, D]),
            "or_mprgp_solve": (I, [P, P, P, P, P, D, D, I64, D, I64, P]),
            "or_pgs_solve": (I, [P, P, P, P, P, D, I64, I, I64, P]),
            "or_pcg_solve": (I, [P, P, P, P, P, D, D, D, I64, P]),
            "or_agglomerate": (ctypes.c_int32, [P, I64, D, P]),
            "or_mmg_solve": (I, [P, P, P, P, P, I, P, I, I, I64, D, D, D, D, I64, P]),
            "or_pg_solve": (I, [P, P, P, P, P, D, D, D, I64, P]),
            "or_set_threads": (None, [I]),
            "or_get_threads": (I, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _L = L
    return _L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class Mat:
    """Keeps numpy arrays alive behind an or_csr."""

    def __init__(self, rowptr, col, val, shape):
        self.rowptr = np.ascontiguousarray(rowptr, np.int32)
        self.col = np.ascontiguousarray(col, np.int32)
        self.val = np.ascontiguousarray(val, np.float64)
        self.shape = tuple(shape)
        self.s = or_csr(shape[0], shape[1], int(self.rowptr[-1]), self.rowptr.ctypes.data, self.col.ctypes.data,
                        self.val.ctypes.data)

    @classmethod
    def of(cls, A):
        return cls(A.rowptr, A.col, A.val, A.shape)

    @property
    def ref(self):
        return ctypes.byref(self.s)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=self.shape)


def _from_owned(s: or_csr):
    """Copy an oracle-allocated or_csr to numpy and free it."""
    n, nnz = s.nrows, s.nnz
    rp = np.ctypeslib.as_array(ctypes.cast(s.rowptr, ctypes.POINTER(ctypes.c_int32)), (n + 1,)).copy()
    col = np.ctypeslib.as_array(ctypes.cast(s.col, ctypes.POINTER(ctypes.c_int32)), (max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(ctypes.cast(s.val, ctypes.POINTER(ctypes.c_double)), (max(nnz, 1),))[:nnz].copy()
    shape = (s.nrows, s.ncols)
    lib().or_csr_free(ctypes.byref(s))
    return rp, col, val, shape


def spmv(A, x):
    M = Mat.of(A)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(A.shape[0], np.float64)
    lib().or_spmv(M.ref, _p(x), _p(y))
    return y


def spmv_f32(A, x):
    rp = np.ascontiguousarray(A.rowptr, np.int32)
    col = np.ascontiguousarray(A.col, np.int32)
    val = np.ascontiguousarray(A.val, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(A.shape[0], np.float32)
    lib().or_spmv_f32(A.shape[0], _p(rp), _p(col), _p(val), _p(x), _p(y))
    return y


def dot(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().or_dot(x.size, _p(x), _p(y))


def transpose(A):
    out = or_csr()
    M = Mat.of(A)
    lib().or_transpose(M.ref, ctypes.byref(out))
    return _from_owned(out)


def spgemm(A, B):
    out = or_csr()
    MA, MB = Mat.of(A), Mat.of(B)
    if MA.shape[1] != MB.shape[0]:
        raise ValueError(f"spgemm: shapes {MA.shape} and {MB.shape} do not chain")
    lib().or_spgemm(MA.ref, MB.ref, ctypes.byref(out))
    return _from_owned(out)


def ptap(A, P):
    out = or_csr()
    MA, MP = Mat.of(A), Mat.of(P)
    if MA.shape[0] != MA.shape[1] or MA.shape[1] != MP.shape[0]:
        raise ValueError(f"ptap: A {MA.shape} is not square or does not match P {MP.shape}")
    lib().or_ptap(MA.ref, MP.ref, ctypes.byref(out))
    return _from_owned(out)


def agglomerate(A, bmax=3, weight=1.0 / 3):
    """Agglomerate transfer (interpolation) of A as (rowptr, col, val, shape)."""
    out = or_csr()
    lib().or_agglomerate(Mat.of(A).ref, bmax, weight, ctypes.byref(out))
    return _from_owned(out)


def agglomerate_blocks(A, B, bmax=3, weight=1.0 / 3, component=None):
    """BlockAgglomerate::create_transfer (solvers/multilevel/agglomerate/
    utopia_BlockAgglomerate.hpp:37-141) for block size B, as (rowptr, col, val, shape).
    Block (I, J) of the B x B block graph exists where any entry of its B x B tile is
    stored; its weight is the tile's (component, component) entry (0 where absent;
    component defaults to B - 1, :160). Pure-Python loops: small inputs only."""
    M = Mat.of(A)
    rp, col, val = np.asarray(M.rowptr), np.asarray(M.col), np.asarray(M.val)
    n = M.shape[0]
    c = B - 1 if component is None else component
    assert n % B == 0 and 0 <= c < B
    nb = n // B
    rows = []  # per block row: [(J, weight)] in ascending J (the block CRS row view)
    for I in range(nb):
        w = {}
        for r in range(I * B, (I + 1) * B):
            for e in range(rp[r], rp[r + 1]):
                w.setdefault(int(col[e]) // B, 0.0)
        r = I * B + c
        for e in range(rp[r], rp[r + 1]):
            if int(col[e]) % B == c:
                w[int(col[e]) // B] = float(val[e])
        rows.append(sorted(w.items()))
    a_max = [max([abs(v) for J, v in rows[I] if J != I], default=0.0) for I in range(nb)]  # :55-63
    parent = [-1] * nb
    nc = 0
    for I in range(nb):  # :66-118
        if parent[I] != -1:
            continue
        parent[I] = nc
        count = 0
        for J, v in rows[I]:
            if weight * a_max[I] < abs(v) and parent[J] == -1:
                parent[J] = nc
                count += 1
            if count >= bmax:
                break
        if count == 0:  # not a cluster centre: join the aggregate of the strongest block
            parent[I] = -1
            best, arg = 0.0, -1
            for J, v in rows[I]:
                if abs(v) > best:
                    best, arg = abs(v), J
            if arg != -1:
                parent[I] = parent[arg]
        else:
            nc += 1
    for I in range(nb):  # :112-118
        if parent[I] == -1:
            parent[I] = nc
            nc += 1
    pcol = np.array([parent[i // B] * B + i % B for i in range(n)], np.int32)  # :127-137
    return np.arange(n + 1, dtype=np.int32), pcol, np.ones(n), (n, nc * B)


def amg_transfers(A, n_levels, bmax=3, weight=1.0 / 3, block_size=1):
    """AlgebraicMultigridBuilder::build (solvers/multilevel/utopia_AlgebraicMultigrid.hpp:23-52):
    P_l = agglomerate(A_{l+1}), A_l = P_l^T A_{l+1} P_l from the finest level down;
    block_size 2-4 aggregates with BlockAgglomerate (:87-103).
    Returns the interpolations and the operators, both coarse to fine, as Mat."""
    Ps, As = [], [Mat.of(A)]
    for _ in range(n_levels - 1):
        if 2 <= block_size <= 4:
            P = Mat(*agglomerate_blocks(As[0], block_size, bmax, weight))
        else:
            P = Mat(*agglomerate(As[0], bmax, weight))
        As.insert(0, Mat(*ptap(As[0], P)))
        Ps.insert(0, P)
    return Ps, As


def cholesky(Adense):
    L = np.ascontiguousarray(Adense, np.float64).copy()
    info = lib().or_cholesky(L.shape[0], _p(L))
    return L, info


def cholesky_solve(L, b):
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b)
    lib().or_cholesky_solve(L.shape[0], _p(L), _p(b), _p(x))
    return x


def status_dict(st):
    return {"iterates": st.iterates, "gradient_norm": st.gradient_norm, "reason": st.reason,
            "converged": st.converged}


def cg_solve(A, b, x0=None, precond=True, atol=1e-9, rtol=1e-9, stol=1e-11, max_it=1000):
    M = Mat.of(A)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    st = or_status()
    lib().or_cg_solve(M.ref, _p(b), _p(x), 1 if precond else 0, atol, rtol, stol, max_it, ctypes.byref(st))
    return x, status_dict(st)


def power_method(A, tol=1e-2, max_it=30):
    M = Mat.of(A)
    return lib().or_power_method(M.ref, tol, max_it)


def chebyshev_smooth(A, b, x, nu, eig_max, eig_min):
    M = Mat.of(A)
    b = np.ascontiguousarray(b, np.float64)
    x = np.ascontiguousarray(x, np.float64).copy()
    lib().or_chebyshev_smooth(M.ref, _p(b), _p(x), nu, eig_max, eig_min)
    return x


def jacobi_smooth(A, b, x, sweeps, point=False):
    """Jacobi (point=False) or PointJacobi (point=True) sweeps from x."""
    M = Mat.of(A)
    b = np.ascontiguousarray(b, np.float64)
    x = np.ascontiguousarray(x, np.float64).copy()
    f = lib().or_point_jacobi_smooth if point else lib().or_jacobi_smooth
    f(M.ref, _p(b), _p(x), int(sweeps))
    return x


class MG:
    """Oracle Multigrid (+ CG-MG)."""

    SMOOTHER_CHEBYSHEV, SMOOTHER_JACOBI, SMOOTHER_GAUSS_SEIDEL, SMOOTHER_POINT_JACOBI = 0, 1, 2, 3
    MULTIPLICATIVE_CYCLE, FULL_CYCLE = 1, 3

    def __init__(self, interps, smoother=0, pre=3, post=3, mg_type=1, masks=True, fix_semidefinite=False,
                 cycle_type=1, v_cycle_repetition=10, use_line_search=False):
        self._I = [Mat.of(I) for I in interps]
        arr = (ctypes.c_void_p * len(self._I))(*[ctypes.addressof(m.s) for m in self._I])
        self.h = lib().or_mg_create(len(interps) + 1, arr, smoother, pre, post, mg_type, 1 if masks else 0,
                                    1 if fix_semidefinite else 0)
        if not self.h:
            raise ValueError("or_mg_create failed")
        lib().or_mg_set_cycle(self.h, int(cycle_type), int(v_cycle_repetition), 1 if use_line_search else 0)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mg_destroy(self.h)
            self.h = None

    def update(self, A):
        self._A = Mat.of(A)
        return lib().or_mg_update(self.h, self._A.ref)

    def apply(self, b, x=None, max_it=1, atol=1e-9, rtol=1e-9):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(x, np.float64).copy()
        lib().or_mg_apply(self.h, _p(b), _p(x), max_it, atol, rtol)
        return x

    def cg_solve(self, b, x0=None, atol=1e-9, max_it=1000):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        st = or_status()
        lib().or_cg_mg_solve(self.h, _p(b), _p(x), atol, max_it, ctypes.byref(st))
        return x, status_dict(st)

    def nlevels(self):
        return lib().or_mg_nlevels(self.h)

    def level(self, l):
        p = lib().or_mg_level(self.h, l)
        s = or_csr.from_address(p)
        n, nnz = s.nrows, s.nnz
        rp = np.ctypeslib.as_array(ctypes.cast(s.rowptr, ctypes.POINTER(ctypes.c_int32)), (n + 1,)).copy()
        col = np.ctypeslib.as_array(ctypes.cast(s.col, ctypes.POINTER(ctypes.c_int32)), (max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(ctypes.cast(s.val, ctypes.POINTER(ctypes.c_double)), (max(nnz, 1),))[:nnz].copy()
        return rp, col, val, (s.nrows, s.ncols)

    def mask(self, l):
        p = lib().or_mg_mask(self.h, l)
        n = self.level(l)[3][0]
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), (n,)).copy()

    def eigs(self, l):
        a, b = ctypes.c_double(), ctypes.c_double()
        lib().or_mg_eigs(self.h, l, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def set_eigs(self, l, emax, emin):
        lib().or_mg_set_eigs(self.h, l, emax, emin)

    def set_colors(self, l, color, ncolors):
        color = np.ascontiguousarray(color, np.int32)
        lib().or_mg_set_colors(self.h, l, _p(color), int(ncolors))


def jp_color(A, seed=0x5eed):
    """Jones-Plassmann colouring: (colour per row, number of colours)."""
    M = Mat.of(A)
    color = np.zeros(max(A.shape[0], 1), np.int32)
    nc = ctypes.c_int32()
    lib().or_jp_color(M.ref, ctypes.c_uint64(seed), _p(color), ctypes.byref(nc))
    return color[:A.shape[0]], nc.value


def gs_smooth(A, dinv, color, ncolors, b, x0=None, sweeps=1, symmetric=True):
    """Multicolour symmetric Gauss-Seidel sweeps (GaussSeidel::smooth)."""
    M = Mat.of(A)
    dinv = np.ascontiguousarray(dinv, np.float64)
    color = np.ascontiguousarray(color, np.int32)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    lib().or_gs_smooth(M.ref, _p(dinv), _p(color), int(ncolors), _p(b), _p(x), sweeps, 1 if symmetric else 0)
    return x


def mprgp_solve(A, b, lb=None, ub=None, x0=None, atol=1e-9, rtol=0.0, max_it=100000, eps_eig_est=1e-1,
                power_method_max_it=10):
    M = Mat.of(A)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    lb = None if lb is None else np.ascontiguousarray(lb, np.float64)
    ub = None if ub is None else np.ascontiguousarray(ub, np.float64)
    st = or_status()
    lib().or_mprgp_solve(M.ref, _p(b), _p(lb), _p(ub), _p(x), atol, rtol, max_it, eps_eig_est, power_method_max_it,
                         ctypes.byref(st))
    return x, status_dict(st)


def _qp_call(fn, A, b, lb, ub, x0, atol, rtol, stol, max_it):
    M = Mat.of(A)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    lb = None if lb is None else np.ascontiguousarray(lb, np.float64)
    ub = None if ub is None else np.ascontiguousarray(ub, np.float64)
    st = or_status()
    fn(M.ref, _p(b), _p(lb), _p(ub), _p(x), atol, rtol, stol, max_it, ctypes.byref(st))
    return x, status_dict(st)


def mmg_solve(A, b, interps, lb=None, ub=None, x0=None, pre=3, post=3, coarse_space_it=1, active_set_tol=0.0,
              atol=1e-9, rtol=1e-9, stol=1e-11, max_it=1000):
    """MonotoneMultigrid (interps coarse to fine)."""
    M = Mat.of(A)
    Is = [Mat.of(I) for I in interps]
    arr = (ctypes.c_void_p * len(Is))(*[ctypes.addressof(m.s) for m in Is])
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    lb = None if lb is None else np.ascontiguousarray(lb, np.float64)
    ub = None if ub is None else np.ascontiguousarray(ub, np.float64)
    st = or_status()
    lib().or_mmg_solve(M.ref, _p(b), _p(lb), _p(ub), _p(x), len(Is) + 1, arr, pre, post, coarse_space_it,
                       active_set_tol, atol, rtol, stol, max_it, ctypes.byref(st))
    return x, status_dict(st)


def pcg_solve(A, b, lb=None, ub=None, x0=None, atol=1e-9, rtol=1e-9, stol=1e-11, max_it=1000):
    """ProjectedConjugateGradient."""
    return _qp_call(lib().or_pcg_solve, A, b, lb, ub, x0, atol, rtol, stol, max_it)


def pg_solve(A, b, lb=None, ub=None, x0=None, atol=1e-9, rtol=1e-9, stol=1e-11, max_it=1000):
    """ProjectedGradient."""
    return _qp_call(lib().or_pg_solve, A, b, lb, ub, x0, atol, rtol, stol, max_it)


def pgs_solve(A, b, lb=None, ub=None, x0=None, stol=1e-11, max_it=1000, symmetric=True, check_each=10):
    M = Mat.of(A)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    lb = None if lb is None else np.ascontiguousarray(lb, np.float64)
    ub = None if ub is None else np.ascontiguousarray(ub, np.float64)
    st = or_status()
    lib().or_pgs_solve(M.ref, _p(b), _p(lb), _p(ub), _p(x), stol, max_it, 1 if symmetric else 0, check_each,
                       ctypes.byref(st))
    return x, status_dict(st)


def set_threads(n: int):
    lib().or_set_threads(n)


def get_threads() -> int:
    return lib().or_get_threads()


def srand(seed: int = 1):
    ctypes.CDLL(None).srand(ctypes.c_uint(seed))
