"""Kernel-level parity on the GPU: the C ABI against the CPU oracle.

SpMV rows are summed sequentially from LDS, so every y_i must equal the
oracle's sequential CSR loop bit for bit (FP64 and FP32). Pattern work
(transpose, SpGEMM, RAP) must match exactly, and since both sides accumulate
C_ij in the same visiting order, so must the values.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
import utopia_amd as ua
from utopia_amd import petsc_io
from utopia_amd._abi import uhip_cheb_step

pytestmark = pytest.mark.gpu

L = ua.lib


def P(x):
    return ctypes.c_void_p(x.ptr)


def csr_random(n, m, density, seed, empty_rows=(), long_row=None, long_len=0):
    rng = np.random.default_rng(seed)
    M = sp.random(n, m, density=density, random_state=seed, format="lil")
    for r in empty_rows:
        M[r, :] = 0
    if long_row is not None:
        cols = rng.choice(m, size=long_len, replace=False)
        M[long_row, cols] = rng.standard_normal(long_len)
    C = ua.Csr.from_scipy(sp.csr_matrix(M))
    return C


MATRICES = {
    "poisson2d": lambda: ua.gen_stencil(ua.STENCIL_2D5, 37, 29),
    "varcoef3d": lambda: ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 13, 11, 9, seed=3),
    "nine_point": lambda: ua.gen_stencil(ua.STENCIL_2D9, 40, 33, shift=1e-3),
    "prolongation": lambda: ua.gen_prolongation(16, 12, 10),
    "ragged_empty_rows": lambda: csr_random(3000, 2500, 0.003, 1, empty_rows=range(100, 900)),
    "one_row": lambda: csr_random(1, 50, 0.5, 2),
    "rows_up_to_900": lambda: csr_random(400, 5000, 0.1, 3, long_row=17, long_len=900),
}


@pytest.mark.parametrize("name", sorted(MATRICES))
@pytest.mark.parametrize("block_nnz", [1024, 2048, 4096])
def test_spmv_fp64_bit_exact(ctx, name, block_nnz):
    A = MATRICES[name]()
    x = ua.gen_uniform(11, A.shape[1])
    D = ua.DeviceCsr(ctx, A, np.float64, block_nnz)
    xd = ua.DeviceArray.from_numpy(ctx, x)
    yd = ua.DeviceArray(ctx, A.shape[0])
    D.spmv(xd, yd)
    y = yd.numpy()
    ref = orc.spmv(A, x)
    # rows that fit the block's LDS budget (1.25 * block_nnz) are summed sequentially: bit-exact;
    # longer rows are reduced by the whole workgroup: rounding-level agreement
    seq = np.diff(A.rowptr) <= block_nnz + block_nnz // 4
    assert np.array_equal(y[seq], ref[seq]), np.abs(y - ref)[seq].max()
    bound = 1e-14 * (np.abs(A.to_scipy()) @ np.abs(x))
    assert np.all(np.abs(y - ref)[~seq] <= 20 * bound[~seq])
    blocks = D.blocks()
    assert blocks[0] == 0 and blocks[-1] == A.shape[0] and np.all(np.diff(blocks) > 0)


@pytest.mark.parametrize("name", sorted(MATRICES))
@pytest.mark.parametrize("fmt", [ua.FMT_ROWBLOCK, ua.FMT_SELL64])
def test_spmv_fp32_bit_exact(ctx, name, fmt):
    A = MATRICES[name]()
    x = ua.gen_uniform(12, A.shape[1]).astype(np.float32)
    D = ua.DeviceCsr(ctx, A, np.float32, fmt=fmt)
    assert D.format == fmt
    xd = ua.DeviceArray.from_numpy(ctx, x)
    yd = ua.DeviceArray(ctx, A.shape[0], np.float32)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), orc.spmv_f32(A, x))


@pytest.mark.parametrize("name", sorted(MATRICES))
def test_spmv_sell64_fp64_bit_exact(ctx, name):
    A = MATRICES[name]()
    x = ua.gen_uniform(14, A.shape[1])
    D = ua.DeviceCsr(ctx, A, np.float64, fmt=ua.FMT_SELL64)
    assert D.format == ua.FMT_SELL64
    xd = ua.DeviceArray.from_numpy(ctx, x)
    yd = ua.DeviceArray(ctx, A.shape[0])
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), orc.spmv(A, x))


def test_auto_format_choice(ctx):
    # stencils (few distinct offsets col - row) take DIA; block-diagonal systems of
    # different grid sizes (many offsets) take SELL-64; a matrix with one long row
    # per slice keeps row-block CSR; transfers take VI
    D = ua.DeviceCsr(ctx, ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 16, 16, 16, seed=1))
    assert D.format == ua.FMT_DIA and D.m.ndiag == 7
    assert ua.DeviceCsr(ctx, ua.gen_stencil(ua.STENCIL_2D9, 30, 20, shift=1e-3)).m.ndiag == 9
    from utopia_amd import batched as bt
    assert ua.DeviceCsr(ctx, bt.make_batch([9, 13, 17, 21, 25, 29, 33, 37, 41, 45, 49, 53, 57, 61, 65, 69, 73],
                                           levels=1).A).format == ua.FMT_SELL64
    assert ua.DeviceCsr(ctx, MATRICES["rows_up_to_900"]()).format == ua.FMT_ROWBLOCK
    # trilinear transfers: few distinct values, <= 8 / 27 offsets from each row's first
    # column: the restriction fills its 27 slots, the interpolation 3 of 8 (both VI);
    # two entries per row spread over 9 offsets (4.5 slots per entry) are not VI
    P3 = ua.gen_prolongation(16, 16, 16)
    assert ua.DeviceCsr(ctx, P3).format == ua.FMT_VI
    assert ua.DeviceCsr(ctx, ua.Csr.from_scipy(P3.to_scipy().T)).format == ua.FMT_VI
    r = np.arange(200)
    S = sp.csr_matrix((np.tile([1.0, 2.0], 200), np.stack([r, r + 1 + r % 8], 1).ravel(), 2 * np.arange(201)),
                      shape=(200, 300))
    assert ua.DeviceCsr(ctx, ua.Csr.from_scipy(S)).format != ua.FMT_VI


def _galerkin(n):
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=2)
    return ua.Csr(*orc.ptap(A, ua.gen_prolongation(n, n, n)))


DIA_MATRICES = {
    "2d5": lambda: ua.gen_stencil(ua.STENCIL_2D5, 37, 29),
    "3d7_varcoef": lambda: ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 21, 17, 12, seed=5),
    "2d9_shift": lambda: ua.gen_stencil(ua.STENCIL_2D9, 33, 31, shift=1e-3),
    "galerkin27": lambda: _galerkin(18),
    "edgecoef_odd": lambda: ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, 45, 13, seed=3),
    "3d7_planes": lambda: ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 16, 8, 11, seed=6),
    "3d7_planes_one_slab": lambda: ua.gen_stencil(ua.STENCIL_3D7, 8, 8, 3),
    "asym_values": lambda: _perturbed(ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 9, 8, 7, seed=4), 100, 1),
    "asym_pattern": lambda: _dropped(ua.gen_stencil(ua.STENCIL_2D5, 11, 9), 40),
}

# symmetric DIA storage (offsets >= 0 only) is taken exactly when the values are
# symmetric bit for bit
DIA_SYMMETRIC = {"2d5": 1, "3d7_varcoef": 1, "2d9_shift": 1, "edgecoef_odd": 1, "asym_values": 0,
                 "asym_pattern": 0, "3d7_planes": 2, "3d7_planes_one_slab": 2}


def _perturbed(A, row, k):
    """A with entry k of row `row` (an off-diagonal) scaled by 1 + 2^-20 (kept in FP32 too)."""
    val = A.val.copy()
    e = A.rowptr[row] + k
    assert A.col[e] != row
    val[e] *= 1 + 2.0 ** -20
    return ua.Csr(A.rowptr, A.col, val, A.shape)


def _dropped(A, row):
    """A without the first entry of row `row` (an off-diagonal): pattern no longer symmetric."""
    e = A.rowptr[row]
    assert A.col[e] != row
    keep = np.ones(A.nnz, bool)
    keep[e] = False
    rp = A.rowptr.copy()
    rp[row + 1:] -= 1
    return ua.Csr(rp, A.col[keep], A.val[keep], A.shape)


@pytest.mark.parametrize("name", sorted(DIA_MATRICES))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_spmv_dia_bit_exact(ctx, name, dt):
    A = DIA_MATRICES[name]()
    n = A.shape[0]
    D = ua.DeviceCsr(ctx, A, dtype=dt, fmt=ua.FMT_DIA)
    assert D.format == ua.FMT_DIA
    if name in DIA_SYMMETRIC:
        # 2: the plane-walking 7-point layout, which every product of either precision runs
        assert D.m.dia_sym == DIA_SYMMETRIC[name]
    x = ua.gen_uniform(17, n)
    if dt is np.float32:
        x = x.astype(np.float32)
        ref = orc.spmv_f32(A, x)
    else:
        ref = orc.spmv(A, x)
    xd = ua.DeviceArray.from_numpy(ctx, x)
    yd = ua.DeviceArray(ctx, n, dt)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), ref)


def _galerkin_dims(nx, ny, nz, seed=2):
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, nx, ny, nz, seed=seed)
    return ua.Csr(*orc.ptap(A, ua.gen_prolongation(nx, ny, nz)))


def _symmetrized(A):
    """(A + A^T) / 2: symmetric bit for bit (a_ij + a_ji and a_ji + a_ij round alike)."""
    S = A.to_scipy()
    return ua.Csr.from_scipy(((S + S.T) * 0.5).tocsr())


def _with_nan(A, row):
    val = A.val.copy()
    val[A.rowptr[row] + 2] = np.nan
    return ua.Csr(A.rowptr, A.col, val, A.shape)


def _sign_flipped(A, row):
    """A with the first (lower) entry of `row` negated: no longer symmetric up to rounding."""
    val = A.val.copy()
    assert A.col[A.rowptr[row]] < row
    val[A.rowptr[row]] = -val[A.rowptr[row]]
    return ua.Csr(A.rowptr, A.col, val, A.shape)


# 27-point operators whose planes are whole multiples of 64 rows (the plane-walking kernel):
# (coarse grid, expected FP64 layout: "full", "sym", or "delta" = the upper half plus the lower
# values' bit-pattern differences from their mirrors). Galerkin products are symmetric only to
# rounding (delta); symmetrized ones and the constant box stencil are symmetric bit for bit; a
# NaN value or a lower entry of the wrong sign keeps the full layout. Plane counts leave
# partial 4-plane slabs; 16 x 16 and 32 x 32 planes span several 64-row segments (the mirrors
# one y line and one plane back come from other waves).
DIA27 = {
    "galerkin_8x8x11": (lambda: _galerkin_dims(16, 16, 22), "delta"),
    "galerkin_16x16x20": (lambda: _galerkin_dims(32, 32, 40), "delta"),
    "galerkin_32x32x6": (lambda: _galerkin_dims(64, 64, 12), "delta"),
    "galerkin_sym_16x8x10": (lambda: _symmetrized(_galerkin_dims(32, 16, 20)), "sym"),
    "box27_16x8x13": (lambda: _box27(16, 8, 13), "sym"),
    "galerkin_nan_8x8x11": (lambda: _with_nan(_galerkin_dims(16, 16, 22), 77), "full"),
    "galerkin_signflip_16x16x20": (lambda: _sign_flipped(_galerkin_dims(32, 32, 40), 1000), "full"),
    "random_delta32_16x16x20": (lambda: _random27(16, 16, 20, seed=4, layout="delta32"), "delta"),
}


@pytest.mark.parametrize("name", sorted(DIA27))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_spmv_dia27_walk_bit_exact(ctx, name, dt):
    """The plane-walking 27-point kernel (full layout selecting on NaN-encoded absent entries or
    on the mask, symmetric layout with mirrors, the upper half with the lower values' bit-pattern
    differences): y = A x equal to the CSR sum bit for bit; a NaN entry keeps the mask and gives
    NaN exactly in its row."""
    make, layout = DIA27[name]
    A = make()
    n = A.shape[0]
    D = ua.DeviceCsr(ctx, A, dtype=dt, fmt=ua.FMT_DIA)
    assert D.format == ua.FMT_DIA and D.m.ndiag == 27 and D.m.dia_w27 == 1
    assert D.m.dia_sym == (0 if layout == "full" else 1)
    if dt is np.float64:
        assert bool(D.m.dia_delta) == (layout == "delta")
        assert D.m.dia_nan == (0 if "nan" in name else 1)
    elif "delta32" in name:
        assert D.m.dia_delta  # FP32 values a few ulps apart: int8 deltas
    x = ua.gen_uniform(17, n)
    if dt is np.float32:
        x = x.astype(np.float32)
        ref = orc.spmv_f32(A, x)
    else:
        ref = orc.spmv(A, x)
    xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray(ctx, n, dt)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), ref, equal_nan=True)


@pytest.mark.parametrize("name", ["galerkin_16x16x20", "galerkin_sym_16x8x10", "box27_16x8x13",
                                  "galerkin_signflip_16x16x20", "random_delta32_16x16x20"])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_dia27_walk_epilogues(ctx, name, dt):
    """Every fused epilogue on the 27-point kernel against the row-block plan of the same matrix
    (bit-exact to the CSR sums): residual, masked, add, axpby, axpbz, the product with its dot,
    and a middle and a last Chebyshev step (x taken from the kernel's registers), with z.r."""
    A = DIA27[name][0]()
    n = A.shape[0]
    D = ua.DeviceCsr(ctx, A, dtype=dt, fmt=ua.FMT_DIA)
    R = ua.DeviceCsr(ctx, A, dtype=dt, fmt=ua.FMT_ROWBLOCK)
    assert D.m.dia_w27 == 1
    code = ua.F64 if dt is np.float64 else ua.F32
    x, b, z, m, pp, rp = (ua.gen_uniform(s, n).astype(dt) for s in (1, 2, 3, 4, 5, 6))
    m = (m > 0).astype(dt)
    dw = ua.gen_uniform(7, n)
    h = ctx.h
    dev = lambda v: ua.DeviceArray.from_numpy(ctx, v)  # noqa: E731
    xd, bd, md, ppd, rpd, dwd = (dev(v) for v in (x, b, m, pp, rp, dw))

    def both(run):
        outs = []
        for M in (D, R):
            o = [dev(z), dev(z), dev(z)]
            dd = ua.DeviceArray(ctx, 1)
            run(M, o, dd)
            outs.append(([v.numpy() for v in o], dd.numpy()[0]))
        return outs

    for run in (lambda M, o, dd: ua.check(L().uhip_spmv_residual(h, M.ref, P(xd), P(bd), code, P(o[0]))),
                lambda M, o, dd: ua.check(L().uhip_spmv_masked(h, M.ref, P(xd), P(md), P(o[0]))),
                lambda M, o, dd: ua.check(L().uhip_spmv_add(h, M.ref, P(xd), P(o[0]))),
                lambda M, o, dd: ua.check(L().uhip_spmv_axpby(h, M.ref, 2.5, P(xd), -0.5, P(o[0]))),
                lambda M, o, dd: ua.check(L().uhip_spmv_axpbz(h, M.ref, 1.0, P(xd), -1.0, P(bd), P(o[0])))):
        (o1, _), (o2, _) = both(run)
        assert np.array_equal(o1[0], o2[0])
    # A x with x.(A x)
    s = ctx.scalars
    outs = []
    for M in (D, R):
        o = ua.DeviceArray(ctx, n, dt)
        ua.check(L().uhip_spmv_dot(h, M.ref, P(xd), P(o), ctypes.c_void_p(s)))
        outs.append((o.numpy(), ctx.read_scalars(s, 1)[0]))
    assert np.array_equal(outs[0][0], outs[1][0])
    Ax = outs[1][0].astype(np.float64)
    bound = 2 * n * np.finfo(dt).eps * float(np.abs(x.astype(np.float64)) @ np.abs(Ax))
    assert abs(outs[0][1] - float(x.astype(np.float64) @ Ax)) <= bound

    # Chebyshev: a middle step (x, r, p out) and a last step with z.r
    def cheb_mid(M, o, dd):
        st = uhip_cheb_step(pprev=ppd.ptr, p_dtype=code, rprev=rpd.ptr, xin=xd.ptr, xout=o[0].ptr, x_dtype=code,
                            rout=o[1].ptr, pout=o[2].ptr, a_prev=0.41, beta=0.073, a_cur=0.29, last=0,
                            dot_with=None, dot_d=None)
        ua.check(L().uhip_cheb_step_apply(h, M.ref, ctypes.byref(st)))

    def cheb_last(M, o, dd):
        st = uhip_cheb_step(pprev=ppd.ptr, p_dtype=code, rprev=None, xin=xd.ptr, xout=o[0].ptr, x_dtype=code,
                            rout=None, pout=None, a_prev=0.41, beta=0.073, a_cur=0.29, last=1, dot_with=dwd.ptr,
                            dot_d=dd.ptr)
        ua.check(L().uhip_cheb_step_apply(h, M.ref, ctypes.byref(st)))

    (o1, _), (o2, _) = both(cheb_mid)
    for a, c in zip(o1, o2):
        assert np.array_equal(a, c)
    (o1, d1), (o2, d2) = both(cheb_last)
    assert np.array_equal(o1[0], o2[0])
    zr = o1[0].astype(np.float64)
    bound = 2 * n * np.finfo(np.float64).eps * float(np.abs(zr) @ np.abs(dw))
    assert abs(d1 - float(zr @ dw)) <= bound and abs(d2 - float(zr @ dw)) <= bound


def _random27(nx, ny, nz, seed, layout):
    """The 27-point pattern of an nx x ny x nz grid with random values: "full" as drawn, "sym"
    symmetrized bit for bit, "delta" symmetrized and then each lower value moved by up to 3 ulps,
    "delta32" the same with FP32 values moved by up to 3 FP32 ulps (held exactly in FP64)."""
    T = lambda k: sp.diags([np.ones(k - 1), np.ones(k), np.ones(k - 1)], [-1, 0, 1])  # noqa: E731
    B = sp.kron(sp.kron(T(nz), T(ny)), T(nx)).tocsr()
    rng = np.random.default_rng(seed)
    B.data = rng.uniform(-1.0, 1.0, B.nnz)
    if layout != "full":
        B = ((B + B.T) * 0.5).tocsr()
        B.sort_indices()
    rows = np.repeat(np.arange(B.shape[0]), np.diff(B.indptr))
    lower = B.indices < rows
    if layout == "delta":
        bits = B.data.view(np.int64)
        bits[lower] += rng.integers(-3, 4, int(lower.sum()))
    elif layout == "delta32":
        v32 = B.data.astype(np.float32)
        v32.view(np.int32)[lower] += rng.integers(-3, 4, int(lower.sum())).astype(np.int32)
        B.data = v32.astype(np.float64)
    return ua.Csr.from_scipy(B)


@pytest.mark.parametrize("dims,layout,dt", [((96, 96, 96), "full", np.float64), ((112, 112, 112), "sym", np.float64),
                                            ((112, 112, 112), "delta", np.float64),
                                            ((136, 136, 136), "delta32", np.float32)])
def test_dia27_walk_streams_large_levels(ctx, dims, layout, dt):
    """Value arrays above 128 MB are streamed (non-temporal loads) instead of read through the
    caches: FP64 27-point operators of 0.88 M rows (full layout, 191 MB) and 1.40 M rows
    (symmetric layout, 157 MB, with and without lower-triangle deltas), and an FP32 one of
    2.52 M rows (symmetric with int8 deltas, 141 MB) give the CSR sums bit for bit, plain and as
    a residual."""
    A = _random27(*dims, seed=3, layout=layout)
    n = A.shape[0]
    vb = np.dtype(dt).itemsize
    D = ua.DeviceCsr(ctx, A, dtype=dt, fmt=ua.FMT_DIA)
    assert D.m.dia_w27 == 1 and D.m.dia_sym == (0 if layout == "full" else 1)
    assert bool(D.m.dia_delta) == layout.startswith("delta")
    assert (27 if layout == "full" else 14) * n * vb > 128 << 20
    x, b = ua.gen_uniform(5, n).astype(dt), ua.gen_uniform(6, n).astype(dt)
    Ax = orc.spmv(A, x) if dt is np.float64 else orc.spmv_f32(A, x)
    xd, bd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray.from_numpy(ctx, b), ua.DeviceArray(ctx, n, dt)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), Ax)
    D.residual(xd, bd, yd)
    assert np.array_equal(yd.numpy(), b - Ax)


def test_dia27_autotune_keeps_a_valid_plan(ctx):
    """The autotuner times the 27-point DIA plan with and without the plane walk (and with it the
    near-symmetric layout); whichever it keeps computes the CSR sum bit for bit. An automatic plan
    walks only grids with enough waves to fill the GPU (1024): a 16 x 8 x 13 grid has 8."""
    B = ua.DeviceCsr(ctx, _box27(16, 8, 13))
    assert B.format == ua.FMT_DIA and B.m.dia_w27 == 0
    assert ua.DeviceCsr(ctx, _box27(16, 8, 13), fmt=ua.FMT_DIA).m.dia_w27 == 1
    A = _galerkin_dims(32, 32, 40)
    n = A.shape[0]
    fmt = ctypes.c_int32()
    ms = ctypes.c_float()
    D = ua.DeviceCsr(ctx, A, fmt=ua.FMT_DIA)
    ua.check(L().uhip_csr_autotune(ctx.h, D.ref, 3, ctypes.byref(fmt), ctypes.byref(ms)))
    assert fmt.value == D.format
    if D.format == ua.FMT_DIA:
        assert D.m.dia_w27 in (0, 1) and (D.m.dia_s2 == 256) == (D.m.dia_w27 == 1)
    x = ua.gen_uniform(9, n)
    xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray(ctx, n)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), orc.spmv(A, x))


def test_dia_skips_absent_entries_exactly(ctx):
    # x holds inf at a column that a row's padded (absent) diagonal slot would touch:
    # the masked sum must equal the CSR sum (finite wherever the CSR row avoids it)
    A = ua.gen_stencil(ua.STENCIL_2D5, 8, 8)
    x = ua.gen_uniform(3, 64)
    x[8] = np.inf  # row 7 (end of grid line 0) has no entry at column 8, offset +1 is absent there
    D = ua.DeviceCsr(ctx, A, fmt=ua.FMT_DIA)
    xd = ua.DeviceArray.from_numpy(ctx, x)
    yd = ua.DeviceArray(ctx, 64)
    D.spmv(xd, yd)
    y = yd.numpy()
    ref = orc.spmv(A, x)
    assert np.isfinite(y[7]) and np.array_equal(np.isfinite(y), np.isfinite(ref))
    assert np.array_equal(y[np.isfinite(ref)], ref[np.isfinite(ref)])
    # the mirrored lower diagonal: row 8 (start of grid line 1) has no entry at column 7
    assert D.m.dia_sym == 1
    x = ua.gen_uniform(4, 64)
    x[7] = np.inf
    xd = ua.DeviceArray.from_numpy(ctx, x)
    D.spmv(xd, yd)
    y = yd.numpy()
    ref = orc.spmv(A, x)
    assert np.isfinite(y[8]) and np.array_equal(np.isfinite(y), np.isfinite(ref))
    assert np.array_equal(y[np.isfinite(ref)], ref[np.isfinite(ref)])
    # full layout of a non-symmetric matrix, absent entries NaN-encoded and selected on the value
    N = _perturbed(A, 10, 3)
    D = ua.DeviceCsr(ctx, N, fmt=ua.FMT_DIA)
    assert D.m.dia_sym == 0 and D.m.dia_nan == 1
    for col in (7, 8):
        x = ua.gen_uniform(5, 64)
        x[col] = np.inf
        D.spmv(ua.DeviceArray.from_numpy(ctx, x), yd)
        y, ref = yd.numpy(), orc.spmv(N, x)
        assert np.array_equal(np.isfinite(y), np.isfinite(ref))
        assert np.array_equal(y[np.isfinite(ref)], ref[np.isfinite(ref)])


def test_dia_rejects_unsorted_and_rectangular(ctx):
    A = ua.gen_stencil(ua.STENCIL_2D5, 6, 6)
    col = A.col.copy()
    col[A.rowptr[10]], col[A.rowptr[10] + 1] = col[A.rowptr[10] + 1], col[A.rowptr[10]]
    val = A.val.copy()
    val[A.rowptr[10]], val[A.rowptr[10] + 1] = val[A.rowptr[10] + 1], val[A.rowptr[10]]
    U = ua.Csr(A.rowptr, col, val, A.shape)
    with pytest.raises(ua.HipError):
        ua.DeviceCsr(ctx, U, fmt=ua.FMT_DIA)
    assert ua.DeviceCsr(ctx, U).format != ua.FMT_DIA  # automatic choice falls back
    with pytest.raises(ua.HipError):
        ua.DeviceCsr(ctx, ua.gen_prolongation(8, 8), fmt=ua.FMT_DIA)


@pytest.mark.parametrize("fmt,dims", [(ua.FMT_ROWBLOCK, (20, 18, 13)), (ua.FMT_SELL64, (20, 18, 13)),
                                      (ua.FMT_DIA, (20, 18, 13)), (ua.FMT_DIA, (16, 16, 13)),
                                      (ua.FMT_DIA, (256, 256, 21))])
def test_fused_epilogues_both_formats_with_dot(ctx, fmt, dims):
    """(256, 256, 21): the benchmark's plane size (S2 = 65,536) with 2 full 8-plane slabs and a
    partial one in FP64, 5 full 4-plane slabs and a partial one in FP32."""
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, *dims, seed=9)
    n = A.shape[0]
    x, b = ua.gen_uniform(1, n), ua.gen_uniform(2, n)
    Ax = orc.spmv(A, x)
    D = ua.DeviceCsr(ctx, A, fmt=fmt)
    D32 = ua.DeviceCsr(ctx, A, dtype=np.float32, fmt=fmt)
    if fmt == ua.FMT_DIA and dims[0] * dims[1] % 64 == 0:
        # the plane-walking 7-point kernel runs every epilogue in both precisions, here with
        # a partial last slab
        assert D32.m.dia_sym == 2 and D.m.dia_sym == 2 and D.m.dia_nan == 1
    x32, b32 = x.astype(np.float32), b.astype(np.float32)
    x32d, b32d = ua.DeviceArray.from_numpy(ctx, x32), ua.DeviceArray.from_numpy(ctx, b32)
    out32 = ua.DeviceArray(ctx, n, np.float32)
    D32.residual(x32d, b32d, out32)
    assert np.array_equal(out32.numpy(), b32 - orc.spmv_f32(A, x32))
    xd, bd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray.from_numpy(ctx, b)
    out = ua.DeviceArray(ctx, n)
    D.residual(xd, bd, out)
    assert np.array_equal(out.numpy(), b - Ax)
    s = ctx.scalars
    ua.check(L().uhip_spmv_dot(ctx.h, D.ref, P(xd), P(out), ctypes.c_void_p(s)))
    assert np.array_equal(out.numpy(), Ax)
    assert abs(ctx.read_scalars(s, 1)[0] - np.dot(x, Ax)) <= 1e-13 * np.dot(np.abs(x), np.abs(Ax))


def test_spmv_long_rows_take_block_reduction(ctx):
    # rows longer than the LDS budget are reduced by a whole workgroup (tree order)
    A = csr_random(64, 20000, 0.001, 5, long_row=5, long_len=15000)
    x = ua.gen_uniform(13, A.shape[1])
    D = ua.DeviceCsr(ctx, A, np.float64, 1024)
    xd = ua.DeviceArray.from_numpy(ctx, x)
    yd = ua.DeviceArray(ctx, A.shape[0])
    D.spmv(xd, yd)
    y, ref = yd.numpy(), orc.spmv(A, x)
    bound = 1e-14 * (np.abs(A.to_scipy()) @ np.abs(x)) + 1e-300
    assert np.all(np.abs(y - ref) <= 20 * bound)
    mask = np.ones(A.shape[0], bool)
    mask[5] = False
    assert np.array_equal(y[mask], ref[mask])


def test_spmv_empty_matrix_rows(ctx):
    A = ua.Csr(np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0), (5, 4))
    D = ua.DeviceCsr(ctx, A)
    xd = ua.DeviceArray.from_numpy(ctx, np.ones(4))
    yd = ua.DeviceArray.from_numpy(ctx, np.full(5, 7.0))
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), np.zeros(5))


@pytest.mark.parametrize("dims", [(12, 10, 8), (16, 8, 10)])
def test_fused_spmv_epilogues(ctx, dims):
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, *dims, seed=9)
    n = A.shape[0]
    x, b, z, m = (ua.gen_uniform(s, n) for s in (1, 2, 3, 4))
    m = (m > 0).astype(np.float64)
    Ax = orc.spmv(A, x)
    D = ua.DeviceCsr(ctx, A)
    xd, bd, zd, md = (ua.DeviceArray.from_numpy(ctx, v) for v in (x, b, z, m))
    out = ua.DeviceArray(ctx, n)
    h = ctx.h
    ua.check(L().uhip_spmv_residual(h, D.ref, P(xd), P(bd), ua.F64, P(out)))
    assert np.array_equal(out.numpy(), b - Ax)
    ua.check(L().uhip_spmv_masked(h, D.ref, P(xd), P(md), P(out)))
    assert np.array_equal(out.numpy(), m * Ax)
    out.upload(z)
    ua.check(L().uhip_spmv_add(h, D.ref, P(xd), P(out)))
    assert np.array_equal(out.numpy(), z + Ax)
    out.upload(z)
    ua.check(L().uhip_spmv_axpby(h, D.ref, 2.5, P(xd), -0.5, P(out)))
    assert np.array_equal(out.numpy(), 2.5 * Ax + (-0.5) * z)
    ua.check(L().uhip_spmv_axpbz(h, D.ref, 1.0, P(xd), -1.0, P(bd), P(out)))
    assert np.array_equal(out.numpy(), Ax - b)
    s = ctx.scalars
    ua.check(L().uhip_spmv_dot(h, D.ref, P(xd), P(out), ctypes.c_void_p(s)))
    assert np.array_equal(out.numpy(), Ax)
    got = ctx.read_scalars(s, 1)[0]
    assert abs(got - np.dot(x, Ax)) <= 1e-13 * np.dot(np.abs(x), np.abs(Ax))


def test_mixed_residual_fp64_rhs_on_fp32_matrix(ctx):
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 10, 9, 8, seed=4)
    n = A.shape[0]
    x = ua.gen_uniform(5, n).astype(np.float32)
    b = ua.gen_uniform(6, n)
    D = ua.DeviceCsr(ctx, A, np.float32)
    xd, bd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray.from_numpy(ctx, b)
    r = ua.DeviceArray(ctx, n, np.float32)
    D.residual(xd, bd, r)
    assert np.array_equal(r.numpy(), b.astype(np.float32) - orc.spmv_f32(A, x))


def test_blas1_and_reductions(ctx):
    n = 100_003
    x, y = ua.gen_uniform(21, n), ua.gen_uniform(22, n)
    h, l = ctx.h, L()
    xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray.from_numpy(ctx, y)
    w = ua.DeviceArray(ctx, n)
    ua.check(l.uhip_vec_waxpby_f64(h, n, 2.0, P(xd), -3.0, P(yd), P(w)))
    assert np.array_equal(w.numpy(), 2.0 * x + (-3.0) * y)
    w.upload(y)
    ua.check(l.uhip_vec_axpy_f64(h, n, 0.25, P(xd), P(w)))
    assert np.array_equal(w.numpy(), y + 0.25 * x)
    w.upload(y)
    ua.check(l.uhip_vec_axpby_f64(h, n, 0.25, P(xd), 3.0, P(w)))
    assert np.array_equal(w.numpy(), 0.25 * x + 3.0 * y)
    ua.check(l.uhip_vec_emul_f64(h, n, P(xd), P(yd), P(w)))
    assert np.array_equal(w.numpy(), x * y)
    ua.check(l.uhip_vec_ediv_f64(h, n, P(xd), P(yd), P(w)))
    assert np.array_equal(w.numpy(), x / y)
    xs = x.copy()
    xs[::7] = 1e-17
    xsd = ua.DeviceArray.from_numpy(ctx, xs)
    ua.check(l.uhip_vec_pseudo_inv_f64(h, n, P(xsd), P(w), np.finfo(float).eps))
    ref = np.where(np.abs(xs) > np.finfo(float).eps, 1.0 / np.where(xs == 0, 1, xs), 0.0)
    assert np.array_equal(w.numpy(), ref)
    w.upload(x)
    ua.check(l.uhip_vec_shift_f64(h, n, -1.0, P(w)))
    ua.check(l.uhip_vec_abs_f64(h, n, P(w), P(w)))
    assert np.array_equal(w.numpy(), np.abs(x + (-1.0)))
    ua.check(l.uhip_vec_scale_f64(h, n, -2.0, P(w)))
    ua.check(l.uhip_vec_set_f64(h, 17, 3.5, P(w)))
    assert np.array_equal(w.numpy()[:17], np.full(17, 3.5))
    s = ctx.scalars
    arr = (ctypes.c_void_p * 4)
    xs4 = arr(xd.ptr, xd.ptr, yd.ptr, xd.ptr)
    ys4 = arr(yd.ptr, xd.ptr, yd.ptr, xd.ptr)
    ua.check(l.uhip_vec_dots_f64(h, n, 4, xs4, ys4, ctypes.c_void_p(s)))
    got = ctx.read_scalars(s, 4)
    exact = [np.dot(x, y), np.dot(x, x), np.dot(y, y), np.dot(x, x)]
    for g, e, a in zip(got, exact, [np.abs(x) @ np.abs(y), x @ x, y @ y, x @ x]):
        assert abs(g - e) <= 1e-13 * a
    ua.check(l.uhip_vec_norm2_f64(h, n, P(xd), ctypes.c_void_p(s)))
    assert abs(ctx.read_scalars(s, 1)[0] - np.linalg.norm(x)) <= 1e-14 * np.linalg.norm(x)
    bad = x.copy()
    bad[[3, 99]] = [np.nan, np.inf]
    bd = ua.DeviceArray.from_numpy(ctx, bad)
    for op, ref in [(0, x.sum()), (1, x.min()), (2, x.max()), (3, np.abs(x).max())]:
        ua.check(l.uhip_vec_reduce_f64(h, n, P(xd), op, ctypes.c_void_p(s)))
        v = ctx.read_scalars(s, 1)[0]
        assert (v == ref) if op else abs(v - ref) <= 1e-12 * np.abs(x).sum()
    ua.check(l.uhip_vec_reduce_f64(h, n, P(bd), 4, ctypes.c_void_p(s)))
    assert ctx.read_scalars(s, 1)[0] == 2
    # deterministic: the same reduction twice gives the same bits
    ua.check(l.uhip_vec_dot_f64(h, n, P(xd), P(yd), ctypes.c_void_p(s)))
    a1 = ctx.read_scalars(s, 1)[0]
    ua.check(l.uhip_vec_dot_f64(h, n, P(xd), P(yd), ctypes.c_void_p(s)))
    assert ctx.read_scalars(s, 1)[0] == a1


def test_axpy_cast_boundary(ctx):
    n = 5000
    p = ua.gen_uniform(31, n)
    x = ua.gen_uniform(32, n).astype(np.float32)
    pd, xd = ua.DeviceArray.from_numpy(ctx, p), ua.DeviceArray.from_numpy(ctx, x)
    out = ua.DeviceArray(ctx, n)
    ua.check(L().uhip_vec_axpy_cast(ctx.h, n, 0.5, P(pd), ua.F64, P(xd), ua.F32, P(out), ua.F64))
    assert np.array_equal(out.numpy(), x.astype(np.float64) + 0.5 * p)
    out32 = ua.DeviceArray(ctx, n, np.float32)
    ua.check(L().uhip_vec_axpy_cast(ctx.h, n, 0.5, P(pd), ua.F64, None, ua.F32, P(out32), ua.F32))
    assert np.array_equal(out32.numpy(), (np.float32(0.5) * p.astype(np.float32)).astype(np.float32))


def _owned(ctx, m):
    C = ua.csr_from_raw(ctx, m)
    ua.check(L().uhip_csr_free(ctx.h, ctypes.byref(m)))
    return C


@pytest.mark.parametrize("case", ["varcoef3d", "fixture_fA"])
def test_transpose_spgemm_ptap_match_oracle_exactly(ctx, golden, case):
    if case == "varcoef3d":
        A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 8, 8, 8, seed=2)
        Pm = ua.gen_prolongation(8, 8, 8)
    else:
        A = petsc_io.read_mat(golden / "utopia_data" / "f_A")
        Pm = petsc_io.read_mat(golden / "utopia_data" / "I_3")
    DA, DP = ua.DeviceCsr(ctx, A), ua.DeviceCsr(ctx, Pm)
    T = ua.uhip_csr()
    ua.check(L().uhip_csr_transpose(ctx.h, DP.ref, ctypes.byref(T)))
    Th = _owned(ctx, T)
    rp, col, val, shape = orc.transpose(Pm)
    assert Th.shape == shape and np.array_equal(Th.rowptr, rp) and np.array_equal(Th.col, col)
    assert np.array_equal(Th.val, val)
    C = ua.uhip_csr()
    ua.check(L().uhip_csr_spgemm(ctx.h, DA.ref, DP.ref, ctypes.byref(C)))
    Ch = _owned(ctx, C)
    rp, col, val, shape = orc.spgemm(A, Pm)
    assert np.array_equal(Ch.rowptr, rp) and np.array_equal(Ch.col, col) and np.array_equal(Ch.val, val)
    Ac = ua.uhip_csr()
    ua.check(L().uhip_csr_ptap(ctx.h, DA.ref, DP.ref, ctypes.byref(Ac)))
    Ah = _owned(ctx, Ac)
    rp, col, val, shape = orc.ptap(A, Pm)
    assert np.array_equal(Ah.rowptr, rp) and np.array_equal(Ah.col, col) and np.array_equal(Ah.val, val)


def test_spgemm_wide_rows_use_large_workgroups(ctx):
    # candidate products per row > 1024 select the 256-thread / 4096-8192 entry path
    A = csr_random(300, 300, 0.2, 8)
    B = csr_random(300, 400, 0.1, 9)
    DA, DB = ua.DeviceCsr(ctx, A), ua.DeviceCsr(ctx, B)
    C = ua.uhip_csr()
    ua.check(L().uhip_csr_spgemm(ctx.h, DA.ref, DB.ref, ctypes.byref(C)))
    Ch = _owned(ctx, C)
    rp, col, val, _ = orc.spgemm(A, B)
    assert np.array_equal(Ch.rowptr, rp) and np.array_equal(Ch.col, col) and np.array_equal(Ch.val, val)


@pytest.mark.parametrize("n", [1, 63, 200, 333])
def test_dense_cholesky_and_inverse(ctx, n):
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    Ad = ua.DeviceArray.from_numpy(ctx, A.ravel())
    info = ctypes.c_int()
    ua.check(L().uhip_dense_cholesky(ctx.h, n, P(Ad), n, ctypes.byref(info)))
    assert info.value == 0
    Lg = np.tril(Ad.numpy().reshape(n, n))
    Lref = np.linalg.cholesky(A)
    assert np.abs(Lg - Lref).max() <= 1e-12 * np.abs(Lref).max()
    Inv = ua.DeviceArray(ctx, n * n)
    ua.check(L().uhip_dense_cholesky_inverse(ctx.h, n, P(Ad), n, P(Inv)))
    Ainv = Inv.numpy().reshape(n, n)
    assert np.abs(Ainv @ A - np.eye(n)).max() <= 1e-12
    x = rng.standard_normal(n)
    xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray(ctx, n)
    ua.check(L().uhip_dense_gemv(ctx.h, n, P(Inv), P(xd), ua.F64, P(yd), ua.F64))
    assert np.abs(yd.numpy() - Ainv @ x).max() <= 1e-13 * np.abs(Ainv).sum(1).max() * np.abs(x).max()


def test_dense_cholesky_and_inverse_at_the_coarse_size(ctx):
    """n = 4096, the coarsest operator of the 256^3 hierarchy: here the 16^3 variable-coefficient
    operator made dense (condition ~1e3). Backward error of the factor, and the inverse against
    the identity within the conditioning bound; the GEMV against numpy."""
    import scipy.sparse.linalg as spla
    S = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 16, 16, 16, seed=3).to_scipy()
    n = S.shape[0]
    A = S.toarray()
    lmax = spla.eigsh(S, k=1, which="LA", return_eigenvectors=False)[0]
    lmin = spla.eigsh(S.tocsc(), k=1, sigma=0, which="LM", return_eigenvectors=False)[0]
    Ad = ua.DeviceArray.from_numpy(ctx, A.ravel())
    info = ctypes.c_int()
    ua.check(L().uhip_dense_cholesky(ctx.h, n, P(Ad), n, ctypes.byref(info)))
    assert info.value == 0
    Lg = np.tril(Ad.numpy().reshape(n, n))
    assert np.linalg.norm(Lg @ Lg.T - A) <= 1e-14 * np.linalg.norm(A)
    Inv = ua.DeviceArray(ctx, n * n)
    ua.check(L().uhip_dense_cholesky_inverse(ctx.h, n, P(Ad), n, P(Inv)))
    Ainv = Inv.numpy().reshape(n, n)
    assert np.abs(Ainv - Ainv.T).max() <= 1e-14 * np.abs(Ainv).max() * (lmax / lmin)
    assert np.abs(Ainv @ A - np.eye(n)).max() <= 1e-15 * n * (lmax / lmin)
    x = ua.gen_uniform(9, n)
    xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray(ctx, n)
    ua.check(L().uhip_dense_gemv(ctx.h, n, P(Inv), P(xd), ua.F64, P(yd), ua.F64))
    assert np.abs(yd.numpy() - Ainv @ x).max() <= 1e-14 * np.abs(Ainv).sum(1).max() * np.abs(x).max()


def test_dense_cholesky_reports_indefinite(ctx):
    A = np.diag([1.0, 2.0, -1.0, 4.0])
    Ad = ua.DeviceArray.from_numpy(ctx, A.ravel())
    info = ctypes.c_int()
    ua.check(L().uhip_dense_cholesky(ctx.h, 4, P(Ad), 4, ctypes.byref(info)))
    assert info.value == 3


def test_graph_coloring_equals_oracle(ctx):
    for A in (ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 12, 11, 9, seed=2), ua.gen_stencil(ua.STENCIL_2D9, 40, 31)):
        D = ua.DeviceCsr(ctx, A)
        col = ua.DeviceArray(ctx, A.shape[0], np.int32)
        nc = ctypes.c_int32()
        ua.check(L().uhip_graph_color(ctx.h, D.ref, ctypes.c_uint64(0x5eed), P(col), ctypes.byref(nc)))
        co, nco = orc.jp_color(A, 0x5eed)
        assert nc.value == nco and np.array_equal(col.numpy(), co)


def test_graph_coloring_is_proper(ctx):
    A = ua.gen_stencil(ua.STENCIL_2D9, 60, 50, shift=1e-3)
    D = ua.DeviceCsr(ctx, A)
    n = A.shape[0]
    color = ua.DeviceArray(ctx, n, np.int32)
    k = ctypes.c_int32()
    ua.check(L().uhip_graph_color(ctx.h, D.ref, ctypes.c_uint64(7), P(color), ctypes.byref(k)))
    c = color.numpy()
    assert c.min() == 0 and c.max() == k.value - 1
    rows = np.repeat(np.arange(n), np.diff(A.rowptr))
    off = rows != A.col
    assert not np.any(c[rows[off]] == c[A.col[off]])
    assert k.value <= 9  # at most max degree + 1 for the 9-point stencil


def test_autotuner_picks_a_plan_and_keeps_results(ctx):
    """The autotuner times every applicable layout and keeps one; row sums stay
    the sequential CSR sums whichever it picks."""
    for A in (ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, 40, 32, 24, seed=3), MATRICES["rows_up_to_900"](),
              ua.gen_prolongation(24, 24, 24)):
        D = ua.DeviceCsr(ctx, A, fmt=ua.FMT_ROWBLOCK, block_nnz=2048)
        fmt = ctypes.c_int32()
        ms = ctypes.c_float()
        ua.check(L().uhip_csr_autotune(ctx.h, D.ref, 3, ctypes.byref(fmt), ctypes.byref(ms)))
        assert fmt.value in (ua.FMT_ROWBLOCK, ua.FMT_SELL64, ua.FMT_DIA, ua.FMT_VI) and fmt.value == D.format
        assert 0 < ms.value < 100
        x = ua.gen_uniform(5, A.shape[1])
        xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray(ctx, A.shape[0])
        D.spmv(xd, yd)
        y, ref = yd.numpy(), orc.spmv(A, x)
        short = np.diff(A.rowptr) <= 256  # long rows may take the block reduction (tree order)
        assert np.array_equal(y[short], ref[short])
        assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_context_autotune_applies_to_automatic_plans(ctx):
    c2 = ua.Context(0)
    c2.set_autotune(2)
    A = ua.gen_stencil(ua.STENCIL_2D9, 64, 48, shift=1e-3)
    D = ua.DeviceCsr(c2, A)
    assert D.format in (ua.FMT_ROWBLOCK, ua.FMT_SELL64, ua.FMT_DIA, ua.FMT_VI)
    x = ua.gen_uniform(2, A.shape[0])
    xd, yd = ua.DeviceArray.from_numpy(c2, x), ua.DeviceArray(c2, A.shape[0])
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), orc.spmv(A, x))
    del D, xd, yd
    c2.close()


def test_spmv_matrix_market_reference_fixture(ctx, golden):
    """gre_343_343_crg.mm (general, non-symmetric): every plan and both precisions give
    the oracle's bits."""
    from utopia_amd import petsc_io
    A = petsc_io.read_matrix_market(golden / "utopia_data" / "gre_343_343_crg.mm")
    x = ua.gen_uniform(31, A.shape[1])
    for fmt in (ua.FMT_AUTO, ua.FMT_ROWBLOCK, ua.FMT_SELL64):
        D = ua.DeviceCsr(ctx, A, fmt=fmt)
        xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray(ctx, A.shape[0])
        D.spmv(xd, yd)
        assert np.array_equal(yd.numpy(), orc.spmv(A, x)), fmt
    D = ua.DeviceCsr(ctx, A, dtype=np.float32)
    x32 = x.astype(np.float32)
    xd, yd = ua.DeviceArray.from_numpy(ctx, x32), ua.DeviceArray(ctx, A.shape[0], np.float32)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), orc.spmv_f32(A, x32))


def test_spmv_dot2(ctx):
    A = ua.gen_stencil(ua.STENCIL_2D9, 40, 37, shift=1e-3)
    n = A.shape[0]
    x, z = ua.gen_uniform(1, n), ua.gen_uniform(2, n)
    D = ua.DeviceCsr(ctx, A)
    xd, zd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray.from_numpy(ctx, z), ua.DeviceArray(ctx, n)
    s = ctx.scalars
    ua.check(L().uhip_spmv_dot2_f64(ctx.h, D.ref, P(xd), P(yd), P(zd), ctypes.c_void_p(s)))
    Ax = orc.spmv(A, x)
    assert np.array_equal(yd.numpy(), Ax)
    d = ctx.read_scalars(s, 2)
    assert abs(d[0] - x @ Ax) <= 1e-13 * (np.abs(x) @ np.abs(Ax))
    assert abs(d[1] - z @ x) <= 1e-13 * (np.abs(z) @ np.abs(x))


def _agg_interp(n, seed):
    """A piecewise-constant aggregation transfer (one entry per row, values 1 or 0)."""
    rng = np.random.default_rng(seed)
    parent = np.sort(rng.integers(0, n // 3, n)).astype(np.int32)
    val = np.where(rng.random(n) < 0.05, 0.0, 1.0)
    return ua.Csr(np.arange(n + 1, dtype=np.int32), parent, val, (n, n // 3))


VI_MATRICES = {
    "P3d": lambda: ua.gen_prolongation(16, 12, 10),
    "R3d": lambda: ua.Csr.from_scipy(ua.gen_prolongation(16, 12, 10).to_scipy().T),
    "P2d": lambda: ua.gen_prolongation(34, 20),
    "R2d": lambda: ua.Csr.from_scipy(ua.gen_prolongation(34, 20).to_scipy().T),
    "agg": lambda: _agg_interp(3001, 4),
}


@pytest.mark.parametrize("name", sorted(VI_MATRICES))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_spmv_vi_bit_exact(ctx, name, dt):
    """VI plans (value table + 8-bit codes + row-relative offsets) give the CSR sum's bits
    for the plain product and the transfer epilogues (y += A x, y = mask .* A x)."""
    A = VI_MATRICES[name]()
    m, n = A.shape
    D = ua.DeviceCsr(ctx, A, dtype=dt, fmt=ua.FMT_VI)
    assert D.format == ua.FMT_VI and 1 <= D.m.vi_nvals <= 256 and 1 <= D.m.ndiag <= 32
    assert 1 <= D.m.vi_nsig <= 256  # interior, face, edge and corner rows: signature ids per row
    x = ua.gen_uniform(17, n).astype(dt)
    ref = orc.spmv_f32(A, x) if dt is np.float32 else orc.spmv(A, x)
    xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray(ctx, m, dt)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), ref)
    if dt is np.float64:
        z = ua.gen_uniform(18, m)
        yd.upload(z)
        ua.check(L().uhip_spmv_add(ctx.h, D.ref, P(xd), P(yd)))
        assert np.array_equal(yd.numpy(), z + ref)
        mk = (ua.gen_uniform(19, m) > 0).astype(np.float64)
        mkd = ua.DeviceArray.from_numpy(ctx, mk)
        ua.check(L().uhip_spmv_masked(ctx.h, D.ref, P(xd), P(mkd), P(yd)))
        assert np.array_equal(yd.numpy(), mk * ref)


def _box27(nx, ny, nz):
    """27-point box stencil: -1 off the diagonal, 27 on it."""
    T = lambda k: sp.diags([np.ones(k - 1), np.ones(k), np.ones(k - 1)], [-1, 0, 1])
    B = sp.kron(sp.kron(T(nz), T(ny)), T(nx)).tocsr()
    return ua.Csr.from_scipy((28 * sp.identity(B.shape[0]) - B).tocsr())


VI_SQUARE = {
    "2d9": (lambda: ua.gen_stencil(ua.STENCIL_2D9, 30, 20), 9),
    "3d7": (lambda: ua.gen_stencil(ua.STENCIL_3D7, 10, 9, 8), 13),
    "box27": (lambda: _box27(9, 8, 7), 27),
}


@pytest.mark.parametrize("name", sorted(VI_SQUARE))
def test_vi_square_epilogues(ctx, name):
    """Constant-coefficient stencils (two values) forced into VI: rows of 9-16 and 17-32
    slots from their first column take the two- and four-word code paths. y = A x, the
    residual and A p with p.Ap match the CSR sums bit for bit (the dot to rounding)."""
    make, slots = VI_SQUARE[name]
    A = make()
    n = A.shape[0]
    D = ua.DeviceCsr(ctx, A, fmt=ua.FMT_VI)
    assert D.format == ua.FMT_VI and D.m.ndiag == slots and D.m.vi_nvals == 2 and D.m.vi_nsig > 0
    x, b = ua.gen_uniform(21, n), ua.gen_uniform(22, n)
    Ax = orc.spmv(A, x)
    xd, bd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray.from_numpy(ctx, b), ua.DeviceArray(ctx, n)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), Ax)
    D.residual(xd, bd, yd)
    assert np.array_equal(yd.numpy(), b - Ax)
    s = ctx.scalars
    ua.check(L().uhip_spmv_dot(ctx.h, D.ref, P(xd), P(yd), ctypes.c_void_p(s)))
    assert np.array_equal(yd.numpy(), Ax)
    assert abs(ctx.read_scalars(s, 1)[0] - x @ Ax) <= 1e-13 * (np.abs(x) @ np.abs(Ax))


@pytest.mark.parametrize("slots", [8, 16, 32])
def test_vi_many_signatures_keep_row_codes(ctx, slots):
    """Rows with random slot subsets and values (thousands of distinct signatures) keep their
    per-row masks and codes (one, two and four code words); the product stays bit-exact."""
    rng = np.random.default_rng(slots)
    n = 6000
    rows, cols, vals = [], [], []
    for i in range(n):
        k = np.flatnonzero(rng.random(slots) < 0.5)
        k = np.union1d(k, [0])
        rows += [i] * len(k)
        cols += list(i + k)
        vals += list(rng.integers(1, 9, len(k)) * 0.25)
    A = ua.Csr.from_scipy(sp.csr_matrix((vals, (rows, cols)), shape=(n, n + slots)))
    D = ua.DeviceCsr(ctx, A, fmt=ua.FMT_VI)
    assert D.format == ua.FMT_VI and D.m.vi_nsig == 0 and D.m.ndiag <= slots
    x = ua.gen_uniform(23, n + slots)
    xd, yd = ua.DeviceArray.from_numpy(ctx, x), ua.DeviceArray(ctx, n)
    D.spmv(xd, yd)
    assert np.array_equal(yd.numpy(), orc.spmv(A, x))


def test_vi_rejects_many_values_and_nan(ctx):
    A = ua.gen_prolongation(40, 30)
    assert A.nnz > 300
    val = A.val.copy()
    val[:300] = np.arange(300) * 0.001 + 2.0  # 300 + distinct values
    with pytest.raises(ua.HipError):
        ua.DeviceCsr(ctx, ua.Csr(A.rowptr, A.col, val, A.shape), fmt=ua.FMT_VI)
    val = A.val.copy()
    val[5] = np.nan
    with pytest.raises(ua.HipError):
        ua.DeviceCsr(ctx, ua.Csr(A.rowptr, A.col, val, A.shape), fmt=ua.FMT_VI)
    assert ua.DeviceCsr(ctx, ua.Csr(A.rowptr, A.col, val, A.shape)).format != ua.FMT_VI


@pytest.mark.parametrize("name", ["prolongation", "ragged_empty_rows", "one_row", "varcoef3d", "rows_up_to_900"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_transpose_bit_exact(ctx, name, dtype):
    """y = A^T x (MatMultTranspose, PetscMatrix::transpose_multiply) through the transpose plan:
    equal bit for bit to the oracle's product with the explicit transpose (each y_c adds the
    column's entries by increasing row, as the scatter over A's rows does), on rectangular,
    ragged and single-row matrices; the plan is built once and reused."""
    A = MATRICES[name]()
    x = ua.gen_uniform(15, A.shape[0]).astype(dtype)
    D = ua.DeviceCsr(ctx, A, dtype)
    xd = ua.DeviceArray.from_numpy(ctx, x)
    yd = ua.DeviceArray(ctx, A.shape[1], dtype)
    L = ua.lib()
    for _ in range(2):
        ua.check(L.uhip_spmv_transpose(ctx.h, D.ref, ctypes.c_void_p(xd.ptr), ctypes.c_void_p(yd.ptr)),
                 "uhip_spmv_transpose")
    assert D.m.tr  # the plan stays with A
    At = A.to_scipy().T.tocsr()
    At.sort_indices()
    T = ua.Csr.from_scipy(At)
    ref = orc.spmv(T, x.astype(np.float64)) if dtype == np.float64 else orc.spmv_f32(T, x)
    assert np.array_equal(yd.numpy(), ref)
