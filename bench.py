"""Benchmark: time-to-rtol 1e-10 of MG-preconditioned CG on the 3D 7-point
variable-coefficient operator, 256^3 (16.8M rows, 117M nonzeros), FP32
V-cycle inside FP64 CG ("mixed") versus the same solver in FP64, on one
MI355X; plus the HBM bandwidth of the CSR SpMV kernel.

One step = one full solve from x0 = 0 to |b - Ax| <= 1e-10 |b| with the
matrix, hierarchy and right-hand side resident in HBM (setup excluded and
reported separately). value = mean mixed-precision seconds per solve.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, val_bytes):
    # CSR, int32 indices: values + cols per nonzero; rowptr + x (read once) + y per row
    return (val_bytes + 4) * nnz + (4 + 2 * val_bytes) * n + 4


def dia_bytes(n, ndiag, val_bytes, sym=0, nan_select=False, delta=False):
    # DIA: ndiag value slots per row (absent entries NaN - or zero - in the value arrays;
    # symmetric storage keeps the ndiag // 2 + 1 offsets >= 0, near-symmetric storage adds the
    # lower values' bit-pattern deltas, 4 B in FP64 and 1 B in FP32), the per-row mask (1/2/4 B
    # for <= 8/16/32 offsets; not read where the kernel selects on NaN-encoded absent entries:
    # the 7-point and full layouts), x read once, y written; offsets table
    mb = 1 if ndiag <= 8 else (2 if ndiag <= 16 else 4)
    stored = ndiag // 2 + 1 if sym else ndiag
    mask = 0 if (nan_select and sym != 1) else mb * n
    dl = (ndiag // 2) * (4 if val_bytes == 8 else 1) * n if delta else 0
    return val_bytes * stored * n + mask + dl + 2 * val_bytes * n + 4 * ndiag


def vi_bytes(D, val_bytes, add=False):
    # value-indexed transfer: per row the base column (4 B) and either a signature id (1 B; the
    # signature table is read per workgroup from cache) or its slot mask (1/2/4 B) and 8-bit codes
    # (8 B per 8 slots), the value table, x read once (ncols), y written (and read by y += A x)
    m = D.m
    nw = (m.ndiag + 7) // 8
    mb = 1 if m.ndiag <= 8 else (2 if m.ndiag <= 16 else 4)
    rows, cols = m.nrows, m.ncols
    per_row = 4 + 1 if m.vi_nsig > 0 else 4 + mb + 8 * nw
    tables = val_bytes * m.vi_nvals + (8 * (nw + 1) * m.vi_nsig if m.vi_nsig > 0 else 0)
    return per_row * rows + tables + val_bytes * cols + val_bytes * rows * (2 if add else 1)


def vi_sig_note(D):
    return f", {D.m.vi_nsig} row signatures" if D.m.vi_nsig > 0 else ""


def time_kernel(ua, ctx, launch, reps):
    L = ua.lib()
    for _ in range(3):
        launch()
    ev = ctypes.c_void_p()
    ua.check(L.uhip_event_timer_start(ctx.h, ctypes.byref(ev)))
    for _ in range(reps):
        launch()
    ms = ctypes.c_float()
    ua.check(L.uhip_event_timer_stop(ctx.h, ev, ctypes.byref(ms)))
    return ms.value / reps


def kernel_rooflines(ua, ctx, A, n_rows, reps=20):
    """Average launch time (HIP events on the context stream) of the CSR SpMV
    kernels on the fine-level operator, and a streaming probe (w = a x + b y)."""
    L = ua.lib()
    out = {}
    x64 = ua.DeviceArray.from_numpy(ctx, ua.gen_uniform(5, n_rows))
    y64 = ua.DeviceArray(ctx, n_rows)
    for dt, name, vb in ((np.float64, "spmv_csr_f64", 8), (np.float32, "spmv_csr_f32", 4)):
        D = ua.DeviceCsr(ctx, A, dt)
        if dt is np.float64:
            x, y = x64, y64
        else:
            x = ua.DeviceArray.from_numpy(ctx, ua.gen_uniform(5, n_rows).astype(np.float32))
            y = ua.DeviceArray(ctx, n_rows, np.float32)
        ms = time_kernel(ua, ctx, lambda: D.spmv(x, y), reps)
        csr_b = spmv_bytes(n_rows, A.nnz, vb)
        ty = "double" if dt is np.float64 else "float"
        if D.format == ua.FMT_DIA:
            nbytes = dia_bytes(n_rows, D.m.ndiag, vb, D.m.dia_sym, D.m.dia_nan == 1)
            stored = f"{D.m.ndiag // 2 + 1} of {D.m.ndiag} offsets stored (symmetric)" if D.m.dia_sym else \
                f"{D.m.ndiag} offsets"
            pb = D.m.dia_sym == 2
            kname = "k_spmv_dia7r" if pb else "k_spmv_dia"
            walk = ", plane-walking" if pb else ""
            sel = "NaN-encoded absent entries" if (pb and D.m.dia_nan) else "per-row mask"
            label = f"{kname}<{ty}, EpiStore<{ty}>> (DIA, {stored}{walk}, {sel})"
        elif D.format == ua.FMT_SELL64:
            nbytes, label = csr_b, f"k_spmv_sell<{ty}, EpiStore<{ty}>> (SELL-64)"
        else:
            nbytes, label = csr_b, f"k_spmv_rowblock<{ty}, EpiStore<{ty}>> (row-block CSR)"
        out[name] = {"ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9, "label": label,
                     "csr_bytes": csr_b, "csr_equivalent_gbs": csr_b / (ms * 1e-3) / 1e9}
        del D
    # streaming probe: w = 2x - 3y on 3 x 1.6 GB vectors
    m = 200_000_000
    a = ua.DeviceArray(ctx, m)
    b = ua.DeviceArray(ctx, m)
    w = ua.DeviceArray(ctx, m)
    ua.check(L.uhip_vec_set_f64(ctx.h, m, 1.0, ctypes.c_void_p(a.ptr)))
    ua.check(L.uhip_vec_set_f64(ctx.h, m, 2.0, ctypes.c_void_p(b.ptr)))
    ms = time_kernel(ua, ctx, lambda: L.uhip_vec_waxpby_f64(ctx.h, m, 2.0, ctypes.c_void_p(a.ptr), -3.0,
                                                           ctypes.c_void_p(b.ptr), ctypes.c_void_p(w.ptr)), reps)
    out["stream_waxpby_f64"] = {"ms": ms, "bytes": 24 * m, "gbs": 24 * m / (ms * 1e-3) / 1e9,
                                "label": "k_waxpby<double> (w = a x + b y, 200M entries)"}
    for v in (a, b, w):
        v.free()
    return out


def generic_csr_kernels(ua, ctx, n, seed, reps=20):
    """y = A x with the general sparse formats on a matrix with no diagonal structure: the n^3
    variable-coefficient operator renumbered by a seeded permutation of its 64-row blocks
    (generator kind 5; at 256^3 the FP64 CSR arrays are 1.47 GB). Row-block CSR and SELL-64 in
    FP64, and the automatic plan in FP32; algorithmic bytes are those of the CSR arrays."""
    A = ua.gen_stencil(ua.STENCIL_3D7_BLOCKPERM, n, n, n, seed=seed)
    N = A.shape[0]
    out = {}
    names = {ua.FMT_ROWBLOCK: "k_spmv_rowblock", ua.FMT_SELL64: "k_spmv_sell", ua.FMT_DIA: "k_spmv_dia",
             ua.FMT_VI: "k_spmv_vi"}
    for dt, fmt, key in ((np.float64, ua.FMT_ROWBLOCK, "perm_rowblock_f64"), (np.float64, ua.FMT_SELL64, "perm_sell64_f64"),
                         (np.float32, ua.FMT_AUTO, "perm_auto_f32")):
        vb = 8 if dt is np.float64 else 4
        D = ua.DeviceCsr(ctx, A, dt, fmt=fmt)
        x = ua.DeviceArray.from_numpy(ctx, ua.gen_uniform(5, N).astype(dt))
        y = ua.DeviceArray(ctx, N, dt)
        ms = time_kernel(ua, ctx, lambda: D.spmv(x, y), reps)
        nb = spmv_bytes(N, A.nnz, vb)
        ty = "double" if dt is np.float64 else "float"
        out[key] = {"ms": ms, "bytes": nb, "gbs": nb / (ms * 1e-3) / 1e9,
                    "label": f"{names.get(D.format, D.format)}<{ty}, EpiStore<{ty}>> on the block-permuted "
                             f"{n}^3 operator ({A.nnz} nonzeros, format {D.format})"}
        del D
        x.free()
        y.free()
    return out


class ChebStep(ctypes.Structure):
    _fields_ = [("pprev", ctypes.c_void_p), ("p_dtype", ctypes.c_int32), ("rprev", ctypes.c_void_p),
                ("xin", ctypes.c_void_p), ("xout", ctypes.c_void_p), ("x_dtype", ctypes.c_int32),
                ("rout", ctypes.c_void_p), ("pout", ctypes.c_void_p), ("a_prev", ctypes.c_double),
                ("beta", ctypes.c_double), ("a_cur", ctypes.c_double), ("last", ctypes.c_int32),
                ("dot_with", ctypes.c_void_p), ("dot_d", ctypes.c_void_p)]


def solve_kernels(ua, ctx, A, interps, n_rows, reps=20):
    """Average launch time and algorithmic bytes of the fine-level launches of one CG iteration
    of the 256^3 solve (V(2,2) Chebyshev from a zero guess), in FP64 (the FP64 form) and FP32 (the
    mixed form): the fused smoothing pairs the solve runs (the pre-smoothing step with r_R = r - A c;
    p0 = r - A c with the last smoothing step and z.r), the single-product launches they replace,
    the restriction and the interpolation-add, and the CG's A p with p.Ap (FP64 in both forms)."""
    L = ua.lib()
    out = {}
    N = n_rows
    F64, F32 = ua.F64, ua.F32
    vec = {}
    for dt in (np.float64, np.float32):
        vec[dt] = [ua.DeviceArray.from_numpy(ctx, ua.gen_uniform(20 + i, N).astype(dt)) for i in range(4)]
    r64 = ua.DeviceArray.from_numpy(ctx, ua.gen_uniform(30, N))
    dd = ua.DeviceArray(ctx, 2)
    P = ua.DeviceCsr(ctx, interps[-1], np.float32)   # 128^3 -> 256^3 interpolation (auto plan)
    R = ua.DeviceCsr(ctx, ua.Csr.from_scipy(interps[-1].to_scipy().T.tocsr()), np.float32)
    nc = interps[-1].shape[1]
    xc = ua.DeviceArray.from_numpy(ctx, ua.gen_uniform(31, nc).astype(np.float32))
    mc = ua.DeviceArray.from_numpy(ctx, np.ones(nc, np.float32))
    for dt, vb, tag, code in ((np.float64, 8, "f64", F64), (np.float32, 4, "f32", F32)):
        D = ua.DeviceCsr(ctx, A, dt)
        base = dia_bytes(N, D.m.ndiag, vb, D.m.dia_sym, D.m.dia_nan == 1) if D.format == ua.FMT_DIA else \
            spmv_bytes(N, A.nnz, vb)
        v0, v1, v2, v3 = (ctypes.c_void_p(a.ptr) for a in vec[dt])
        ty = "double" if dt is np.float64 else "float"

        def cheb(pprev, xin, xout, dot):
            st = ChebStep(pprev=pprev, p_dtype=code, rprev=None, xin=xin, xout=xout, x_dtype=code, rout=None,
                          pout=None, a_prev=0.3, beta=0.1, a_cur=0.2, last=1,
                          dot_with=ctypes.c_void_p(r64.ptr) if dot else None,
                          dot_d=ctypes.c_void_p(dd.ptr) if dot else None)
            return lambda: L.uhip_cheb_step_apply(ctx.h, D.ref, ctypes.byref(st))

        forms = [
            ("pre_smooth", f"EpiCheb<{ty}> last step, zero guess", cheb(v0, None, v1, False), base),
            ("residual", f"EpiResidual<{ty}> r_R = r - A c",
             lambda: L.uhip_spmv_residual(ctx.h, D.ref, v1, v0, code, v2), base + vb * N),
            ("post_smooth_dot", f"EpiCheb<{ty}, dot> last step with z.r (r FP64)", cheb(v2, v1, v3, True),
             base + vb * N + 8 * N),
        ]
        if dt is np.float64:
            forms.append(("cg_Ap_dot", "EpiStoreDot<double> A p with p.Ap",
                          lambda: L.uhip_spmv_dot(ctx.h, D.ref, v0, v1, ctypes.c_void_p(dd.ptr)), base))
        if L.uhip_cheb_pair_supported(D.ref):
            # the fused smoothing pairs the V-cycle runs on this level: values once, b in, c and
            # r_R out; x, b in, z out and r (FP64) for z.r
            forms.append(("pre_pair", f"k_dia7_pair<{ty}> last step from a zero guess, then r_R = b - A c",
                          lambda: L.uhip_cheb_pre_residual(ctx.h, D.ref, v0, 0.3, 0.1, 0.2, v1, v2),
                          base + vb * N))
            forms.append(("post_pair", f"k_dia7_pair<{ty}, dot> p0 = b - A c, then the last step with z.r",
                          lambda: L.uhip_cheb_residual_post(ctx.h, D.ref, v0, v1, 0.3, 0.1, 0.2, v3,
                                                            ctypes.c_void_p(r64.ptr), ctypes.c_void_p(dd.ptr)),
                          base + vb * N + 8 * N))
        for key, label, launch, nbytes in forms:
            def run(launch=launch):
                ua.check(launch())
            ms = time_kernel(ua, ctx, run, reps)
            kname = "" if "pair" in key else f" ({'k_spmv_dia7r' if D.m.dia_sym == 2 else 'format ' + str(D.format)})"
            out[f"{key}_{tag}"] = {"ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9,
                                   "label": f"{label}{kname}"}
        del D
    xs = vec[np.float32]
    ms = time_kernel(ua, ctx, lambda: ua.check(L.uhip_spmv_masked(ctx.h, R.ref, ctypes.c_void_p(xs[0].ptr),
                                                                  ctypes.c_void_p(mc.ptr), ctypes.c_void_p(xc.ptr))),
                     reps)
    nb = vi_bytes(R, 4) + 4 * nc
    out["restrict_f32"] = {"ms": ms, "bytes": nb, "gbs": nb / (ms * 1e-3) / 1e9,
                           "label": f"EpiMaskStore<float> R r_R, format {R.format} (VI = 3), {R.m.ndiag} slots"
                                    f"{vi_sig_note(R)}"}
    ms = time_kernel(ua, ctx, lambda: ua.check(L.uhip_spmv_add(ctx.h, P.ref, ctypes.c_void_p(xc.ptr),
                                                               ctypes.c_void_p(xs[1].ptr))), reps)
    nb = vi_bytes(P, 4, add=True)
    out["interpolate_f32"] = {"ms": ms, "bytes": nb, "gbs": nb / (ms * 1e-3) / 1e9,
                              "label": f"EpiAddTo<float> c += P c_c, format {P.format} (VI = 3), {P.m.ndiag} slots"
                                       f"{vi_sig_note(P)}"}
    del P, R
    # level 1: the 128^3 Galerkin operator P^T A P (27 points), in the plan the solve gives it
    DA = ua.DeviceCsr(ctx, A, np.float64, fmt=ua.FMT_ROWBLOCK)
    DP = ua.DeviceCsr(ctx, interps[-1], np.float64, fmt=ua.FMT_ROWBLOCK)
    raw = ua.uhip_csr()
    ua.check(L.uhip_csr_ptap(ctx.h, DA.ref, DP.ref, ctypes.byref(raw)))
    C = ua.csr_from_raw(ctx, raw)
    ua.check(L.uhip_csr_free(ctx.h, ctypes.byref(raw)))
    del DA, DP
    for dt, vb, tag, code in ((np.float64, 8, "f64", F64), (np.float32, 4, "f32", F32)):
        D = ua.DeviceCsr(ctx, C, dt)
        base = dia_bytes(nc, D.m.ndiag, vb, D.m.dia_sym, D.m.dia_nan == 1, bool(D.m.dia_delta)) \
            if D.format == ua.FMT_DIA else spmv_bytes(nc, C.nnz, vb)
        u = [ua.DeviceArray.from_numpy(ctx, ua.gen_uniform(40 + i, nc).astype(dt)) for i in range(3)]
        u0, u1, u2 = (ctypes.c_void_p(a.ptr) for a in u)
        st = ChebStep(pprev=u0, p_dtype=code, rprev=None, xin=u1, xout=u2, x_dtype=code, rout=None, pout=None,
                      a_prev=0.3, beta=0.1, a_cur=0.2, last=1, dot_with=None, dot_d=None)
        layout = ("near-symmetric, bit-pattern deltas" if D.m.dia_delta else
                  ("symmetric" if D.m.dia_sym else "full")) if D.format == ua.FMT_DIA else f"format {D.format}"
        kname = "k_spmv_dia27r" if D.m.dia_w27 else f"format {D.format}"
        for key, label, launch, nbytes in (
                ("l1_residual", f"EpiResidual<{'double' if vb == 8 else 'float'}> r_R = r - A c, 128^3 Galerkin",
                 lambda: L.uhip_spmv_residual(ctx.h, D.ref, u1, u0, code, u2), base + vb * nc),
                ("l1_cheb", f"EpiCheb<{'double' if vb == 8 else 'float'}> last step, 128^3 Galerkin",
                 lambda: L.uhip_cheb_step_apply(ctx.h, D.ref, ctypes.byref(st)), base + vb * nc)):
            ms = time_kernel(ua, ctx, lambda launch=launch: ua.check(launch()), reps)
            out[f"{key}_{tag}"] = {"ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9,
                                   "label": f"{label} ({kname}, {layout}, {D.m.ndiag} offsets)"}
        del D
    return out


def cpu_baseline(n, levels, seed, gpu_iters, x_fp64, x_mixed, reps=1, iters_1core=4):
    """The oracle (plain C, OpenMP, FP64) MG-PCG on the benchmark's own problem: the same 256^3
    operator, hierarchy, right-hand side and tolerance. Setup (Galerkin products, power method,
    dense Cholesky) is timed apart from the solve; the solve runs to convergence `reps` times (the
    median is the value); its iteration count and solution are compared with the GPU's. The one-core
    figure times `iters_1core` iterations of the same solve on one thread and scales them to the
    iteration count."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    import utopia_amd as ua

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = min(threads, 16)
    orc.set_threads(threads)
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=seed)
    b = ua.gen_uniform(seed + 100, A.shape[0])
    atol = 1e-10 * float(np.linalg.norm(b))
    interps = ua.mg_hierarchy(n, n, n, levels)
    ua.srand(1)
    t0 = time.perf_counter()
    mg = orc.MG(interps, pre=2, post=2)
    mg.update(A)
    t_setup = time.perf_counter() - t0
    log(f"cpu baseline: oracle setup {t_setup:.1f} s ({threads} threads)")
    times, x, st = [], None, None
    for r in range(reps):
        t0 = time.perf_counter()
        x, st = mg.cg_solve(b, atol=atol, max_it=1000)
        times.append(time.perf_counter() - t0)
        log(f"cpu baseline: oracle solve {r}: {times[-1]:.1f} s, {st['iterates'] + 1} iterations")
    its = st["iterates"] + 1
    t_solve = float(np.median(times))
    orc.set_threads(1)
    t0 = time.perf_counter()
    _, st1 = mg.cg_solve(b, atol=0.0, max_it=iters_1core)
    its1 = st1["iterates"] + 1
    per_it1 = (time.perf_counter() - t0) / its1
    orc.set_threads(threads)
    nx = float(np.linalg.norm(x))
    return {
        "value": t_solve, "unit": "s", "cores": threads, "kind": "port",
        "value_1core": per_it1 * its,
        "setup_s": t_setup, "solve_s": times, "iterations": its,
        "gpu_iterations": gpu_iters,
        "rel_error_fp64_vs_oracle": float(np.linalg.norm(x_fp64 - x) / nx),
        "rel_error_mixed_vs_oracle": float(np.linalg.norm(x_mixed - x) / nx),
        "sample": (f"oracle FP64 MG-PCG (V(2,2) Chebyshev, {levels} levels, dense Cholesky coarse solve) on the "
                   f"same {n}^3 variable-coefficient problem to rtol 1e-10: {its} iterations, solve "
                   f"{t_solve:.1f} s (median of {reps}), setup {t_setup:.1f} s apart; value_1core: {its1} "
                   f"iterations of the same solve on one thread ({per_it1:.2f} s each) times {its}"),
    }


def bench_batched(a):
    """64 SPD 2D 5-point systems (sides log-uniform in [64, 512], per-edge
    coefficients exp(N(0, 0.5))), 8 right-hand sides each, batched MG-PCG with
    an FP32 V-cycle in FP64 CG to rtol 1e-8, 16 CG iterations of the whole
    batch per hipGraph replay; plus the adjoint solve and pattern gradient."""
    import utopia_amd as ua
    from utopia_amd import batched as bt
    L = ua.lib()
    ctx = ua.Context(0)
    k = 8
    t0 = time.perf_counter()
    sides = bt.batch_sides(a.seed, 64)
    p = bt.make_batch(sides, seed=a.seed * 1000)
    B = ua.gen_uniform(a.seed + 7, p.n * k)
    Xs = ua.gen_normal(a.seed + 8, p.n * k)
    log(f"batched: 64 systems, sides {min(sides)}..{max(sides)}, rows {p.n}, nnz {p.A.nnz}, levels {p.levels}, "
        f"generated in {time.perf_counter() - t0:.1f} s")
    bs = bt.BatchedSolver(ctx, k, mixed=True, rtol=1e-8, max_it=500, pre_smoothing_steps=2, post_smoothing_steps=2,
                          autotune=a.autotune)
    t0 = time.perf_counter()
    bs.setup(p)
    setup = time.perf_counter() - t0
    Bd = ua.DeviceArray.from_numpy(ctx, B)
    Xd = ua.DeviceArray(ctx, p.n * k)
    Gd = ua.DeviceArray(ctx, p.n * k)
    Ld = ua.DeviceArray(ctx, p.n * k)
    grad = ua.DeviceArray(ctx, p.A.nnz)
    Xsd = ua.DeviceArray.from_numpy(ctx, Xs)

    fwd = []

    def step():
        ua.check(L.uhip_vec_set_f64(ctx.h, p.n * k, 0.0, ctypes.c_void_p(Xd.ptr)))
        its = bs.solve_device(Bd, Xd)
        fwd.append(bs.info("solve_seconds"))
        # adjoint: G = X - X*, lambda = A^-1 G, dL/dA on the pattern
        ua.check(L.uhip_vec_waxpby_f64(ctx.h, p.n * k, 1.0, ctypes.c_void_p(Xd.ptr), -1.0, ctypes.c_void_p(Xsd.ptr),
                                       ctypes.c_void_p(Gd.ptr)))
        ua.check(L.uhip_vec_set_f64(ctx.h, p.n * k, 0.0, ctypes.c_void_p(Ld.ptr)))
        its_adj = bs.solve_device(Gd, Ld)
        bs.pattern_grad_device(Ld, Xd, grad)
        return its, its_adj

    for _ in range(a.warmup):
        step()
    ctx.sync()
    times = []
    fwd.clear()
    for i in range(a.steps):
        t = time.perf_counter()
        its, its_adj = step()
        ctx.sync()
        times.append(time.perf_counter() - t)
        log(f"batched: step {i} {times[-1] * 1e3:.1f} ms (forward solve {fwd[-1] * 1e3:.1f} ms), "
            f"iterations {its} / {its_adj}")
    conv, flags = bs.convergence()
    X = Xd.numpy().reshape(-1, k)
    R = B.reshape(-1, k) - p.A.to_scipy() @ X
    worst = max(float(np.max(np.linalg.norm(R[p.seg[s]:p.seg[s + 1]], axis=0)
                             / np.linalg.norm(B.reshape(-1, k)[p.seg[s]:p.seg[s + 1]], axis=0)))
                for s in range(p.nseg))
    per = float(np.mean(times))
    line = {
        "metric": "batched MG-PCG forward + adjoint, 64 systems x 8 RHS, rtol 1e-8",
        "value": per, "unit": "s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": per * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 CG", "data": "synthetic (seeded sides, edge coefficients, RHS, x*)",
        "config": {"workload": "64 SPD 2D 5-point systems, sides log-uniform [64,512], 8 RHS, forward + adjoint",
                   "rows": p.n, "nnz": p.A.nnz, "levels": p.levels, "rhs": k,
                   "graph": f"{int(bs.info('graph_iterations'))} CG iterations of the batch per hipGraph replay "
                            f"(kernels gated on the active-pair count)", "parallelism": "single GPU"},
        "graph_replays": int(bs.info("graph_replays")), "graph_captures": int(bs.info("graph_captures")),
        "forward_s": float(np.mean(fwd)), "setup_s": setup,
        "iterations_max": int(conv.max()), "iterations_min": int(conv.min()),
        "converged": int((flags == 1).sum()), "pairs": int(flags.size), "worst_rel_residual": worst,
    }
    print(json.dumps(line))
    return 0


def bench_mprgp(a, n=1024):
    """min 1/2 x'Ax - b'x, l <= x, A = 9-point Laplacian + 1e-3 I on n^2, l = -0.05 - 0.15 u,
    MPRGP to projected-gradient norm 1e-9; active set compared with the oracle's."""
    import utopia_amd as ua
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    ctx = ua.Context(0)
    seed = a.seed
    A = ua.gen_stencil(ua.STENCIL_2D9, n, n, shift=1e-3)
    b = ua.gen_uniform(seed, n * n)
    lb = -0.05 - 0.15 * ua.gen_uniform(seed + 1, n * n, 0, 1)
    q = ua.MPRGP(ctx, atol=1e-9, rtol=0, max_it=1000000)
    q.update(A)
    q.set_box_constraints(lb, None)
    bd = ua.DeviceArray.from_numpy(ctx, b)
    xd = ua.DeviceArray(ctx, n * n)
    L = ua.lib()
    times = []
    for i in range(a.warmup + a.steps):
        ua.check(L.uhip_vec_set_f64(ctx.h, n * n, 0.0, ctypes.c_void_p(xd.ptr)))
        ctx.sync()
        t = time.perf_counter()
        q.apply_device(bd, xd)
        ctx.sync()
        if i >= a.warmup:
            times.append(time.perf_counter() - t)
        log(f"mprgp {n}^2: solve {time.perf_counter() - t:.3f} s, iterations {q.status['iterates']}")
    x = xd.numpy()
    st = q.status
    line = {"metric": f"MPRGP box-constrained QP {n}^2 9-point + 1e-3 I, projected gradient 1e-9",
            "value": float(np.mean(times)), "unit": "s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": float(np.mean(times)) * 1e3, "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded b, bounds)",
            "config": {"workload": f"MPRGP {n}^2 9-point + 1e-3 I, l = -0.05 - 0.15 u", "rows": n * n,
                       "nnz": A.nnz, "parallelism": "single GPU"},
            "iterations": st["iterates"], "reason": st["reason"]}
    if not a.no_cpu_baseline:
        orc.set_threads(min(16, os.cpu_count() or 1))
        ua.srand(1)
        t = time.perf_counter()
        xo, so = orc.mprgp_solve(A, b, lb=lb, atol=1e-9, rtol=0, max_it=1000000)
        tc = time.perf_counter() - t
        act, act_o = np.abs(x - lb) < 1e-12, np.abs(xo - lb) < 1e-12
        ties = np.abs(x - lb) < 1e-9
        line["cpu_baseline"] = {"value": tc, "unit": "s", "cores": orc.get_threads(), "kind": "port",
                                "sample": f"oracle MPRGP on the same {n}^2 problem, {so['iterates']} iterations"}
        line["oracle_iterations"] = so["iterates"]
        line["active_set"] = {"gpu": int(act.sum()), "oracle": int(act_o.sum()),
                              "disagree_beyond_ties": int(np.sum((act != act_o) & ~ties))}
        line["rel_error_vs_oracle"] = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
        log(f"oracle: {tc:.1f} s, {so['iterates']} iterations; active {act.sum()} vs {act_o.sum()}")
    print(json.dumps(line))
    return 0


def host_info():
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k = line.split(":")[0].strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k] = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return info


def bench_shard(a):
    """One GPU's share of the 256^3 variable-coefficient problem split over 8 GPUs
    by z-slabs: the 256 x 256 x 32 local block (2.1M rows). Times the SpMV plans and
    the mixed MG-PCG iteration at that shape; the halo exchange a multi-GPU run would
    add is a model (one 256^2 plane each way per SpMV over one xGMI link), not a
    measurement."""
    import utopia_amd as ua
    n, nz, parts = a.n, a.n // 8, 8
    ctx = ua.Context(0)
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, nz, seed=a.seed)
    N = A.shape[0]
    kern = kernel_rooflines(ua, ctx, A, N)
    for k, v in kern.items():
        log(f"shard kernel {k}: {v['ms'] * 1e3:.1f} us, {v['gbs']:.0f} GB/s")
    b = ua.gen_uniform(a.seed + 100, N)
    interps = ua.mg_hierarchy(n, n, nz, 4)
    atol = 1e-10 * float(np.linalg.norm(b))
    bd = ua.DeviceArray.from_numpy(ctx, b)
    xd = ua.DeviceArray(ctx, N)
    L = ua.lib()
    forms = {}
    for form, kind in (("fp64", "multigrid"), ("mixed", "multigrid_mixed")):
        ua.srand(1)
        cg = ua.ConjugateGradient(ctx, kind, atol=atol, rtol=0, stol=0, max_it=1000, pre_smoothing_steps=2,
                                  post_smoothing_steps=2, autotune=a.autotune)
        cg.set_transfer_operators(interps)
        cg.update(A)
        times = []
        for k in range(a.warmup + a.steps):
            ua.check(L.uhip_vec_set_f64(ctx.h, N, 0.0, ctypes.c_void_p(xd.ptr)))
            ctx.sync()
            t0 = time.perf_counter()
            cg.apply_device(bd, xd)
            ctx.sync()
            if k >= a.warmup:
                times.append(time.perf_counter() - t0)
        its = cg.status["iterates"] + 1
        t = float(np.mean(times))
        forms[form] = {"s": t, "iterations": its, "ms_per_iteration": t / its * 1e3}
        log(f"shard {form}: {t * 1e3:.1f} ms, {its} iterations, {t / its * 1e6:.0f} us/iteration")
        del cg
    plane = n * n
    link = 153e9  # one xGMI link, bytes/s
    halo = {"bytes_per_spmv_fp64": 2 * 8 * plane, "modelled_us_fp64": 8 * plane / link * 1e6,
            "bytes_per_spmv_fp32": 2 * 4 * plane, "modelled_us_fp32": 4 * plane / link * 1e6,
            "note": "model: one plane each way to each of 2 neighbours on separate links, overlapped with the "
                    "interior rows; not measured (one GPU)"}
    print(json.dumps({
        "metric": f"shard shape: {n}x{n}x{nz} slab (1 of {parts} z-slabs of the {n}^3 problem), mixed MG-PCG "
                  "time per iteration and SpMV GB/s", "value": forms["mixed"]["ms_per_iteration"],
        "unit": "ms/iteration", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": forms["mixed"]["s"] * 1e3, "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32 V-cycle / f64 CG", "data": "synthetic (seeded coefficients and b)",
        "config": {"workload": f"MG-PCG 3D 7-point variable-coefficient {n}x{n}x{nz} (local block of an "
                               f"{parts}-way z-slab split), 4 levels, rtol 1e-10", "rows": N, "nnz": A.nnz},
        "forms": forms, "kernels": kern, "halo": halo}))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", type=int, default=5,
                    help="reps > 0: every SpMV plan of the hierarchy is chosen by timing (uhip_csr_autotune)")
    ap.add_argument("--cpu-reps", type=int, default=1,
                    help="oracle solves at the full size timed for the CPU baseline (median reported)")
    ap.add_argument("--solve-only", choices=["fp64", "mixed"],
                    help="time the solves of one form and stop (no kernel table, no CPU baseline): for "
                         "kernel-trace runs that attribute a solve's time to its kernels")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the SpMV / streaming kernel timings (for counter collection runs)")
    ap.add_argument("--workload", choices=["mgpcg256", "batched", "mprgp1024", "shard"], default="mgpcg256",
                    help="mgpcg256: the 256^3 MG-PCG line (default); batched: 64 systems x 8 RHS hipGraph MG-PCG; "
                         "mprgp1024: box-constrained QP on the 1024^2 9-point operator; shard: one GPU's "
                         "256x256x32 slab of the 256^3 problem")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the multi-GPU sharding is a design "
              f"(row-block slabs + halo exchange) and is not measured here.", file=sys.stderr)
        return 2

    if a.workload == "batched":
        return bench_batched(a)
    if a.workload == "mprgp1024":
        return bench_mprgp(a)
    if a.workload == "shard":
        return bench_shard(a)
    import utopia_amd as ua
    L = ua.lib()
    n = a.n
    ctx = ua.Context(0)
    t0 = time.perf_counter()
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=a.seed)
    b = ua.gen_uniform(a.seed + 100, A.shape[0])
    interps = ua.mg_hierarchy(n, n, n, a.levels)
    N, nnz = A.shape[0], A.nnz
    log(f"generated {n}^3 variable-coefficient operator: N={N} nnz={nnz} in {time.perf_counter() - t0:.1f} s")
    if a.kernels_only:
        kern = kernel_rooflines(ua, ctx, A, N)
        kern.update(solve_kernels(ua, ctx, A, interps, N))
        kern.update(generic_csr_kernels(ua, ctx, n, a.seed))
        for k, v in kern.items():
            log(f"kernel {k}: {v['ms']:.3f} ms, {v['gbs']:.0f} GB/s, {v['bytes']} algorithmic bytes, {v.get('label', '')}")
        print(json.dumps({"kernels": kern}))
        return 0
    atol = 1e-10 * float(np.linalg.norm(b))
    bd = ua.DeviceArray.from_numpy(ctx, b)
    xd = ua.DeviceArray(ctx, N)

    results = {}
    forms = (("fp64", "multigrid"), ("mixed", "multigrid_mixed"))
    if a.solve_only:
        forms = tuple(f for f in forms if f[0] == a.solve_only)
    for form, kind in forms:
        ua.srand(1)
        cg = ua.ConjugateGradient(ctx, kind, atol=atol, rtol=0, stol=0, max_it=1000, pre_smoothing_steps=2,
                                  post_smoothing_steps=2, autotune=a.autotune)
        cg.set_transfer_operators(interps)
        t0 = time.perf_counter()
        cg.update(A)
        setup = time.perf_counter() - t0
        log(f"{form}: setup {setup:.2f} s, levels {int(cg.info('n_levels'))}, coarse n {int(cg.info('coarse_n'))}")

        def step():
            ua.check(L.uhip_vec_set_f64(ctx.h, N, 0.0, ctypes.c_void_p(xd.ptr)))
            cg.apply_device(bd, xd)

        for _ in range(a.warmup):
            step()
        ctx.sync()
        times = []
        t_all = time.perf_counter()
        for k in range(a.steps):
            t = time.perf_counter()
            step()
            ctx.sync()
            times.append(time.perf_counter() - t)
            log(f"{form}: step {k} {times[-1] * 1e3:.1f} ms, iterations {cg.status['iterates'] + 1}")
        total = time.perf_counter() - t_all
        x = xd.numpy()
        res = float(np.linalg.norm(b - A.to_scipy() @ x))
        results[form] = {"setup_s": setup, "s_per_solve": total / a.steps, "times": times,
                         "iterations": int(cg.info("total_iterations")), "refinements": int(cg.info("refinements")),
                         "rel_residual": res / float(np.linalg.norm(b)), "x": x}
        del cg

    if a.solve_only:
        r = results[a.solve_only]
        r.pop("x")
        print(json.dumps({"solve_only": a.solve_only, **r}))
        return 0
    x64 = results["fp64"].pop("x")
    xmx = results["mixed"].pop("x")
    err_mixed = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))
    log(f"mixed vs fp64: rel. error {err_mixed:.2e}")

    kern = kernel_rooflines(ua, ctx, A, N)
    kern.update(solve_kernels(ua, ctx, A, interps, N))
    kern.update(generic_csr_kernels(ua, ctx, n, a.seed))
    for k, v in kern.items():
        log(f"kernel {k}: {v['ms']:.3f} ms, {v['gbs']:.0f} GB/s ({100 * v['gbs'] / HBM_PEAK_GBS:.1f}% of 8 TB/s) "
            f"{v['label']}")

    cpu = None
    if not a.no_cpu_baseline:
        t0 = time.perf_counter()
        cpu = cpu_baseline(n, a.levels, a.seed, results["mixed"]["iterations"], x64, xmx, reps=a.cpu_reps)
        log(f"cpu baseline ({time.perf_counter() - t0:.1f} s): {cpu['value']:.1f} s, {cpu['sample']}")

    # The roofline line is for the launch with the largest share of a mixed CG iteration (the
    # CG's A p with p.Ap, the fused smoothing pairs of the fine level, the transfers); the other
    # launch forms and the plain products are in "kernels". HBM bytes per launch come from
    # the committed counter runs (FETCH_SIZE / WRITE_SIZE), used only when they were taken on
    # the kernel measured here.
    per_iteration = {"cg_Ap_dot_f64": 1, "pre_pair_f32": 1, "post_pair_f32": 1, "restrict_f32": 1,
                     "interpolate_f32": 1}
    if "pre_pair_f32" not in kern:  # planes that do not split into whole chunks: one launch per product
        per_iteration = {"cg_Ap_dot_f64": 1, "residual_f32": 2, "pre_smooth_f32": 1, "post_smooth_dot_f32": 1,
                         "restrict_f32": 1, "interpolate_f32": 1}
    rk = max(per_iteration, key=lambda k: kern[k]["ms"] * per_iteration[k])
    traffic = {}
    tf = ROOT / "profiles" / "spmv_varcoef256_traffic.json"
    if tf.exists():
        try:
            tj = json.loads(tf.read_text())
            for k, v in tj.get("kernels", {}).items():
                if k in kern and v.get("label") == kern[k]["label"] and tj.get("n") == n:
                    traffic[k] = v.get("hbm_bytes_per_launch")
        except Exception:
            traffic = {}

    mixed, fp64 = results["mixed"], results["fp64"]
    sp = kern[rk]
    line = {
        "metric": "Time-to-rtol 1e-10, MG-PCG 3D 256³: mixed vs FP64 on GPU; SpMV HBM GB/s",
        "value": mixed["s_per_solve"],
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": mixed["s_per_solve"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 CG",
        "data": "synthetic (seeded generator: kappa = exp(N(0,1)) per cell, b ~ U[-1,1])",
        "config": {"workload": f"MG-PCG 3D 7-point variable-coefficient {n}^3, rtol 1e-10, x0 = 0",
                   "rows": N, "nnz": nnz, "levels": a.levels, "smoother": "Chebyshev3level degree 2, V(2,2)",
                   "coarse": "dense FP64 Cholesky (explicit inverse)", "parallelism": "single GPU"},
        "fp64": {"s_per_solve": fp64["s_per_solve"], "iterations": fp64["iterations"], "setup_s": fp64["setup_s"],
                 "rel_residual": fp64["rel_residual"]},
        "mixed": {"s_per_solve": mixed["s_per_solve"], "iterations": mixed["iterations"],
                  "refinements": mixed["refinements"], "setup_s": mixed["setup_s"],
                  "rel_residual": mixed["rel_residual"], "rel_error_vs_fp64": err_mixed},
        "speedup_mixed_vs_fp64": fp64["s_per_solve"] / mixed["s_per_solve"],
        "roofline": {"bound": "hbm", "kernel": f"{rk}: {sp['label']}",
                     "launches_per_cg_iteration": per_iteration[rk],
                     "achieved": sp["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": sp["gbs"] / HBM_PEAK_GBS,
                     "traffic": traffic.get(rk), "bytes_per_launch": sp["bytes"], "avg_launch_ms": sp["ms"]},
        "kernels": {k: {"gbs": v["gbs"], "ms": v["ms"], "frac": v["gbs"] / HBM_PEAK_GBS, "label": v["label"],
                        "bytes": v["bytes"], "traffic": traffic.get(k),
                        "csr_equivalent_gbs": v.get("csr_equivalent_gbs")}
                    for k, v in kern.items()},
        "cpu_baseline": cpu,
        "host": host_info(),
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
