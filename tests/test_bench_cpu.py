"""bench.py's contract on the CPU: the multi-GPU refusal, the algorithmic byte counts
behind the roofline figures, and the CPU-baseline record (on a tiny sample)."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_algorithmic_bytes_of_the_256_cube_operator():
    n = 256 ** 3
    # symmetric 7-point DIA: 4 stored diagonals, 1 B mask per row, x and y, 7 offsets
    assert bench.dia_bytes(n, 7, 4, 2) == 4 * 4 * n + n + 8 * n + 4 * 7 == 419_430_428
    assert bench.dia_bytes(n, 7, 8, 1) == 822_083_612
    assert bench.dia_bytes(n, 7, 8, 0) == 1_224_736_796
    # CSR: 12 B per nonzero (FP64 value + int32 column), rowptr, x and y
    assert bench.spmv_bytes(n, 117_047_296, 8) == 12 * 117_047_296 + 4 * (n + 1) + 16 * n


def test_nan_selected_plane_kernel_reads_no_mask():
    n = 256 ** 3
    assert bench.dia_bytes(n, 7, 8, 2, nan_select=True) == 4 * 8 * n + 16 * n + 4 * 7
    assert bench.dia_bytes(n, 7, 8, 2, nan_select=False) == 4 * 8 * n + n + 16 * n + 4 * 7


def test_cpu_baseline_record_on_a_small_problem():
    """The record on a 16^3 problem: the oracle solve's iterations and its distance to the
    solutions handed in (the oracle's own solution, so the distance is 0)."""
    import numpy as np
    import utopia_amd as ua
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    n, levels, seed = 16, 3, 3
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, n, n, n, seed=seed)
    b = ua.gen_uniform(seed + 100, A.shape[0])
    ua.srand(1)
    mg = orc.MG(ua.mg_hierarchy(n, n, n, levels), pre=2, post=2)
    mg.update(A)
    x, st = mg.cg_solve(b, atol=1e-10 * float(np.linalg.norm(b)), max_it=1000)
    r = bench.cpu_baseline(n, levels, seed, st["iterates"] + 1, x, x, reps=1, iters_1core=2)
    assert set(r) >= {"value", "unit", "cores", "kind", "sample", "value_1core", "iterations"}
    assert r["unit"] == "s" and r["kind"] == "port" and r["cores"] >= 1
    assert r["value"] > 0 and r["value_1core"] > 0 and "16^3" in r["sample"]
    assert r["iterations"] == st["iterates"] + 1 == r["gpu_iterations"]
    assert r["rel_error_fp64_vs_oracle"] == 0.0
