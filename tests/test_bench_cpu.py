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


def test_cpu_baseline_record_on_a_tiny_sample():
    r = bench.cpu_baseline(32, 10, sample_n=16)
    assert set(r) >= {"value", "unit", "cores", "kind", "sample", "value_1core"}
    assert r["unit"] == "s" and r["kind"] == "port" and r["cores"] >= 1
    assert r["value"] > 0 and r["value_1core"] > 0 and "16^3" in r["sample"]
