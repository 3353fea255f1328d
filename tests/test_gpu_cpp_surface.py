"""The host backend's tensor surface from C++: a solver template written against that surface
only (tests/cpp/template_surface.cpp, built by __graft_entry__.build()) runs on HipVector /
HipCrsMatrix and agrees with the backend's fused ConjugateGradient."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "cpp" / "template_surface"


def test_templated_solver_on_hip_tensors():
    assert BIN.exists(), f"{BIN} not built (run __graft_entry__.build())"
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK")
