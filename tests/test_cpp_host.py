"""Host-only parts of the C++ API (no GPU): the reordering of sets
(test/reorder-tests.cpp's cases, the Hilbert curve, the bunny) and the
tetgen loader, run from the compiled program-test driver."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "simit_amd" / "bin" / "simit_program_tests"


@pytest.mark.parametrize("name", ["reorder_int", "reorder_double", "hilbert_curve",
                                  "reorder_bunny_host", "mesh_load", "grid_set"])
def test_cpp_host(name):
    if not BIN.exists():
        pytest.skip(f"{BIN} is not built (run __graft_entry__.build())")
    r = subprocess.run([str(BIN), name], capture_output=True, text=True, timeout=120,
                       cwd=str(ROOT))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"[  OK  ] {name}" in r.stdout
