"""The C++ simit::Set / Program / Function API on the HIP backend.

Runs the compiled program tests (simit_amd/csrc/tests_cpp/program_tests.cpp),
which follow the reference's gtest program tests (test/program-isprings.cpp,
test/program-cg-test.cpp) and assert the same values.
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "simit_amd" / "bin" / "simit_program_tests"


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["isprings", "cg", "springs_lattice", "mesh_bunny",
                                  "reorder_bunny_springs"])
def test_cpp_program(name):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (run __graft_entry__.build())")
    r = subprocess.run([str(BIN), name], capture_output=True, text=True, timeout=90,
                       cwd=str(ROOT))
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"[  OK  ] {name}" in r.stdout
