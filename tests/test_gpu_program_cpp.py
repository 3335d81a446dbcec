"""The C++ simit::Set / Program / Function API on the HIP backend.

Runs the compiled program tests (simit_amd/csrc/tests_cpp/program_tests.cpp),
which follow the reference's gtest program tests (test/program-isprings.cpp,
test/program-cg-test.cpp, test/program-fem-test.cpp) and assert the same
values.  The sim_* cases run the same tests through Simit source parsed by
Program::loadFile (tests/programs/*.sim): element functions generated as
HIP and compiled at init, the exported function interpreted over the C ABI.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "simit_amd" / "bin" / "simit_program_tests"


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["isprings", "cg", "springs_lattice", "mesh_bunny",
                                  "reorder_bunny_springs", "extern_csr", "fem_tet",
                                  "reorder_fem_tet", "sim_isprings", "sim_cg", "sim_fem_tet",
                                  "sim_reorder_fem_tet"])
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


@pytest.mark.gpu
def test_cpp_extern_csr_tet_system(tmp_path, oracle):
    """A 10^3 x 5 tet system assembled by the oracle, handed to the C++ API
    as CSR TensorData (Function::bind(name, TensorData&)): y = A b through
    the exact-order SpMV equals the oracle's product bit for bit, and
    x = cg(A, b) reaches 1e-8 on the oracle's matrix."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simit_amd import meshes
    prob = meshes.tet_problem(10, seed=5, field="grf", std_log10=1.0, corr_len=3.0)
    rp, ci = oracle.index_vev(prob["nverts"], [prob["endpoints"]])
    ov, of = oracle.tet_assemble(oracle.tet_params(), prob, rp, ci)
    rp.astype(np.int32).tofile(tmp_path / "rowptr.i32")
    ci.astype(np.int32).tofile(tmp_path / "colidx.i32")
    ov.astype(np.float64).tofile(tmp_path / "vals.f64")
    of.astype(np.float64).tofile(tmp_path / "b.f64")
    r = subprocess.run([str(BIN), "extern_file", str(tmp_path)], capture_output=True, text=True,
                       timeout=90, cwd=str(ROOT))
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    y = np.fromfile(tmp_path / "y.f64", dtype=np.float64)
    x = np.fromfile(tmp_path / "x.f64", dtype=np.float64)
    np.testing.assert_array_equal(y, oracle.bsrmv(rp, ci, ov, of, 3))
    res = of - oracle.bsrmv(rp, ci, ov, x, 3)
    assert np.linalg.norm(res) <= 1.01e-8 * np.linalg.norm(of)
