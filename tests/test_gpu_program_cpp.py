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
                                  "sim_reorder_fem_tet", "sim_gemv", "sim_gemv_grid",
                                  "sim_gemv_grid_2d"])
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


# linear elasticity in the Voigt form of apps/fem/fem_linear.sim:95-174
# (strain-displacement matrix S, Ke = W S' E S), written for these tests
LINEAR_STIFFNESS = """
func corner_block(M : tensor[12,12](float), r : int, c : int) -> (X : tensor[3,3](float))
  for i in 0:3
    for j in 0:3
      X(i,j) = M(r+i, c+j);
    end
  end
end

func linear_stiffness(h : float, t : Tet, v : (Vert*4))
    -> (K : tensor[verts,verts](tensor[3,3](float)))
  var E : tensor[6,6](float) = 0.0;
  for i in 0:3
    for j in 0:3
      E(i,j) = t.l;
    end
    E(i,i) = t.l + 2.0*t.u;
    E(i+3,i+3) = t.u;
  end
  var g : tensor[4,3](float) = 0.0;
  for a in 0:3
    for d in 0:3
      g(a,d) = t.B(a,d);
      g(3,d) = g(3,d) - t.B(a,d);
    end
  end
  var S : tensor[6,12](float) = 0.0;
  for a in 0:4
    for d in 0:3
      S(d, 3*a+d) = g(a,d);
    end
    S(3, 3*a) = g(a,1);
    S(3, 3*a+1) = g(a,0);
    S(4, 3*a+1) = g(a,2);
    S(4, 3*a+2) = g(a,1);
    S(5, 3*a) = g(a,2);
    S(5, 3*a+2) = g(a,0);
  end
  Ke = t.W*S'*E*S;
  for a in 0:4
    for b in 0:4
      if (v(a).c <= 0) and (v(b).c <= 0)
        K(v(a), v(b)) = h*h*corner_block(Ke, 3*a, 3*b);
      end
    end
  end
  mass = 0.25*1000.0*t.W;
  for a in 0:4
    K(v(a), v(a)) = mass*[1.0, 0.0, 0.0; 0.0, 1.0, 0.0; 0.0, 0.0, 1.0];
  end
end
"""


def operators_program(tmp_path, stiffness="stiffness"):
    """tests/programs/neohookean.sim with two more vertex fields and an
    exported `operators`: main's force map into verts.f, a stiffness map,
    and verts.y = A*verts.v."""
    text = (ROOT / "tests" / "programs" / "neohookean.sim").read_text()
    text = text.replace("  m  : float;\nend", "  m  : float;\n  y  : tensor[3](float);\n"
                        "  f  : tensor[3](float);\nend", 1)
    assert "y  : tensor[3]" in text
    text += LINEAR_STIFFNESS
    text += ("\nexport func operators()\n  h = 0.01;\n"
             "  verts.f = map internal_force(h) to tets reduce +;\n"
             f"  A = map {stiffness}(h) to tets reduce +;\n  verts.y = A*verts.v;\nend\n")
    path = tmp_path / f"operators_{stiffness}.sim"
    path.write_text(text)
    return path


def run_operators(tmp_path, stiffness):
    prog = operators_program(tmp_path, stiffness)
    r = subprocess.run([str(BIN), "sim_fem_ops", str(prog), str(tmp_path)], capture_output=True,
                       text=True, timeout=120, cwd=str(ROOT))
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    return {n: np.fromfile(tmp_path / f"{n}.f64") for n in ("W", "B", "m", "f", "y", "v")}


@pytest.mark.gpu
def test_generated_femtet_kernels_match_the_oracle(tmp_path, oracle):
    """The generated element kernels of the parsed neo-Hookean program on the
    bar2k mesh: initializeTet's B, W and masses, main's force and the
    stiffness applied to v (y = A v, the exact-order SpMV) equal the oracle's
    restatement of femTet.sim in FP64 (same operations in the same order)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simit_amd import meshes
    got = run_operators(tmp_path, "stiffness")
    golden = ROOT / "tests" / "golden" / "fem-bar2k"
    prob = meshes.femtet_problem(golden / "bar2k.node", golden / "bar2k.ele")
    W, B = oracle.femtet_precompute(prob)
    np.testing.assert_array_equal(got["W"], W)
    np.testing.assert_array_equal(got["B"], B)
    np.testing.assert_array_equal(got["m"], prob["m"])
    prob["W"], prob["B"] = W, B
    rp, ci = oracle.index_vev(prob["nverts"], [prob["endpoints"]], force_diag=True)
    ov, of = oracle.femtet_assemble(prob, rp, ci)
    np.testing.assert_array_equal(got["f"], of)
    np.testing.assert_array_equal(got["v"], prob["v"].ravel())
    y = oracle.bsrmv(rp, ci, ov, prob["v"].ravel(), 3)
    np.testing.assert_array_equal(got["y"], y)


@pytest.mark.gpu
def test_generated_linear_stiffness_matches_the_closed_form(tmp_path, oracle):
    """A fem_linear-style stiffness written with the Voigt matrices
    (W S' E S, a 12x12 slice taken by a function with computed offsets) on
    bar2k: A v against the oracle's closed-form blocks of the same element
    (apps/fem/fem_linear.sim:95-174 with h^2 scaling, clamped pairs dropped,
    lumped mass) to rounding."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simit_amd import meshes
    got = run_operators(tmp_path, "linear_stiffness")
    golden = ROOT / "tests" / "golden" / "fem-bar2k"
    prob = meshes.femtet_problem(golden / "bar2k.node", golden / "bar2k.ele")
    prob["mu"], prob["lam"] = prob["u"], prob["l"]
    rp, ci = oracle.index_vev(prob["nverts"], [prob["endpoints"]], force_diag=True)
    ov, _ = oracle.tet_assemble(oracle.tet_params(stiff_scale=0.01 * 0.01, rho=1000.0,
                                                  load=(0.0, 0.0, 0.0), fixed_diag=0.0),
                                prob, rp, ci)
    y = oracle.bsrmv(rp, ci, ov, prob["v"].ravel(), 3)
    err = np.abs(got["y"] - y).max() / np.abs(y).max()
    assert err < 1e-12, err
