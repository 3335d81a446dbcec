"""The C++ host layer under AddressSanitizer and UndefinedBehaviorSanitizer
(no GPU): the reordering and mesh-loading cases, grid edge sets, and the
front-end (parser, code generator) on its own programs, the reference's
programs and the error paths.  Any sanitizer report fails the test."""
import os
import subprocess
from pathlib import Path

import pytest

import test_frontend as FE

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "simit_amd" / "bin" / "simit_program_tests_asan"
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def run(*args, tmp=None):
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (run __graft_entry__.build())")
    r = subprocess.run([str(BIN), *args], capture_output=True, text=True, timeout=300,
                       cwd=str(tmp or ROOT), env=ENV)
    out = r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out
    return r.returncode, out


@pytest.mark.parametrize("name", ["reorder_int", "reorder_double", "hilbert_curve",
                                  "reorder_bunny_host", "mesh_load", "grid_set"])
def test_host_case(name):
    rc, out = run(name)
    assert rc == 0 and f"[  OK  ] {name}" in out, out


def test_own_programs():
    for name, _ in FE.OWN_PROGRAMS:
        rc, out = run("compile", str(ROOT / "tests" / "programs" / name))
        assert rc == 0 and "code object" in out, out


def test_reference_programs():
    paths = [FE.REF / rel for rel in FE.REF_PROGRAMS if (FE.REF / rel).exists()]
    if not paths:
        pytest.skip("reference tree not present")
    for p in paths:
        rc, out = run("compile", str(p))
        assert rc == 0 and "COMPILE FAILED" not in out, out


def test_error_paths(tmp_path):
    bad = tmp_path / "bad.sim"
    bad.write_text("element P\n  x : float;\nend\n\nfunc f(\n")
    rc, out = run("compile", str(bad), tmp=tmp_path)
    assert rc == 1, out
    text = ("element P\n  x : float;\nend\nelement S\n  k : float;\nend\n"
            "extern ps : set{P};\nextern ss : set{S}(ps, ps);\n"
            "func f(s : S, p : (P*2)) -> (y : tensor[ps](float))\n  y(p(2)) = 1.0;\nend\n"
            "export func main()\n  y = map f to ss reduce +;\nend\n")
    (tmp_path / "range.sim").write_text(text)
    rc, out = run("compile", str(tmp_path / "range.sim"), "main", tmp=tmp_path)
    assert rc == 1 and "out of range" in out, out
