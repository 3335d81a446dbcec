"""The Simit-source front-end without a device: Program::loadFile parses,
checks and generates the HIP kernels of every exported function, and the
generated source compiles for gfx950 with hiprtc.

Covers this backend's own programs (tests/programs/) and, where the
reference tree is present, the reference's own hot-path programs read in
place (test/input/program/{isprings,cg,femTet,...}.sim,
apps/fem/fem_linear.sim, apps/springs/isprings.sim).
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "simit_amd" / "bin" / "simit_program_tests"
REF = Path("/root/reference")


def compile_program(path, *functions, text=None, tmp=None):
    if text is not None:
        path = tmp / "prog.sim"
        path.write_text(text)
    r = subprocess.run([str(BIN), "compile", str(path), *functions], capture_output=True,
                       text=True, timeout=300, cwd=str(ROOT))
    return r.returncode, r.stdout + r.stderr


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (run __graft_entry__.build())")


OWN_PROGRAMS = [
    ("springs.sim", ["main"]),
    ("cg_points.sim", ["main"]),
    ("neohookean.sim", ["initializeTet", "main"]),
    ("gemv.sim", ["main"]),
    ("grid_gemv.sim", ["main"]),
    ("grid_gemv_2d.sim", ["main"]),
]


@pytest.mark.parametrize("name,functions", OWN_PROGRAMS)
def test_own_programs_compile(name, functions):
    rc, out = compile_program(ROOT / "tests" / "programs" / name)
    assert rc == 0, out
    for f in functions:
        assert f"{name}: {f}:" in out and "code object" in out, out


REF_PROGRAMS = [
    "test/input/program/isprings.sim",
    "test/input/program/cg.sim",
    "test/input/program/femTet.sim",
    "test/input/program/isprings_simple.sim",
    "test/input/program/esprings.sim",
    "apps/fem/fem_linear.sim",
    "apps/fem/fem_neohookean.sim",
    "apps/springs/isprings.sim",
    "apps/springs/esprings.sim",
    "test/input/system/gemv.sim",
    "test/input/system/gemv_stencil.sim",
    "test/input/system/gemv_stencil_2d.sim",
    "test/input/system/gemv_stencil_indexless.sim",
    "test/input/system/gemv_stencil_2d_indexless.sim",
]


@pytest.mark.parametrize("rel", REF_PROGRAMS)
def test_reference_programs_compile_unchanged(rel):
    path = REF / rel
    if not path.exists():
        pytest.skip("reference tree not present")
    rc, out = compile_program(path)
    assert rc == 0, out
    assert "COMPILE FAILED" not in out and "not compiled" not in out, out


def test_parse_error_names_line_and_column(tmp_path):
    rc, out = compile_program(None, text="element P\n  x : float;\nend\n\nfunc f(\n", tmp=tmp_path)
    assert rc == 1
    assert "6:1:" in out or "5:" in out, out


@pytest.mark.parametrize("body,message", [
    # a result written inside a while loop has no static write slot
    ("func f(p : P) -> (y : tensor[ps](float))\n  var i = 0;\n  while i < 2\n"
     "    y(p) = 1.0;\n    i = i + 1;\n  end\nend\n", "while loop"),
    # a constant endpoint index beyond the tuple
    ("func f(s : S, p : (P*2)) -> (y : tensor[ps](float))\n  y(p(2)) = 1.0;\nend\n",
     "out of range"),
    # a tensor index beyond the tensor
    ("func f(p : P) -> (y : tensor[ps](float))\n  var t : tensor[3](float);\n"
     "  for i in 0:4\n    t(i) = 1.0;\n  end\n  y(p) = t(0);\nend\n", "out of range"),
])
def test_unsupported_or_wrong_element_code_is_rejected(tmp_path, body, message):
    target = "ss" if "(P*2)" in body else "ps"
    text = ("element P\n  x : float;\nend\nelement S\n  k : float;\nend\n"
            "extern ps : set{P};\nextern ss : set{S}(ps, ps);\n" + body +
            f"export func main()\n  y = map f to {target} reduce +;\nend\n")
    rc, out = compile_program(None, "main", text=text, tmp=tmp_path)
    assert rc == 1, out
    assert message in out, out


def test_grid_reads_are_checked(tmp_path):
    """A link must join two neighbouring grid points; grid matrices are
    written in their own row within one step (27-point stencil)."""
    head = ("element P\n  b : float;\nend\nelement L\n  a : float;\nend\n"
            "extern ps : set{P};\nextern g : grid[2]{L}(ps);\n")
    tail = "export func main()\n  B = map f to ps through g;\n  ps.b = B*ps.b;\nend\n"
    cases = [
        ("  B(p,p) = l[0,0;1,1].a;\n", "neighbouring grid points"),
        ("  B(p,ps[2,0]) = l[0,0;1,0].a;\n", "within one step"),
        ("  B(ps[1,0],p) = l[0,0;1,0].a;\n", "own row"),
    ]
    for body, message in cases:
        text = (head + "func f(p : P, l : grid[2]{L}(ps)) -> (B : tensor[ps,ps](float))\n" + body +
                "end\n" + tail)
        rc, out = compile_program(None, "main", text=text, tmp=tmp_path)
        assert rc == 1 and message in out, out
