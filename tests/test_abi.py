"""The C ABI library loads and exports every entry point include/simit_hip.h
declares (no compute: this runs without a GPU)."""
import ctypes
import re

import pytest

import simit_amd


def declared_functions():
    text = simit_amd.HEADER_PATH.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("smh_index_build", "smh_assemble_springs", "smh_assemble_tet_linear",
                 "smh_assemble_hex_poisson", "smh_bsrmv", "smh_dot", "smh_cg_solve"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = simit_amd.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_abi():
    lib = simit_amd.load()
    for n in declared_functions():
        assert getattr(lib, n).argtypes is not None, f"{n} has no ctypes signature"


# ctypes mirror -> C struct name in include/simit_hip.h
STRUCTS = {"IndexView": "smh_index_view", "Bsr": "smh_bsr", "Stencil27": "smh_stencil27",
           "SpringsParams": "smh_springs_params", "SpringsSet": "smh_springs_set",
           "TetParams": "smh_tet_params", "TetSet": "smh_tet_set",
           "HexParams": "smh_hex_params", "HexSet": "smh_hex_set",
           "FemTetParams": "smh_femtet_params", "FemTetSet": "smh_femtet_set",
           "EdgeBlocks": "smh_edge_blocks", "CGOptions": "smh_cg_opts",
           "CGBatchStats": "smh_cg_batch_stats", "CGStats": "smh_cg_stats"}


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror has the size and field offsets gcc gives the
    header's struct."""
    import subprocess
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "simit_hip.h"',
             "int main(void) {"]
    for py, cs in STRUCTS.items():
        lines.append(f'  printf("{py} size %zu\\n", sizeof({cs}));')
        for f, _ in getattr(simit_amd, py)._fields_:
            lines.append(f'  printf("{py} {f} %zu\\n", offsetof({cs}, {f}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(simit_amd.HEADER_PATH.parent), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        py, f, v = line.split()
        got[(py, f)] = int(v)
    for py in STRUCTS:
        cls = getattr(simit_amd, py)
        assert ctypes.sizeof(cls) == got[(py, "size")], py
        for f, _ in cls._fields_:
            assert getattr(cls, f).offset == got[(py, f)], (py, f)


def test_context_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = simit_amd.load()
    h = ctypes.c_void_p()
    assert lib.smh_ctx_create(0, ctypes.byref(h)) != 0
    assert not h.value
