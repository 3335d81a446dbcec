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


def test_struct_layouts_match_header():
    # sizes a C compiler gives these structs on x86-64
    assert ctypes.sizeof(simit_amd.Bsr) == 56
    assert ctypes.sizeof(simit_amd.CGOptions) == 80
    assert ctypes.sizeof(simit_amd.CGBatchStats) == 16
    assert ctypes.sizeof(simit_amd.CGStats) == 40
    assert ctypes.sizeof(simit_amd.SpringsParams) == 56


def test_context_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = simit_amd.load()
    h = ctypes.c_void_p()
    assert lib.smh_ctx_create(0, ctypes.byref(h)) != 0
    assert not h.value
