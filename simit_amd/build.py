"""Build recipe for libsimit_hip.so (gfx950) and the C++ host-API test driver.

Compiles in-tree with hipcc: the shared library lands in simit_amd/lib/,
next to this file, so it travels with the source tree.
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
BINDIR = PKG / "bin"
LIB = LIBDIR / "libsimit_hip.so"
HOST_LIB = LIBDIR / "libsimit.so"
PROGRAM_TESTS = BINDIR / "simit_program_tests"

HIP_SOURCES = ["smh_core.hip", "smh_index.hip", "smh_assemble.hip", "smh_cg.hip", "smh_gnn.hip",
               "smh_jit.hip"]
HOST_SOURCES = ["simit/graph.cpp", "simit/program.cpp", "simit/mesh.cpp",
                "simit/reorder.cpp", "simit/parser.cpp", "simit/codegen.cpp",
                "simit/sim_function.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed-FP32 instructions are disabled: all FP32 math is scalar VALU
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"), "-Wno-unused-result",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(jobs: int = 4) -> Path:
    """Compile every HIP translation unit and link libsimit_hip.so."""
    LIBDIR.mkdir(exist_ok=True)
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    procs, objs = [], []
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        if len(procs) >= jobs:
            _wait(procs)
    _wait(procs)
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", str(LIB), *map(str, objs), "-L/opt/rocm/lib",
          "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"])
    return LIB


def _wait(procs):
    while procs:
        src, p = procs.pop(0)
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed on {src}")


def build_host() -> Path:
    """The simit::Set / Program / Function work-alike (C++), linked on the C ABI."""
    LIBDIR.mkdir(exist_ok=True)
    srcs = [str(CSRC / s) for s in HOST_SOURCES]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", str(ROOT / "include"),
          "-I", str(CSRC), *srcs, "-L", str(LIBDIR), "-lsimit_hip",
          "-Wl,-rpath,$ORIGIN", "-ldl", "-o", str(HOST_LIB)])
    BINDIR.mkdir(exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "include"), "-I", str(CSRC),
          str(CSRC / "tests_cpp" / "program_tests.cpp"), "-L", str(LIBDIR), "-lsimit",
          "-lsimit_hip", "-Wl,-rpath,$ORIGIN/../lib", "-o", str(PROGRAM_TESTS)])
    return HOST_LIB


PROGRAM_TESTS_ASAN = BINDIR / "simit_program_tests_asan"


def build_host_asan() -> Path:
    """The host layer (sets, graphs, reordering, mesh loading, parser, code
    generator, programs) and its test driver built with AddressSanitizer and
    UndefinedBehaviorSanitizer, as one executable on the unsanitized device
    library.  tests/test_host_asan.py runs its host-only cases."""
    BINDIR.mkdir(exist_ok=True)
    srcs = [str(CSRC / s) for s in HOST_SOURCES]
    _run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
          "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined",
          "-I", str(ROOT / "include"), "-I", str(CSRC), *srcs,
          str(CSRC / "tests_cpp" / "program_tests.cpp"), "-L", str(LIBDIR), "-lsimit_hip",
          "-Wl,-rpath,$ORIGIN/../lib", "-ldl", "-o", str(PROGRAM_TESTS_ASAN)])
    return PROGRAM_TESTS_ASAN


def build_all():
    build_hip()
    if (CSRC / "simit" / "program.cpp").exists():
        build_host()
        build_host_asan()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
