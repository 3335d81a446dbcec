"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the reported CPU baseline.
The product (simit_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
import struct
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
LIB_OMP = HERE / "liboracle_omp.so"  # the same source with OpenMP (bench baseline only)
REF_INDEX = HERE / "_ref" / "ref_index"

_lib = None
_lib_omp = None


def build(ref: bool = False):
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.run(["make", "-C", str(HERE), *targets], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _declare(_lib)
    return _lib


def lib_omp():
    """The OpenMP build of the same restatement: multi-core CPU baseline only
    (its dot products sum in a different order; never a checker)."""
    global _lib_omp
    if _lib_omp is None:
        if not LIB_OMP.exists():
            build()
        _lib_omp = C.CDLL(str(LIB_OMP))
        _declare(_lib_omp)
    return _lib_omp


class SpringsParams(C.Structure):
    _fields_ = [("h", C.c_double), ("viscous", C.c_double), ("grav", C.c_double * 3),
                ("fixed_fill", C.c_double), ("fixed_identity", C.c_int32)]


class TetParams(C.Structure):
    _fields_ = [("stiff_scale", C.c_double), ("rho", C.c_double), ("load", C.c_double * 3),
                ("fixed_diag", C.c_double)]


class HexParams(C.Structure):
    _fields_ = [("h", C.c_double), ("source", C.c_double), ("fixed_diag", C.c_double)]


class CGOpts(C.Structure):
    _fields_ = [("maxiters", C.c_int32), ("rel", C.c_int32), ("replace_every", C.c_int32),
                ("precond", C.c_int32), ("norm_form", C.c_int32), ("tol", C.c_double)]


class CGStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("rr", C.c_double), ("true_rel", C.c_double)]


def _declare(L):
    vp = C.c_void_p
    L.or_index_ve.restype = C.c_int64
    L.or_index_ve.argtypes = [C.c_int32, C.c_int64, C.c_int32, vp, vp, vp]
    L.or_index_vev.restype = C.c_int64
    L.or_index_vev.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, C.c_int32, vp, vp, C.c_int64]
    L.or_loc.restype = C.c_int32
    L.or_loc.argtypes = [vp, vp, C.c_int32, C.c_int32]
    L.or_det3.restype = C.c_double
    L.or_det3.argtypes = [vp]
    L.or_inv3.argtypes = [vp, vp]
    L.or_springs_assemble.argtypes = [C.POINTER(SpringsParams), C.c_int32, C.c_int64] + [vp] * 11
    L.or_tet_block.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_int, vp, vp]
    L.or_tet_assemble.argtypes = [C.POINTER(TetParams), C.c_int32, C.c_int64] + [vp] * 9
    L.or_hex_kref.restype = C.c_double
    L.or_hex_kref.argtypes = [C.c_int, C.c_int]
    L.or_hex_assemble.argtypes = [C.POINTER(HexParams), C.c_int32, C.c_int64] + [vp] * 7
    L.or_bsrmv.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, vp, vp]
    L.or_dot.restype = C.c_double
    L.or_dot.argtypes = [C.c_int64, vp, vp]
    L.or_cg.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, vp, vp, C.POINTER(CGOpts),
                        C.POINTER(CGStats)]
    L.or_springs_step.argtypes = [C.POINTER(SpringsParams), C.POINTER(CGOpts), C.c_int32,
                                  C.c_int64] + [vp] * 9 + [C.POINTER(CGStats)]
    L.or_log.restype = C.c_double
    L.or_log.argtypes = [C.c_double]
    L.or_femtet_geom.argtypes = [vp, vp, vp]
    L.or_femtet_precompute.argtypes = [C.c_double, C.c_int32, C.c_int64] + [vp] * 5
    L.or_femtet_H.argtypes = [vp, C.c_double, C.c_double, C.c_double, vp, vp]
    L.or_femtet_kb.argtypes = [vp, C.c_double, C.c_double, C.c_double, vp, C.c_int, C.c_int, vp]
    L.or_femtet_assemble.argtypes = [C.c_double, C.c_double, vp, C.c_int32, C.c_int64] + [vp] * 12


def _p(a):
    return a.ctypes.data if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---- sparsity ----------------------------------------------------------------------
def index_vev(nv, endpoint_sets, force_diag=False):
    sets = [_c(e, np.int32) for e in endpoint_sets]
    sets = [s if s.ndim == 2 else s.reshape(s.shape[0], -1) for s in sets]
    ns = len(sets)
    ne = np.array([s.shape[0] for s in sets], dtype=np.int64)
    card = np.array([s.shape[1] for s in sets], dtype=np.int32)
    ptrs = (C.c_void_p * ns)(*[s.ctypes.data for s in sets])
    cap = int(sum(int(s.shape[0]) * int(s.shape[1]) ** 2 for s in sets)) + nv
    rowptr = np.zeros(nv + 1, dtype=np.uint32)
    colidx = np.zeros(max(cap, 1), dtype=np.uint32)
    nnz = lib().or_index_vev(nv, ns, _p(ne), _p(card), ptrs, int(force_diag), _p(rowptr),
                             _p(colidx), cap)
    assert nnz >= 0
    return rowptr, colidx[:nnz].copy()


def index_ve(nv, endpoints):
    e = _c(endpoints, np.int32).reshape(endpoints.shape[0], -1)
    ptr = np.zeros(nv + 1, dtype=np.uint32)
    idx = np.zeros(max(e.size, 1), dtype=np.uint32)
    n = lib().or_index_ve(nv, e.shape[0], e.shape[1], _p(e), _p(ptr), _p(idx))
    return ptr, idx[:n].copy()


def loc(rowptr, colidx, i, j):
    return lib().or_loc(_p(rowptr), _p(colidx), i, j)


def det3(a):
    return lib().or_det3(_p(_c(a, np.float64).ravel()))


def inv3(a):
    out = np.zeros(9)
    lib().or_inv3(_p(_c(a, np.float64).ravel()), _p(out))
    return out.reshape(3, 3)


# ---- assembly ----------------------------------------------------------------------
def springs_params(h=1e-2, viscous=0.0, grav=(0.0, 0.0, -9.8), fixed_fill=1.0,
                   fixed_identity=1):
    return SpringsParams(h, viscous, (C.c_double * 3)(*grav), fixed_fill, fixed_identity)


def springs_assemble(P, pts, springs, rowptr, colidx):
    np_ = int(pts["x"].shape[0])
    ep = _c(springs["endpoints"], np.int32)
    nnz = int(rowptr[-1])
    vals = np.zeros(max(nnz, 1) * 9)
    f = np.zeros(np_ * 3)
    lib().or_springs_assemble(C.byref(P), np_, ep.shape[0], _p(ep), _p(_c(pts["x"], np.float64)),
                              _p(_c(pts["v"], np.float64)), _p(_c(pts["m"], np.float64)),
                              _p(_c(pts["fixed"], np.uint8)), _p(_c(springs["k"], np.float64)),
                              _p(_c(springs["l0"], np.float64)), _p(rowptr), _p(colidx),
                              _p(vals), _p(f))
    return vals[: nnz * 9], f


def tet_params(stiff_scale=1.0, rho=0.0, load=(0.0, 0.0, -9.81e3), fixed_diag=1.0):
    return TetParams(stiff_scale, rho, (C.c_double * 3)(*load), fixed_diag)


def tet_block(xs, mu, lam, a, b):
    blk = np.zeros(9)
    W = C.c_double()
    lib().or_tet_block(_p(_c(xs, np.float64).ravel()), mu, lam, a, b, _p(blk), C.addressof(W))
    return blk.reshape(3, 3), W.value


def tet_assemble(P, prob, rowptr, colidx):
    nv = int(prob["nverts"])
    ep = _c(prob["endpoints"], np.int32)
    nnz = int(rowptr[-1])
    vals = np.zeros(max(nnz, 1) * 9)
    f = np.zeros(nv * 3)
    lib().or_tet_assemble(C.byref(P), nv, ep.shape[0], _p(ep), _p(_c(prob["x"], np.float64)),
                          _p(_c(prob["c"], np.int32)), _p(_c(prob["mu"], np.float64)),
                          _p(_c(prob["lam"], np.float64)), _p(rowptr), _p(colidx), _p(vals),
                          _p(f))
    return vals[: nnz * 9], f


# ---- neo-Hookean tets (test/input/program/femTet.sim) -----------------------------------
def log(x):
    return lib().or_log(float(x))


def femtet_precompute(prob, rho=1000.0):
    """initializeTet: per-tet W, B (row-major 3x3) and the vertex m field
    (prob['m'] is updated in place, as the map assigns it)."""
    ep = _c(prob["endpoints"], np.int32)
    nt = ep.shape[0]
    W = np.zeros(nt)
    B = np.zeros(nt * 9)
    m = prob["m"]
    assert m.dtype == np.float64 and m.flags.c_contiguous
    lib().or_femtet_precompute(rho, int(prob["nverts"]), nt, _p(ep), _p(_c(prob["x"], np.float64)),
                               _p(W), _p(B), _p(m))
    return W, B


def femtet_H(xs, u, l, W, B):
    H = np.zeros(9)
    lib().or_femtet_H(_p(_c(xs, np.float64).ravel()), u, l, W, _p(_c(B, np.float64).ravel()),
                      _p(H))
    return H.reshape(3, 3)


def femtet_kb(xs, u, l, W, B, jj, row):
    out = np.zeros(9)
    lib().or_femtet_kb(_p(_c(xs, np.float64).ravel()), u, l, W, _p(_c(B, np.float64).ravel()),
                       jj, row, _p(out))
    return out.reshape(3, 3)


def femtet_assemble(prob, rowptr, colidx, h=0.01, rho=1000.0, grav=(0.0, -10.0, 0.0)):
    """main's f = map compute_force(h) and A = map compute_stiffness(h)."""
    nv = int(prob["nverts"])
    ep = _c(prob["endpoints"], np.int32)
    nnz = int(rowptr[-1])
    vals = np.zeros(max(nnz, 1) * 9)
    f = np.zeros(nv * 3)
    g = _c(grav, np.float64)
    lib().or_femtet_assemble(h, rho, _p(g), nv, ep.shape[0], _p(ep),
                             _p(_c(prob["x"], np.float64)), _p(_c(prob["v"], np.float64)),
                             _p(_c(prob["c"], np.int32)), _p(_c(prob["u"], np.float64)),
                             _p(_c(prob["l"], np.float64)), _p(_c(prob["W"], np.float64)),
                             _p(_c(prob["B"], np.float64)), _p(rowptr), _p(colidx), _p(vals),
                             _p(f))
    return vals[: nnz * 9], f


def femtet_step(prob, rowptr, colidx, h=0.01):
    """One femTet.sim main(): assemble, CG from v (absolute 1e-6 on r'r, 100
    iterations), v = x, x = h*v + x.  Updates prob['x'] and prob['v']."""
    vals, f = femtet_assemble(prob, rowptr, colidx, h)
    x, st = cg(rowptr, colidx, vals, f, prob["v"], 3, cg_opts(maxiters=100, tol=1e-6, rel=0))
    prob["v"][:] = x
    prob["x"][:] = h * x + prob["x"]
    return st


def hex_kref(a, b):
    return lib().or_hex_kref(a, b)


def hex_assemble(P, prob, rowptr, colidx):
    nv = int(prob["nverts"])
    ep = _c(prob["endpoints"], np.int32)
    nnz = int(rowptr[-1])
    vals = np.zeros(max(nnz, 1))
    f = np.zeros(nv)
    lib().or_hex_assemble(C.byref(P), nv, ep.shape[0], _p(ep), _p(_c(prob["dirichlet"], np.int32)),
                          _p(_c(prob["kappa"], np.float64)), _p(rowptr), _p(colidx), _p(vals),
                          _p(f))
    return vals[:nnz], f


# ---- linear algebra ------------------------------------------------------------------
def bsrmv(rowptr, colidx, vals, x, bdim):
    n = len(rowptr) - 1
    y = np.zeros(n * bdim)
    lib().or_bsrmv(n, bdim, _p(rowptr), _p(colidx), _p(_c(vals, np.float64)),
                   _p(_c(x, np.float64)), _p(y))
    return y


def dot(x, y):
    x, y = _c(x, np.float64), _c(y, np.float64)
    return lib().or_dot(x.size, _p(x), _p(y))


def cg_opts(maxiters=1000, tol=1e-8, rel=1, replace_every=0, precond=0, norm_form=0):
    return CGOpts(maxiters, rel, replace_every, precond, norm_form, tol)


def cg(rowptr, colidx, vals, b, x0, bdim, opts, omp=False):
    x = _c(x0, np.float64).copy()
    st = CGStats()
    n = len(rowptr) - 1
    (lib_omp() if omp else lib()).or_cg(n, bdim, _p(rowptr), _p(colidx), _p(_c(vals, np.float64)),
                                        _p(_c(b, np.float64)), _p(x), C.byref(opts), C.byref(st))
    return x, {"iters": st.iters, "rr": st.rr, "true_rel": st.true_rel}


def springs_step(P, opts, pts, springs, rowptr, colidx):
    """One isprings.sim step; updates pts['x'] and pts['v'] in place."""
    st = CGStats()
    ep = _c(springs["endpoints"], np.int32)
    x = pts["x"]
    v = pts["v"]
    assert x.dtype == np.float64 and x.flags.c_contiguous
    assert v.dtype == np.float64 and v.flags.c_contiguous
    lib().or_springs_step(C.byref(P), C.byref(opts), int(x.shape[0]), ep.shape[0], _p(ep), _p(x),
                          _p(v), _p(_c(pts["m"], np.float64)), _p(_c(pts["fixed"], np.uint8)),
                          _p(_c(springs["k"], np.float64)), _p(_c(springs["l0"], np.float64)),
                          _p(rowptr), _p(colidx), C.byref(st))
    return {"iters": st.iters, "rr": st.rr, "true_rel": st.true_rel}


# ---- the reference's own index builder (oracle/_ref, built from the reference) ----------
def ref_index_available():
    return REF_INDEX.exists()


def ref_index_vev(nv, endpoint_sets, tmpdir, mode="vev"):
    inp = Path(tmpdir) / "ref_in.bin"
    out = Path(tmpdir) / "ref_out.bin"
    with open(inp, "wb") as f:
        f.write(struct.pack("<ii", nv, len(endpoint_sets)))
        for e in endpoint_sets:
            e = _c(e, np.int32).reshape(e.shape[0], -1)
            f.write(struct.pack("<qi", e.shape[0], e.shape[1]))
            f.write(e.tobytes())
    subprocess.run([str(REF_INDEX), mode, str(inp), str(out)], check=True)
    b = out.read_bytes()
    n, nnz = struct.unpack("<iq", b[:12])
    rp = np.frombuffer(b[12:12 + 4 * (n + 1)], dtype=np.uint32).copy()
    ci = np.frombuffer(b[12 + 4 * (n + 1):12 + 4 * (n + 1) + 4 * nnz], dtype=np.uint32).copy()
    return rp, ci


def ref_create_box(nx, ny, nz, tmpdir):
    out = Path(tmpdir) / "ref_box.bin"
    subprocess.run([str(REF_INDEX), "box", str(nx), str(ny), str(nz), str(out)], check=True)
    b = out.read_bytes()
    (ne,) = struct.unpack("<q", b[:8])
    return np.frombuffer(b[8:8 + 8 * ne], dtype=np.int32).reshape(ne, 2).copy()
