"""simit_amd -- MI355X execution backend for Simit's sparsity / assembly /
blocked-SpMV / CG path.

The product is the C ABI in include/simit_hip.h, implemented by hand-written
gfx950 kernels in simit_amd/csrc/ and built into simit_amd/lib/libsimit_hip.so.
This module is a thin ctypes binding of that ABI for Python callers (tests,
the benchmark).  Device memory is handed over as raw device pointers; torch
tensors are accepted for convenience (torch is plumbing here, never compute).

There is no fallback: if the HIP library is missing or fails to load, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

__all__ = [
    "SimitError", "load", "Context", "Index", "Bsr", "CGOptions", "CGStats",
    "SpringsParams", "TetParams", "HexParams", "FemTetParams", "Stencil27",
    "F64", "F32", "F32X2", "F32D", "F32P",
    "PRECOND_NONE", "PRECOND_JACOBI", "PRECOND_BLOCK_JACOBI", "PRECOND_LEARNED",
    "INDEX_FORCE_DIAG", "LIB_PATH", "HEADER_PATH",
]

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libsimit_hip.so"
HEADER_PATH = PKG.parent / "include" / "simit_hip.h"

OK, ERR_INVALID, ERR_HIP, ERR_UNSUPPORTED, ERR_NOMEM, ERR_BREAKDOWN, ERR_NOT_CONVERGED = range(7)
F64, F32, F32X2, F32D, F32P = 0, 1, 2, 3, 4  # F32D: float values, double vectors
# F32P: FP32 3x3 values in the lane-planar layout (smh_bsr_planar), float vectors
PRECOND_NONE, PRECOND_JACOBI, PRECOND_BLOCK_JACOBI, PRECOND_LEARNED = 0, 1, 2, 3
STENCIL_PERIODIC = 1
INDEX_FORCE_DIAG = 1

_STATUS = {0: "OK", 1: "INVALID", 2: "HIP", 3: "UNSUPPORTED", 4: "NOMEM",
           5: "BREAKDOWN", 6: "NOT_CONVERGED"}


class SimitError(RuntimeError):
    """A non-zero status from the C ABI (the C++ layer throws SimitException)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{_STATUS.get(code, code)}] {message}")
        self.code = code


# ---- ABI structs (must match include/simit_hip.h) -----------------------------
class IndexView(C.Structure):
    _fields_ = [("nrows", C.c_uint32), ("nnz", C.c_uint64),
                ("rowptr", C.c_void_p), ("colidx", C.c_void_p)]


class Bsr(C.Structure):
    _fields_ = [("nrows", C.c_uint32), ("bdim", C.c_uint32), ("nnz", C.c_uint64),
                ("rowptr", C.c_void_p), ("colidx", C.c_void_p), ("precision", C.c_int32),
                ("vals", C.c_void_p), ("vals_lo", C.c_void_p)]


class Stencil27(C.Structure):
    """Scalar 27-point stencil matrix of an n0 x n1 x n2 grid of rows:
    vals[s*n + i] couples row i with column i + off_s (ascending offsets,
    s = 13 the diagonal); flags STENCIL_PERIODIC wraps the neighbours around."""
    _fields_ = [("n0", C.c_uint32), ("n1", C.c_uint32), ("n2", C.c_uint32),
                ("precision", C.c_int32), ("vals", C.c_void_p), ("vals_lo", C.c_void_p),
                ("flags", C.c_uint32)]


class SpringsParams(C.Structure):
    _fields_ = [("h", C.c_double), ("viscous", C.c_double), ("grav", C.c_double * 3),
                ("fixed_fill", C.c_double), ("fixed_identity", C.c_int32)]


class SpringsSet(C.Structure):
    _fields_ = [("npoints", C.c_uint32), ("nsprings", C.c_uint64), ("endpoints", C.c_void_p),
                ("x", C.c_void_p), ("v", C.c_void_p), ("m", C.c_void_p),
                ("fixed", C.c_void_p), ("k", C.c_void_p), ("l0", C.c_void_p)]


class TetParams(C.Structure):
    _fields_ = [("stiff_scale", C.c_double), ("rho", C.c_double), ("load", C.c_double * 3),
                ("fixed_diag", C.c_double)]


class TetSet(C.Structure):
    _fields_ = [("nverts", C.c_uint32), ("ntets", C.c_uint64), ("endpoints", C.c_void_p),
                ("x", C.c_void_p), ("c", C.c_void_p), ("mu", C.c_void_p), ("lam", C.c_void_p)]


class FemTetParams(C.Structure):
    _fields_ = [("h", C.c_double), ("rho", C.c_double), ("grav", C.c_double * 3)]


class FemTetSet(C.Structure):
    _fields_ = [("nverts", C.c_uint32), ("ntets", C.c_uint64), ("endpoints", C.c_void_p),
                ("x", C.c_void_p), ("v", C.c_void_p), ("c", C.c_void_p), ("u", C.c_void_p),
                ("l", C.c_void_p), ("W", C.c_void_p), ("B", C.c_void_p)]


class HexParams(C.Structure):
    _fields_ = [("h", C.c_double), ("source", C.c_double), ("fixed_diag", C.c_double)]


class HexSet(C.Structure):
    _fields_ = [("nverts", C.c_uint32), ("nhex", C.c_uint64), ("endpoints", C.c_void_p),
                ("dirichlet", C.c_void_p), ("kappa", C.c_void_p)]


class EdgeBlocks(C.Structure):
    _fields_ = [("nverts", C.c_uint32), ("nedges", C.c_uint64), ("bdim", C.c_uint32),
                ("endpoints", C.c_void_p), ("w", C.c_void_p), ("sign", C.c_double * 4),
                ("alpha", C.c_double), ("beta", C.c_double)]


class CGOptions(C.Structure):
    _fields_ = [("maxiters", C.c_int32), ("tol", C.c_double), ("rel", C.c_int32),
                ("norm_form", C.c_int32), ("replace_every", C.c_int32),
                ("precond", C.c_int32), ("mixed", C.c_int32),
                ("reliable_delta", C.c_double), ("check_every", C.c_int32),
                ("learned", C.c_void_p), ("zr_stop", C.c_int32),
                ("learned_corr", C.c_void_p), ("fallback", C.c_int32),
                ("vals_f64", C.c_void_p)]


class CGBatchStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("status", C.c_int32), ("true_rel", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CGStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("reliable_updates", C.c_int32), ("status", C.c_int32),
                ("rr", C.c_double), ("true_rel", C.c_double), ("seconds", C.c_double),
                ("precond_used", C.c_int32), ("fallbacks", C.c_int32),
                ("first_status", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load(path: os.PathLike | None = None) -> C.CDLL:
    """Load libsimit_hip.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise SimitError(ERR_UNSUPPORTED,
                         f"{p} is not built; run simit_amd.build or __graft_entry__.build()")
    lib = C.CDLL(str(p))
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32
    sig = {
        "smh_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "smh_ctx_destroy": (C.c_int, [vp]),
        "smh_ctx_last_error": (C.c_char_p, [vp]),
        "smh_ctx_set_stream": (C.c_int, [vp, vp]),
        "smh_ctx_get_stream": (vp, [vp]),
        "smh_ctx_synchronize": (C.c_int, [vp]),
        "smh_ctx_describe": (C.c_int, [vp, C.c_char_p, C.c_size_t]),
        "smh_malloc": (C.c_int, [vp, C.c_size_t, C.POINTER(vp)]),
        "smh_free": (C.c_int, [vp, vp]),
        "smh_memcpy_htod": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "smh_memcpy_dtoh": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "smh_memset": (C.c_int, [vp, vp, C.c_int, C.c_size_t]),
        "smh_index_build": (C.c_int, [vp, u32, u32, C.POINTER(u64), C.POINTER(u32),
                                      C.POINTER(vp), u32, C.POINTER(vp)]),
        "smh_index_get": (C.c_int, [vp, C.POINTER(IndexView)]),
        "smh_index_get_ve": (C.c_int, [vp, u32, C.POINTER(vp), C.POINTER(vp), C.POINTER(u64)]),
        "smh_index_destroy": (C.c_int, [vp, vp]),
        "smh_index_locs": (C.c_int, [vp, vp, u32, vp]),
        "smh_assemble_springs": (C.c_int, [vp, vp, C.POINTER(SpringsParams),
                                           C.POINTER(SpringsSet), i32, vp, vp, vp]),
        "smh_assemble_tet_linear": (C.c_int, [vp, vp, C.POINTER(TetParams), C.POINTER(TetSet),
                                              i32, vp, vp, vp]),
        "smh_assemble_hex_poisson": (C.c_int, [vp, vp, C.POINTER(HexParams), C.POINTER(HexSet),
                                               i32, vp, vp, vp]),
        "smh_assemble_edge_blocks": (C.c_int, [vp, vp, C.POINTER(EdgeBlocks), i32, vp, vp]),
        "smh_assemble_tet_neohookean": (C.c_int, [vp, vp, C.POINTER(FemTetParams),
                                                  C.POINTER(FemTetSet), i32, vp, vp, vp]),
        "smh_femtet_precompute": (C.c_int, [vp, vp, C.POINTER(FemTetParams),
                                            C.POINTER(FemTetSet), vp, vp, vp]),
        "smh_bsrmv": (C.c_int, [vp, C.POINTER(Bsr), vp, vp, i32]),
        "smh_bsr_planar": (C.c_int, [vp, C.POINTER(Bsr), vp]),
        "smh_dot": (C.c_int, [vp, i32, u64, vp, vp, C.POINTER(C.c_double)]),
        "smh_stencil27_from_bsr": (C.c_int, [vp, C.POINTER(Bsr), u32, u32, u32, vp, vp]),
        "smh_stencil27_correction": (C.c_int, [vp, C.POINTER(Bsr), u32, u32, u32, vp]),
        "smh_stencil27_mv": (C.c_int, [vp, C.POINTER(Stencil27), vp, vp]),
        "smh_cg_solve_stencil27": (C.c_int, [vp, C.POINTER(Stencil27), vp, vp,
                                             C.POINTER(CGOptions), C.POINTER(CGStats)]),
        "smh_block_jacobi": (C.c_int, [vp, C.POINTER(Bsr), vp]),
        "smh_axpby": (C.c_int, [vp, u64, C.c_double, vp, C.c_double, vp]),
        "smh_split_hilo": (C.c_int, [vp, u64, vp, vp, vp]),
        "smh_vec_op": (C.c_int, [vp, i32, u64, C.c_double, vp, C.c_double, vp, vp]),
        "smh_stage_reduce": (C.c_int, [vp, u64, u32, vp, vp, vp, vp, vp]),
        "smh_hiprtc_compile": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(vp),
                                         C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t]),
        "smh_image_free": (None, [vp]),
        "smh_module_load": (C.c_int, [vp, vp, C.c_size_t, C.POINTER(vp)]),
        "smh_module_launch": (C.c_int, [vp, vp, C.c_char_p, u64, vp, C.c_size_t]),
        "smh_module_destroy": (C.c_int, [vp, vp]),
        "smh_bsr_forget": (C.c_int, [vp, vp]),
        "smh_cg_default_opts": (None, [C.POINTER(CGOptions)]),
        "smh_cg_solve": (C.c_int, [vp, C.POINTER(Bsr), vp, vp, C.POINTER(CGOptions),
                                   C.POINTER(CGStats)]),
        "smh_cg_solve_noalloc": (C.c_int, [vp, C.POINTER(Bsr), vp, vp, C.POINTER(CGOptions),
                                           C.POINTER(CGStats)]),
        "smh_cg_solve_batched": (C.c_int, [vp, C.POINTER(Bsr), u32, C.POINTER(u32), vp, vp,
                                           C.POINTER(CGOptions), C.POINTER(CGBatchStats),
                                           C.POINTER(C.c_double)]),
        "smh_springs_solve_batched": (C.c_int, [vp, vp, C.POINTER(SpringsParams),
                                                C.POINTER(SpringsSet), vp, vp, vp, u32,
                                                C.POINTER(u32), vp, C.POINTER(CGOptions),
                                                C.POINTER(CGBatchStats),
                                                C.POINTER(C.c_double)]),
        "smh_gnn_create": (C.c_int, [vp, u64, C.c_double, C.POINTER(vp)]),
        "smh_gnn_destroy": (C.c_int, [vp, vp]),
        "smh_gnn_get_params": (C.c_int, [vp, vp, vp]),
        "smh_gnn_set_params": (C.c_int, [vp, vp, vp]),
        "smh_gnn_precond": (C.c_int, [vp, vp, C.POINTER(Bsr), vp, vp]),
        "smh_gnn_precond_info": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                           C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
        "smh_gnn_loss_grad": (C.c_int, [vp, vp, C.POINTER(Bsr), i32, vp, u64,
                                        C.POINTER(C.c_double), vp]),
        "smh_gnn_train_step": (C.c_int, [vp, vp, C.POINTER(Bsr), i32, u64, C.c_double,
                                         C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ptr(t) -> int:
    """Raw device pointer of a torch tensor, an int, or None."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


class Context:
    """One smh_ctx on one device, ordered on torch's current stream."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.smh_ctx_create(device, C.byref(h))
        if rc:
            raise SimitError(rc, f"smh_ctx_create({device}) failed")
        self.h = h
        self.device = device
        if use_torch_stream:
            import torch
            s = torch.cuda.current_stream(device).cuda_stream
            self._check(self.lib.smh_ctx_set_stream(self.h, C.c_void_p(s)))

    def _check(self, rc: int):
        if rc:
            msg = self.lib.smh_ctx_last_error(self.h).decode(errors="replace")
            raise SimitError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.smh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self._check(self.lib.smh_ctx_synchronize(self.h))

    def describe(self) -> str:
        buf = C.create_string_buffer(512)
        self._check(self.lib.smh_ctx_describe(self.h, buf, len(buf)))
        return buf.value.decode()

    # ---- sparsity -----------------------------------------------------------
    def build_index(self, nverts: int, endpoint_sets, force_diag: bool = False) -> "Index":
        """endpoint_sets: list of int32 device tensors of shape [ne, card]."""
        ns = len(endpoint_sets)
        ne = (C.c_uint64 * ns)(*[int(e.shape[0]) for e in endpoint_sets])
        card = (C.c_uint32 * ns)(*[int(e.shape[1]) if e.dim() == 2 else 1 for e in endpoint_sets])
        eps = (C.c_void_p * ns)(*[_ptr(e) for e in endpoint_sets])
        out = C.c_void_p()
        self._check(self.lib.smh_index_build(self.h, nverts, ns, ne, card, eps,
                                             INDEX_FORCE_DIAG if force_diag else 0, C.byref(out)))
        return Index(self, out, list(endpoint_sets))

    # ---- assembly -----------------------------------------------------------
    def assemble_springs(self, idx, params: SpringsParams, pts: dict, springs: dict,
                         precision: int, vals, vals_lo, rhs):
        s = SpringsSet(idx.nrows, int(springs["endpoints"].shape[0]), _ptr(springs["endpoints"]),
                       _ptr(pts["x"]), _ptr(pts["v"]), _ptr(pts["m"]), _ptr(pts["fixed"]),
                       _ptr(springs["k"]), _ptr(springs["l0"]))
        self._check(self.lib.smh_assemble_springs(self.h, idx.h, C.byref(params), C.byref(s),
                                                  precision, _ptr(vals), _ptr(vals_lo), _ptr(rhs)))

    def assemble_tets(self, idx, params: TetParams, verts: dict, tets: dict, precision: int,
                      vals, vals_lo, rhs):
        s = TetSet(idx.nrows, int(tets["endpoints"].shape[0]), _ptr(tets["endpoints"]),
                   _ptr(verts["x"]), _ptr(verts["c"]), _ptr(tets["mu"]), _ptr(tets["lam"]))
        self._check(self.lib.smh_assemble_tet_linear(self.h, idx.h, C.byref(params), C.byref(s),
                                                     precision, _ptr(vals), _ptr(vals_lo),
                                                     _ptr(rhs)))

    def assemble_hex(self, idx, params: HexParams, verts: dict, hexes: dict, precision: int,
                     vals, vals_lo, rhs):
        s = HexSet(idx.nrows, int(hexes["endpoints"].shape[0]), _ptr(hexes["endpoints"]),
                   _ptr(verts["dirichlet"]), _ptr(hexes["kappa"]))
        self._check(self.lib.smh_assemble_hex_poisson(self.h, idx.h, C.byref(params), C.byref(s),
                                                      precision, _ptr(vals), _ptr(vals_lo),
                                                      _ptr(rhs)))

    def _femtet_set(self, idx, verts: dict, tets: dict):
        return FemTetSet(idx.nrows, int(tets["endpoints"].shape[0]), _ptr(tets["endpoints"]),
                         _ptr(verts["x"]), _ptr(verts.get("v")), _ptr(verts.get("c")),
                         _ptr(tets.get("u")), _ptr(tets.get("l")), _ptr(tets.get("W")),
                         _ptr(tets.get("B")))

    def assemble_femtet(self, idx, params: FemTetParams, verts: dict, tets: dict,
                        precision: int, vals, vals_lo, rhs):
        """femTet.sim main's two maps: A (stiffness + lumped mass) and f."""
        s = self._femtet_set(idx, verts, tets)
        self._check(self.lib.smh_assemble_tet_neohookean(self.h, idx.h, C.byref(params),
                                                         C.byref(s), precision, _ptr(vals),
                                                         _ptr(vals_lo), _ptr(rhs)))

    def femtet_precompute(self, idx, params: FemTetParams, verts: dict, tets: dict, W, B, m):
        """femTet.sim initializeTet: W, B per tet and the vertex masses m."""
        s = self._femtet_set(idx, verts, tets)
        self._check(self.lib.smh_femtet_precompute(self.h, idx.h, C.byref(params), C.byref(s),
                                                   _ptr(W), _ptr(B), _ptr(m)))

    def assemble_edge_blocks(self, idx, endpoints, w, bdim: int, sign=(1, 1, 1, 1),
                             alpha: float = 0.0, beta: float = 1.0, precision: int = F64,
                             vals=None, vals_lo=None):
        e = EdgeBlocks(idx.nrows, int(endpoints.shape[0]), bdim, _ptr(endpoints), _ptr(w),
                       (C.c_double * 4)(*sign), alpha, beta)
        self._check(self.lib.smh_assemble_edge_blocks(self.h, idx.h, C.byref(e), precision,
                                                      _ptr(vals), _ptr(vals_lo)))

    # ---- linear algebra -----------------------------------------------------
    def bsrmv(self, A: Bsr, x, y, exact_order: bool = False):
        self._check(self.lib.smh_bsrmv(self.h, C.byref(A), _ptr(x), _ptr(y), int(exact_order)))

    def split_hilo(self, v, hi, lo):
        """hi = fl32(v), lo = fl32(v - hi) for FP64 values v (the SMH_F32X2 pair
        the assemblies would write)."""
        self._check(self.lib.smh_split_hilo(self.h, v.numel(), _ptr(v), _ptr(hi), _ptr(lo)))

    def bsr_planar(self, A: Bsr, out=None):
        """The lane-planar copy of A's FP32 3x3 values (a float32 tensor of
        576 * ceil(nnz / 64) entries) and the SMH_F32P matrix over it."""
        import torch
        n = ((A.nnz + 63) // 64) * 576
        if out is None:
            out = torch.empty(max(n, 1), dtype=torch.float32, device="cuda")
        self._check(self.lib.smh_bsr_planar(self.h, C.byref(A), _ptr(out)))
        P = Bsr(A.nrows, A.bdim, A.nnz, A.rowptr, A.colidx, F32P, _ptr(out), None)
        P._keep = (out, A)
        return P

    def dot(self, x, y, precision: int | None = None) -> float:
        import torch
        if precision is None:
            precision = F64 if x.dtype == torch.float64 else F32
        out = C.c_double()
        self._check(self.lib.smh_dot(self.h, precision, x.numel(), _ptr(x), _ptr(y),
                                     C.byref(out)))
        return out.value

    def axpby(self, a: float, x, b: float, y):
        """y = a*x + b*y (double vectors)."""
        self._check(self.lib.smh_axpby(self.h, y.numel(), a, _ptr(x), b, _ptr(y)))

    def block_jacobi(self, A: Bsr, dinv):
        self._check(self.lib.smh_block_jacobi(self.h, C.byref(A), _ptr(dinv)))

    def cg(self, A: Bsr, b, x, opts: CGOptions) -> CGStats:
        st = CGStats()
        self._check(self.lib.smh_cg_solve(self.h, C.byref(A), _ptr(b), _ptr(x), C.byref(opts),
                                          C.byref(st)))
        return st

    # ---- 27-point stencil storage ---------------------------------------------
    def stencil27_from_bsr(self, A: Bsr, dims, vals=None, vals_lo=None) -> Stencil27:
        """Scatter a 1x1 matrix into stencil storage on the (n0, n1, n2) grid;
        allocates the [27*n] value arrays unless given."""
        import torch
        n0, n1, n2 = (int(d) for d in dims)
        n = n0 * n1 * n2
        if vals is None:
            dt = torch.float64 if A.precision == F64 else torch.float32
            vals = torch.empty(27 * n, dtype=dt, device="cuda")
        if vals_lo is None and A.precision == F32X2:
            vals_lo = torch.empty(27 * n, dtype=torch.float32, device="cuda")
        self._check(self.lib.smh_stencil27_from_bsr(self.h, C.byref(A), n0, n1, n2, _ptr(vals),
                                                    _ptr(vals_lo)))
        S = Stencil27(n0, n1, n2, A.precision, _ptr(vals), _ptr(vals_lo))
        S._keep = (vals, vals_lo)
        return S

    def stencil27_correction(self, idx, corr, dims, out=None):
        """The float[nnz] learned correction of Gnn.precond (1x1, A's pattern
        idx) as FP16 stencil slots (uint16[27*n]) for cg_stencil27."""
        import torch
        n0, n1, n2 = (int(d) for d in dims)
        if out is None:
            out = torch.empty(27 * n0 * n1 * n2, dtype=torch.int16, device="cuda")
        Cb = idx.bsr(1, F32, corr.view(torch.float32))
        self._check(self.lib.smh_stencil27_correction(self.h, C.byref(Cb), n0, n1, n2,
                                                      _ptr(out)))
        return out

    def stencil27_mv(self, S: Stencil27, x, y):
        self._check(self.lib.smh_stencil27_mv(self.h, C.byref(S), _ptr(x), _ptr(y)))

    def cg_stencil27(self, S: Stencil27, b, x, opts: CGOptions) -> CGStats:
        st = CGStats()
        self._check(self.lib.smh_cg_solve_stencil27(self.h, C.byref(S), _ptr(b), _ptr(x),
                                                    C.byref(opts), C.byref(st)))
        return st

    def cg_batched(self, A: Bsr, rows, b, x, opts: CGOptions):
        """Solve the independent diagonal-block systems rows[m]..rows[m+1];
        returns (per-system stats dicts, seconds)."""
        import numpy as np
        r = np.ascontiguousarray(rows, dtype=np.uint32)
        nb = r.size - 1
        st = (CGBatchStats * max(nb, 1))()
        sec = C.c_double()
        self._check(self.lib.smh_cg_solve_batched(
            self.h, C.byref(A), nb, r.ctypes.data_as(C.POINTER(C.c_uint32)), _ptr(b), _ptr(x),
            C.byref(opts), st, C.byref(sec)))
        return [st[m].as_dict() for m in range(nb)], sec.value

    def springs_solve_batched(self, idx, params: SpringsParams, pts: dict, springs: dict,
                              vals, vals_lo, rhs, rows, x, opts: CGOptions):
        """Assembly of every lattice and the batched CG, in captured graphs;
        returns (per-system stats dicts, seconds)."""
        import numpy as np
        s = SpringsSet(idx.nrows, int(springs["endpoints"].shape[0]), _ptr(springs["endpoints"]),
                       _ptr(pts["x"]), _ptr(pts["v"]), _ptr(pts["m"]), _ptr(pts["fixed"]),
                       _ptr(springs["k"]), _ptr(springs["l0"]))
        r = np.ascontiguousarray(rows, dtype=np.uint32)
        nb = r.size - 1
        st = (CGBatchStats * max(nb, 1))()
        sec = C.c_double()
        self._check(self.lib.smh_springs_solve_batched(
            self.h, idx.h, C.byref(params), C.byref(s), _ptr(vals), _ptr(vals_lo), _ptr(rhs), nb,
            r.ctypes.data_as(C.POINTER(C.c_uint32)), _ptr(x), C.byref(opts), st, C.byref(sec)))
        return [st[m].as_dict() for m in range(nb)], sec.value

    def cg_options(self, **kw) -> CGOptions:
        o = CGOptions()
        self.lib.smh_cg_default_opts(C.byref(o))
        keep = []
        for k, v in kw.items():
            if k in ("learned", "learned_corr", "vals_f64"):
                keep.append(v)  # the struct holds a raw pointer: keep the buffer alive
                v = _ptr(v)
            setattr(o, k, v)
        o._keep = tuple(keep)
        return o

    def bsr_forget(self, A):
        """Drop the cached row-tile plan of A's pattern (call before its
        rowptr buffer is freed or refilled with another pattern)."""
        rp = A.rowptr if isinstance(A, Bsr) else _ptr(A)
        self._check(self.lib.smh_bsr_forget(self.h, rp))

    def gnn(self, seed: int = 0, kappa: float = 0.9) -> "Gnn":
        return Gnn(self, seed, kappa)


GNN_NPARAMS = 18192


class Gnn:
    """The learned block-Jacobi correction network (smh_gnn)."""

    def __init__(self, ctx: Context, seed: int = 0, kappa: float = 0.9):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._check(ctx.lib.smh_gnn_create(ctx.h, seed, kappa, C.byref(h)))
        self.h = h
        self.kappa = kappa

    def get_params(self):
        import numpy as np
        out = np.empty(GNN_NPARAMS, dtype=np.float32)
        self.ctx._check(self.ctx.lib.smh_gnn_get_params(self.ctx.h, self.h,
                                                       out.ctypes.data_as(C.c_void_p)))
        return out

    def set_params(self, p):
        import numpy as np
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == GNN_NPARAMS
        self.ctx._check(self.ctx.lib.smh_gnn_set_params(self.ctx.h, self.h,
                                                       a.ctypes.data_as(C.c_void_p)))

    def precond(self, A: Bsr):
        """(linv, corr) device tensors: 3x3 blocks: float32[nrows*9] and uint8[nnz*12]
        (FP8 blocks); 1x1 blocks: float32[nrows] and uint8[nnz*4] (float32 values)."""
        import torch
        bb = A.bdim * A.bdim
        linv = torch.empty(A.nrows * bb, dtype=torch.float32, device="cuda")
        corr = torch.empty(A.nnz * (12 if A.bdim == 3 else 4), dtype=torch.uint8, device="cuda")
        self.ctx._check(self.ctx.lib.smh_gnn_precond(self.ctx.h, self.h, C.byref(A),
                                                    _ptr(linv), _ptr(corr)))
        return linv, corr

    def precond_info(self) -> dict:
        """The spectral guard of the last precond(): Lanczos estimate of
        lambda_min and the spectral width of the network's C, the scale
        applied to C and the Lanczos steps."""
        lm, w, sc, st = C.c_double(), C.c_double(), C.c_double(), C.c_int32()
        self.ctx._check(self.ctx.lib.smh_gnn_precond_info(self.h, C.byref(lm), C.byref(w),
                                                         C.byref(sc), C.byref(st)))
        return {"lambda_min": lm.value, "spectral_width": w.value, "scale": sc.value,
                "steps": st.value}

    def loss_grad(self, A: Bsr, nprobes: int = 4, probes=None, seed: int = 0, grads=True):
        import numpy as np
        loss = C.c_double()
        g = np.empty(GNN_NPARAMS, dtype=np.float32) if grads else None
        self.ctx._check(self.ctx.lib.smh_gnn_loss_grad(
            self.ctx.h, self.h, C.byref(A), nprobes, _ptr(probes), seed, C.byref(loss),
            g.ctypes.data_as(C.c_void_p) if g is not None else None))
        return loss.value, g

    def train_step(self, A: Bsr, nprobes: int = 4, seed: int = 0, lr: float = 1e-3,
                   want_loss: bool = True):
        loss = C.c_double()
        self.ctx._check(self.ctx.lib.smh_gnn_train_step(
            self.ctx.h, self.h, C.byref(A), nprobes, seed, lr,
            C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.smh_gnn_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Index:
    """A device-built vev index (smh_index)."""

    def __init__(self, ctx: Context, h, endpoint_sets):
        self.ctx = ctx
        self.h = h
        self._eps = endpoint_sets  # keep the borrowed endpoint buffers alive
        v = IndexView()
        ctx._check(ctx.lib.smh_index_get(h, C.byref(v)))
        self.nrows, self.nnz = v.nrows, v.nnz
        self.rowptr_ptr, self.colidx_ptr = v.rowptr, v.colidx

    def to_host(self):
        """(rowptr, colidx) as numpy uint32 arrays (synchronises)."""
        import numpy as np
        rp = np.empty(self.nrows + 1, dtype=np.uint32)
        ci = np.empty(max(self.nnz, 1), dtype=np.uint32)
        lib, h = self.ctx.lib, self.ctx.h
        self.ctx._check(lib.smh_memcpy_dtoh(h, rp.ctypes.data, self.rowptr_ptr, rp.nbytes))
        if self.nnz:
            self.ctx._check(lib.smh_memcpy_dtoh(h, ci.ctypes.data, self.colidx_ptr, 4 * self.nnz))
        return rp, ci[: self.nnz]

    def ve_to_host(self, s: int = 0):
        import numpy as np
        p, e, n = C.c_void_p(), C.c_void_p(), C.c_uint64()
        lib, h = self.ctx.lib, self.ctx.h
        self.ctx._check(lib.smh_index_get_ve(self.h, s, C.byref(p), C.byref(e), C.byref(n)))
        ptr = np.empty(self.nrows + 1, dtype=np.uint32)
        el = np.empty(max(n.value, 1), dtype=np.uint32)
        self.ctx._check(lib.smh_memcpy_dtoh(h, ptr.ctypes.data, p, ptr.nbytes))
        if n.value:
            self.ctx._check(lib.smh_memcpy_dtoh(h, el.ctypes.data, e, 4 * n.value))
        return ptr, el[: n.value]

    def locs(self, s, out):
        self.ctx._check(self.ctx.lib.smh_index_locs(self.ctx.h, self.h, s, _ptr(out)))

    def bsr(self, bdim: int, precision: int, vals, vals_lo=None) -> Bsr:
        b = Bsr(self.nrows, bdim, self.nnz, self.rowptr_ptr, self.colidx_ptr, precision,
                _ptr(vals), _ptr(vals_lo))
        # the struct holds raw pointers: keep the buffers (and the index)
        # alive for as long as it is
        b._keep = (vals, vals_lo, self)
        return b

    def close(self):
        if self.h:
            self.ctx.lib.smh_index_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
