"""smh_cg_solve_noalloc: the allocation-free variant of smh_cg_solve.

Called through the C ABI and checked against smh_cg_solve bit for bit (x,
iterations, statuses), on two calls in a row, the second on a smaller
system, so the shared scratch array's leftovers from the first call cannot
leak into the second.  Also: a system beyond the scratch array's maximum is
rejected before any work is queued, and the learned preconditioner is
refused.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def noalloc(ctx, A, b, x, opts):
    st = S.CGStats()
    rc = ctx.lib.smh_cg_solve_noalloc(ctx.h, C.byref(A), S._ptr(b), S._ptr(x), C.byref(opts),
                                      C.byref(st))
    return rc, st


def tet_system(ctx, n, seed):
    prob = meshes.tet_problem(n, seed=seed, field="grf", std_log10=1.0, corr_len=2.0)
    d = {k: to_dev(prob[k]) for k in ("x", "c", "mu", "lam", "endpoints")}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    P = S.TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    v64 = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(3 * prob["nverts"], dtype=torch.float64, device="cuda")
    ctx.assemble_tets(idx, P, d, d, S.F32X2, hi, lo, rhs)
    ctx.assemble_tets(idx, P, d, d, S.F64, v64, None, rhs)
    return idx, hi, lo, v64, rhs


CASES = [  # (mesh n, seed, precision, cg options)
    (14, 1, "f32x2", dict(precond=S.PRECOND_BLOCK_JACOBI, mixed=1)),
    (9, 2, "f32x2", dict(precond=S.PRECOND_BLOCK_JACOBI, mixed=1)),
    (12, 3, "f64", dict(precond=S.PRECOND_JACOBI)),
    (6, 4, "f64", dict(precond=S.PRECOND_NONE)),
]


def test_noalloc_matches_cg_solve_on_consecutive_calls(ctx):
    for n, seed, prec, kw in CASES:  # each call smaller than the one before
        print("case", n, seed, prec, kw, flush=True)
        idx, hi, lo, v64, rhs = tet_system(ctx, n, seed)
        A = idx.bsr(3, S.F32X2, hi, lo) if prec == "f32x2" else idx.bsr(3, S.F64, v64)
        o = ctx.cg_options(maxiters=5000, tol=1e-8, rel=1, **kw)
        x_ref = torch.zeros_like(rhs)
        st_ref = ctx.cg(A, rhs, x_ref, o)
        x = torch.zeros_like(rhs)
        rc, st = noalloc(ctx, A, rhs, x, o)
        torch.cuda.synchronize()
        assert rc == 0, ctx.lib.smh_ctx_last_error(ctx.h)
        assert st_ref.status == 0 and st_ref.true_rel <= 1e-8, st_ref.as_dict()
        assert torch.equal(x, x_ref), (n, seed)
        for k in ("iters", "reliable_updates", "status", "rr", "true_rel", "precond_used",
                  "fallbacks", "first_status"):
            assert getattr(st, k) == getattr(st_ref, k), (k, st.as_dict(), st_ref.as_dict())


def test_noalloc_fallback_chain_matches(ctx):
    """a zero diagonal: Jacobi breaks down, plain CG finishes -- the chain's
    copy of the initial guess lives in the scratch array too"""
    ep = np.array([[0, 1], [1, 2], [2, 3]], np.int32)
    idx = ctx.build_index(4, [to_dev(ep)])
    rp, ci = idx.to_host()
    diag = [2.0, 0.0, 3.0, 4.0]
    vals = to_dev(np.array([diag[i] if ci[l] == i else 0.0
                            for i in range(4) for l in range(rp[i], rp[i + 1])]))
    b = to_dev(np.array([1.0, 0.0, 1.0, 1.0]))
    A = idx.bsr(1, S.F64, vals)
    o = ctx.cg_options(maxiters=20, tol=1e-12, precond=S.PRECOND_JACOBI)
    x_ref = torch.full((4,), 0.25, dtype=torch.float64, device="cuda")
    st_ref = ctx.cg(A, b, x_ref, o)
    x = torch.full((4,), 0.25, dtype=torch.float64, device="cuda")
    rc, st = noalloc(ctx, A, b, x, o)
    assert rc == 0
    assert torch.equal(x, x_ref)
    assert st.fallbacks == st_ref.fallbacks == 1 and st.precond_used == S.PRECOND_NONE


def test_noalloc_rejects_what_does_not_fit(ctx):
    # 60^3 x 5 tets: ~227k rows of 3x3 blocks in FP64 with block Jacobi
    prob = meshes.tet_problem(60, seed=0)
    d = to_dev(prob["endpoints"])
    idx = ctx.build_index(prob["nverts"], [d])
    v = torch.zeros(idx.nnz * 9, dtype=torch.float64, device="cuda")
    b = torch.ones(3 * prob["nverts"], dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    rc, _ = noalloc(ctx, idx.bsr(3, S.F64, v), b, x,
                    ctx.cg_options(precond=S.PRECOND_BLOCK_JACOBI))
    assert rc == S.ERR_UNSUPPORTED
    assert b"maximum" in ctx.lib.smh_ctx_last_error(ctx.h)
    assert torch.count_nonzero(x).item() == 0  # nothing ran
    # the learned preconditioner is refused
    small = ctx.build_index(4, [to_dev(np.array([[0, 1], [2, 3]], np.int32))])
    vv = torch.ones(small.nnz * 9, dtype=torch.float64, device="cuda")
    bb = torch.ones(12, dtype=torch.float64, device="cuda")
    xx = torch.zeros_like(bb)
    linv = torch.ones(36, dtype=torch.float32, device="cuda")
    corr = torch.zeros(small.nnz * 12, dtype=torch.uint8, device="cuda")
    rc, _ = noalloc(ctx, small.bsr(3, S.F64, vv), bb, xx,
                    ctx.cg_options(precond=S.PRECOND_LEARNED, learned=linv, learned_corr=corr))
    assert rc == S.ERR_UNSUPPORTED
