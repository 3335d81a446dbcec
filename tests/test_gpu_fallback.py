"""The solver's fallback chain: learned -> (block) Jacobi -> no preconditioner.

Every failure is forced from the host through the inputs (an indefinite
injected correction, a NaN in a field, a zero diagonal, an iteration budget
of one); every kernel runs to its normal end.  Each rung is certified by its
own FP64 residual, and the stats name the rung that produced x.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def hex_system(ctx, n=10):
    prob = meshes.hex_problem(n, seed=3, std_log10=1.0, corr_len=3.0, source=1.0)
    d = {k: to_dev(prob[k]) for k in ("dirichlet", "kappa", "endpoints")}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    hi = torch.empty(idx.nnz, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    rhs = torch.empty(prob["nverts"], dtype=torch.float64, device="cuda")
    ctx.assemble_hex(idx, S.HexParams(prob["h"], 1.0, 1.0), d, d, S.F32X2, hi, lo, rhs)
    return prob, idx, hi, lo, rhs


def indefinite_correction(ctx, idx, hi, lo):
    """linv = 1/sqrt(a_ii) and C = -4 on every off-diagonal entry: I + C is
    far from positive definite."""
    rp, ci = idx.to_host()
    d = (hi.double() + lo.double()).cpu().numpy()
    rows = np.repeat(np.arange(idx.nrows), np.diff(rp.astype(np.int64)))
    diag = np.zeros(idx.nrows)
    diag[rows[ci == rows]] = d[ci == rows]
    linv = to_dev((1.0 / np.sqrt(diag)).astype(np.float32))
    corr = np.where(ci == rows, 0.0, -4.0).astype(np.float32)
    return linv, to_dev(corr.view(np.uint8))


def test_indefinite_learned_correction_falls_back_to_jacobi(ctx):
    prob, idx, hi, lo, rhs = hex_system(ctx)
    linv, corr = indefinite_correction(ctx, idx, hi, lo)
    A = idx.bsr(1, S.F32X2, hi, lo)
    kw = dict(maxiters=3000, tol=1e-8, rel=1, precond=S.PRECOND_LEARNED, mixed=1,
              learned=linv, learned_corr=corr)
    # without the chain: a prompt breakdown (z'r <= 0), not 3000 iterations
    x = torch.zeros_like(rhs)
    st = ctx.cg(A, rhs, x, ctx.cg_options(fallback=0, **kw))
    assert st.status == S.ERR_BREAKDOWN and st.iters < 50, st.as_dict()
    assert st.fallbacks == 0 and st.precond_used == S.PRECOND_LEARNED
    # with it: Jacobi finishes, certified
    x = torch.zeros_like(rhs)
    st = ctx.cg(A, rhs, x, ctx.cg_options(**kw))
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    assert st.first_status == S.ERR_BREAKDOWN
    assert st.precond_used == S.PRECOND_JACOBI and st.fallbacks == 1
    # the same over 27-point stencil storage
    n1 = round(prob["nverts"] ** (1 / 3))
    St = ctx.stencil27_from_bsr(A, (n1, n1, n1))
    cst = ctx.stencil27_correction(idx, corr, (n1, n1, n1))
    x2 = torch.zeros_like(rhs)
    kw2 = dict(kw, learned_corr=cst, mixed=2)
    st2 = ctx.cg_stencil27(St, rhs, x2, ctx.cg_options(**kw2))
    assert st2.status == 0 and st2.true_rel <= 1e-8, st2.as_dict()
    assert st2.first_status == S.ERR_BREAKDOWN and st2.precond_used == S.PRECOND_JACOBI


def test_nan_in_a_field_ends_every_rung_in_breakdown(ctx):
    prob = meshes.tet_problem(5, seed=2)
    prob["mu"] = prob["mu"].copy()
    prob["mu"][len(prob["mu"]) // 2] = np.nan  # a tet in the middle of the block
    d = {k: to_dev(prob[k]) for k in ("x", "c", "mu", "lam", "endpoints")}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    v = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(3 * prob["nverts"], dtype=torch.float64, device="cuda")
    P = S.TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
    ctx.assemble_tets(idx, P, d, d, S.F64, v, None, rhs)
    x = torch.zeros_like(rhs)
    st = ctx.cg(idx.bsr(3, S.F64, v), rhs, x,
                ctx.cg_options(maxiters=500, tol=1e-8, precond=S.PRECOND_BLOCK_JACOBI))
    assert st.status == S.ERR_BREAKDOWN, st.as_dict()
    assert st.first_status == S.ERR_BREAKDOWN
    assert st.precond_used == S.PRECOND_NONE and st.fallbacks == 1


def test_zero_diagonal_jacobi_falls_back_to_plain_cg(ctx):
    """diag(2, 0, 3, 4) with b = (1, 0, 1, 1): Jacobi's 1/0 turns z into NaN;
    plain CG solves the consistent system in three steps."""
    ep = np.array([[0, 1], [1, 2], [2, 3]], np.int32)
    idx = ctx.build_index(4, [to_dev(ep)])
    rp, ci = idx.to_host()
    diag = [2.0, 0.0, 3.0, 4.0]
    vals = to_dev(np.array([diag[i] if ci[l] == i else 0.0
                            for i in range(4) for l in range(rp[i], rp[i + 1])]))
    b = to_dev(np.array([1.0, 0.0, 1.0, 1.0]))
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    st = ctx.cg(idx.bsr(1, S.F64, vals), b, x,
                ctx.cg_options(maxiters=20, tol=1e-12, precond=S.PRECOND_JACOBI))
    assert st.first_status == S.ERR_BREAKDOWN, st.as_dict()
    assert st.status == 0 and st.precond_used == S.PRECOND_NONE and st.fallbacks == 1
    np.testing.assert_allclose(x.cpu().numpy(), [0.5, 0.0, 1 / 3, 0.25], rtol=0, atol=1e-14)


def test_iteration_budget_is_not_a_breakdown(ctx):
    prob, idx, hi, lo, rhs = hex_system(ctx, 8)
    A = idx.bsr(1, S.F32X2, hi, lo)
    x = torch.zeros_like(rhs)
    st = ctx.cg(A, rhs, x, ctx.cg_options(maxiters=1, tol=1e-8, precond=S.PRECOND_JACOBI,
                                          mixed=1))
    assert st.status == S.ERR_NOT_CONVERGED and st.fallbacks == 0, st.as_dict()
    # an unconverged learned rung does fall back: to Jacobi, which keeps the status
    linv, corr = indefinite_correction(ctx, idx, hi, lo)
    x = torch.zeros_like(rhs)
    st = ctx.cg(A, rhs, x, ctx.cg_options(maxiters=1, tol=1e-8, precond=S.PRECOND_LEARNED,
                                          mixed=1, learned=linv, learned_corr=corr))
    assert st.status == S.ERR_NOT_CONVERGED and st.fallbacks == 1, st.as_dict()
    assert st.precond_used == S.PRECOND_JACOBI and st.iters <= 2
