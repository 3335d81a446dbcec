"""Mixed-precision solves certified against the FP64 matrix (smh_cg_opts.vals_f64).

hi + lo (SMH_F32X2) carries each FP64 value only to ~2^-48.  On the stiff
scalar hex system (log10 conductivity with std 1.5) that difference alone
moves ||b - A x|| / ||b|| of a 1e-8 solve past 1e-8 against the FP64 matrix
(96^3: 9.8e-9 on hi + lo, 3.8e-8 on the FP64 values).  With the FP64 values
given, the reliable updates and the certification use them, and the result
meets 1e-8 on the oracle's FP64 matrix and load."""
import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def hex_f64(ctx, n, seed=0):
    prob = meshes.hex_problem(n, seed=seed, std_log10=1.5, corr_len=12.0, source=1.0)
    d = {k: to_dev(prob[k]) for k in ("dirichlet", "kappa", "endpoints")}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    P = S.HexParams(prob["h"], prob["source"], 1.0)
    v64 = torch.empty(idx.nnz, dtype=torch.float64, device="cuda")
    rhs = torch.empty(prob["nverts"], dtype=torch.float64, device="cuda")
    ctx.assemble_hex(idx, P, d, d, S.F64, v64, None, rhs)
    return prob, idx, P, d, v64, rhs


def test_split_hilo_is_the_assembly_rounding(ctx):
    prob, idx, P, d, v64, rhs = hex_f64(ctx, 12)
    hi = torch.empty(idx.nnz, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    rhs2 = torch.empty_like(rhs)
    ctx.assemble_hex(idx, P, d, d, S.F32X2, hi, lo, rhs2)
    hs, ls = torch.full_like(hi, float("nan")), torch.full_like(lo, float("nan"))
    ctx.split_hilo(v64, hs, ls)
    assert torch.equal(hs, hi) and torch.equal(ls, lo) and torch.equal(rhs2, rhs)
    # lo is the rounding residual of hi: |v - hi - lo| <= 2^-24 |v - hi|
    v = v64.cpu().numpy()
    h, l = hs.cpu().numpy().astype(np.float64), ls.cpu().numpy().astype(np.float64)
    assert (np.abs(v - h - l) <= 2.0 ** -24 * np.abs(v - h)).all()


@pytest.mark.parametrize("mixed", [1, 2])
def test_hex_stencil_certified_on_fp64(ctx, oracle, mixed):
    n = 96
    prob, idx, P, d, v64, rhs = hex_f64(ctx, n)
    nv = prob["nverts"]
    hi = torch.empty(idx.nnz, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    ctx.split_hilo(v64, hi, lo)
    dims = (n + 1,) * 3
    st_hi = torch.empty(27 * nv, dtype=torch.float32, device="cuda")
    st_lo = torch.empty_like(st_hi)
    st64 = torch.empty(27 * nv, dtype=torch.float64, device="cuda")
    St = ctx.stencil27_from_bsr(idx.bsr(1, S.F32X2, hi, lo), dims, st_hi, st_lo)
    ctx.stencil27_from_bsr(idx.bsr(1, S.F64, v64), dims, st64)
    rp, ci = idx.to_host()
    ov, of = oracle.hex_assemble(oracle.HexParams(prob["h"], prob["source"], 1.0), prob, rp, ci)
    assert np.array_equal(v64.cpu().numpy(), ov) and np.array_equal(rhs.cpu().numpy(), of)
    bn = np.linalg.norm(of)
    x = torch.zeros(nv, dtype=torch.float64, device="cuda")
    st = ctx.cg_stencil27(St, rhs, x, ctx.cg_options(maxiters=40000, tol=1e-8, rel=1,
                                                     check_every=64, precond=S.PRECOND_JACOBI,
                                                     mixed=mixed, vals_f64=st64))
    rel = np.linalg.norm(of - oracle.bsrmv(rp, ci, ov, x.cpu().numpy(), 1)) / bn
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    # the certified residual is the FP64 matrix's: the oracle's agrees with it
    assert rel <= 1.01e-8 and abs(rel - st.true_rel) <= 1e-3 * st.true_rel, (rel, st.as_dict())


def test_tet_csr_certified_on_fp64(ctx, oracle):
    n = 24
    prob = meshes.tet_problem(n, seed=5, field="grf", std_log10=1.0, corr_len=4.0)
    dev = {k: to_dev(prob[k]) for k in ("x", "c", "mu", "lam", "endpoints")}
    nv = prob["nverts"]
    idx = ctx.build_index(nv, [dev["endpoints"]])
    import ctypes as C
    P = S.TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
    v64 = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(3 * nv, dtype=torch.float64, device="cuda")
    ctx.assemble_tets(idx, P, dev, dev, S.F64, v64, None, rhs)
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    ctx.split_hilo(v64, hi, lo)
    rp, ci = idx.to_host()
    ov, of = oracle.tet_assemble(oracle.tet_params(load=prob["load"]), prob, rp, ci)
    bn = np.linalg.norm(of)
    for mixed in (1, 2):
        x = torch.zeros_like(rhs)
        st = ctx.cg(idx.bsr(3, S.F32X2, hi, lo), rhs, x,
                    ctx.cg_options(maxiters=20000, tol=1e-8, rel=1, mixed=mixed,
                                   precond=S.PRECOND_BLOCK_JACOBI, vals_f64=v64))
        rel = np.linalg.norm(of - oracle.bsrmv(rp, ci, ov, x.cpu().numpy(), 3)) / bn
        assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
        assert rel <= 1.01e-8 and abs(rel - st.true_rel) <= 1e-3 * st.true_rel, (mixed, rel)


def test_vals_f64_needs_mixed_and_is_not_batched(ctx):
    prob, idx, P, d, v64, rhs = hex_f64(ctx, 4)
    x = torch.zeros_like(rhs)
    with pytest.raises(S.SimitError) as e:
        ctx.cg(idx.bsr(1, S.F64, v64), rhs, x, ctx.cg_options(vals_f64=v64))
    assert e.value.code == S.ERR_INVALID
    hi = torch.empty(idx.nnz, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    ctx.split_hilo(v64, hi, lo)
    rows = np.array([0, prob["nverts"]], dtype=np.uint32)
    with pytest.raises(S.SimitError) as e:
        ctx.cg_batched(idx.bsr(1, S.F32X2, hi, lo), rows, rhs, x,
                       ctx.cg_options(mixed=1, vals_f64=v64))
    assert e.value.code == S.ERR_UNSUPPORTED
