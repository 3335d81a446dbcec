"""The benchmark configurations as parity cases at their real sizes.

C1  springs 20^3 (createBox order), FP64 plain CG to 1e-8, at most 500 iterations:
    index and FP64 assembly bit-exact, same iterations and solution as the oracle.
C2  tets 64^3 x 5, log10 E ~ N(5, 0.5) i.i.d., mixed Jacobi-PCG to 1e-8: FP64
    assembly bit-exact, the certified residual re-checked on the oracle's matrix.
C3  256 spring lattices n in 8..36 solved together (batched mixed CG): every
    lattice's residual re-checked on the oracle's matrix.
C5  hex Poisson 192^3: size-independent properties of the full assembly
    (symmetry, zero row sums away from the Dirichlet face, the stencil storage's
    product equal to the CSR one bit for bit) and certified mixed Jacobi-PCG
    solves over both storages.
C4  tets 96^3 x 5, log10 E = 5 + GRF (std 1, correlation 8 cells): the learned
    mixed PCG of the benchmark step (a network trained briefly on 16^3
    samples), its solution re-checked on the oracle's FP64 matrix.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev
from simit_amd import learned, meshes

pytestmark = pytest.mark.gpu


def springs_params():
    return S.SpringsParams(0.01, 0.0, (C.c_double * 3)(0.0, 0.0, -9.8), 1.0, 1)


def test_c1_springs20(ctx, oracle):
    pts = meshes.springs_lattice(20, seed=0)
    assert pts["npoints"] == 8000 and pts["endpoints"].shape[0] == 22800
    d = {k: to_dev(pts[k]) for k in ("x", "v", "m", "fixed", "k", "l0")}
    d["endpoints"] = to_dev(pts["endpoints"].astype(np.int32))
    idx = ctx.build_index(8000, [d["endpoints"]])
    rp, ci = idx.to_host()
    orp, oci = oracle.index_vev(8000, [pts["endpoints"]])
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(ci, oci)
    assert idx.nnz == 53600
    vals = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(8000 * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_springs(idx, springs_params(), d, d, S.F64, vals, None, rhs)
    ov, of = oracle.springs_assemble(oracle.springs_params(), pts, pts, rp, ci)
    np.testing.assert_array_equal(vals.cpu().numpy(), ov)
    np.testing.assert_array_equal(rhs.cpu().numpy(), of)
    x = torch.zeros_like(rhs)
    st = ctx.cg(idx.bsr(3, S.F64, vals), rhs, x, ctx.cg_options(maxiters=500, tol=1e-8, rel=1))
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    xo, so = oracle.cg(rp, ci, ov, of, np.zeros(24000), 3, oracle.cg_opts(maxiters=500, tol=1e-8))
    assert abs(st.iters - so["iters"]) <= 1, (st.iters, so)
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-7 * np.linalg.norm(xo)


def test_c2_tets64_jacobi_mixed(ctx, oracle):
    prob = meshes.tet_problem(64, seed=0, field="iid", mean_log10=5.0, std_log10=0.5)
    assert prob["nverts"] == 274625 and prob["endpoints"].shape[0] == 1310720
    d = {k: to_dev(prob[k]) for k in ("x", "c", "mu", "lam", "endpoints")}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    assert idx.nnz == 3494465
    rp, ci = idx.to_host()
    P = S.TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
    v64 = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(prob["nverts"] * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_tets(idx, P, d, d, S.F64, v64, None, rhs)
    ov, of = oracle.tet_assemble(oracle.tet_params(), prob, rp, ci)
    np.testing.assert_array_equal(v64.cpu().numpy(), ov)
    np.testing.assert_array_equal(rhs.cpu().numpy(), of)
    del v64
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    ctx.assemble_tets(idx, P, d, d, S.F32X2, hi, lo, rhs)
    x = torch.zeros_like(rhs)
    st = ctx.cg(idx.bsr(3, S.F32X2, hi, lo), rhs, x,
                ctx.cg_options(maxiters=20000, tol=1e-8, rel=1, precond=S.PRECOND_JACOBI,
                               mixed=1, check_every=64))
    print(st.as_dict(), flush=True)
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    r = of - oracle.bsrmv(rp, ci, ov, x.cpu().numpy(), 3)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(of)


def test_c4_tets96_learned_full_size(ctx, oracle):
    prob = meshes.tet_problem(96, seed=0, field="grf", mean_log10=5.0, std_log10=1.0,
                              corr_len=8.0)
    nv = prob["nverts"]
    assert nv == 912673 and prob["endpoints"].shape[0] == 4423680
    d = {k: to_dev(prob[k]) for k in ("x", "c", "mu", "lam", "endpoints")}
    idx = ctx.build_index(nv, [d["endpoints"]])
    assert idx.nnz == 11695969
    net = ctx.gnn(seed=0, kappa=0.9)
    learned.train(ctx, net, n=16, nsamples=300)
    P = S.TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    rhs = torch.empty(3 * nv, dtype=torch.float64, device="cuda")
    ctx.assemble_tets(idx, P, d, d, S.F32X2, hi, lo, rhs)
    A2 = idx.bsr(3, S.F32X2, hi, lo)
    linv, corr = net.precond(A2)
    x = torch.zeros_like(rhs)
    st = ctx.cg(A2, rhs, x, ctx.cg_options(maxiters=40000, tol=1e-8, rel=1, check_every=64,
                                           precond=S.PRECOND_LEARNED, mixed=1, learned=linv,
                                           learned_corr=corr))
    print(st.as_dict(), net.precond_info(), flush=True)
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    assert st.precond_used == S.PRECOND_LEARNED and st.fallbacks == 0
    rp, ci = idx.to_host()
    ov, of = oracle.tet_assemble(oracle.tet_params(), prob, rp, ci)
    r = of - oracle.bsrmv(rp, ci, ov, x.cpu().numpy(), 3)
    rel = np.linalg.norm(r) / np.linalg.norm(of)
    print("oracle relative residual", rel, flush=True)
    assert rel <= 1.01e-8
    net.close()


@pytest.mark.parametrize("delta", [0.3, 0.01])
def test_c3_batch256(ctx, oracle, delta):
    batch = meshes.springs_batch(256, 8, 36, seed=0)
    npts = batch["npoints"]
    d = {k: to_dev(batch[k]) for k in ("x", "v", "m", "fixed", "k", "l0", "endpoints")}
    idx = ctx.build_index(npts, [d["endpoints"]])
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    rhs = torch.empty(3 * npts, dtype=torch.float64, device="cuda")
    ctx.assemble_springs(idx, springs_params(), d, d, S.F32X2, hi, lo, rhs)
    x = torch.zeros_like(rhs)
    stats, _ = ctx.cg_batched(idx.bsr(3, S.F32X2, hi, lo), batch["rows"], rhs, x,
                              ctx.cg_options(maxiters=2000, tol=1e-8, rel=1, mixed=1,
                                             check_every=8, reliable_delta=delta))
    assert all(s["status"] == 0 and s["true_rel"] <= 1e-8 for s in stats)
    # every lattice's FP64 residual on the oracle's matrix
    rp, ci = idx.to_host()
    ov, of = oracle.springs_assemble(oracle.springs_params(), batch, batch, rp, ci)
    r = of - oracle.bsrmv(rp, ci, ov, x.cpu().numpy(), 3)
    rows = batch["rows"].astype(np.int64)
    for m in range(256):
        s0, s1 = 3 * rows[m], 3 * rows[m + 1]
        assert np.linalg.norm(r[s0:s1]) <= 1.01e-8 * np.linalg.norm(of[s0:s1]), m


def test_c5_hex192_properties_and_solve(ctx):
    prob = meshes.hex_problem(192, seed=0, std_log10=1.5, corr_len=12.0, source=1.0)
    nv = prob["nverts"]
    assert nv == 7189057 and prob["endpoints"].shape[0] == 7077888
    d = {k: to_dev(prob[k]) for k in ("dirichlet", "kappa", "endpoints")}
    idx = ctx.build_index(nv, [d["endpoints"]])
    assert idx.nnz == 577 ** 3  # 27-point rows, (3*193 - 2)^3 entries
    HP = S.HexParams(prob["h"], prob["source"], 1.0)
    v64 = torch.empty(idx.nnz, dtype=torch.float64, device="cuda")
    rhs = torch.empty(nv, dtype=torch.float64, device="cuda")
    ctx.assemble_hex(idx, HP, d, d, S.F64, v64, None, rhs)
    A = idx.bsr(1, S.F64, v64)
    g = torch.Generator(device="cpu").manual_seed(7)
    u = torch.randn(nv, dtype=torch.float64, generator=g).cuda()
    w = torch.randn(nv, dtype=torch.float64, generator=g).cuda()
    Au, Aw = torch.empty_like(u), torch.empty_like(w)
    ctx.bsrmv(A, u, Au)
    ctx.bsrmv(A, w, Aw)
    # symmetric: w'Au = u'Aw
    a, b = float(torch.dot(w, Au)), float(torch.dot(u, Aw))
    assert abs(a - b) <= 1e-10 * (abs(a) + abs(b))
    # K 1 = 0 on every row with no Dirichlet neighbour (x index >= 2)
    one = torch.ones(nv, dtype=torch.float64, device="cuda")
    A1 = torch.empty_like(one)
    ctx.bsrmv(A, one, A1)
    N1 = 193
    xi = torch.arange(nv, device="cuda") // (N1 * N1)
    inner = A1[xi >= 2].abs().max().item()
    diag_scale = (v64.abs().max().item())
    assert inner <= 1e-12 * diag_scale, (inner, diag_scale)
    # the 27-point stencil storage of the same matrix: the product bit for bit
    St = ctx.stencil27_from_bsr(A, (N1, N1, N1))
    Su = torch.empty_like(u)
    ctx.stencil27_mv(St, u, Su)
    assert torch.equal(Su, Au)
    del v64, A, St, Su
    hi = torch.empty(idx.nnz, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    ctx.assemble_hex(idx, HP, d, d, S.F32X2, hi, lo, rhs)
    x = torch.zeros_like(rhs)
    st = ctx.cg(idx.bsr(1, S.F32X2, hi, lo), rhs, x,
                ctx.cg_options(maxiters=40000, tol=1e-8, rel=1, precond=S.PRECOND_JACOBI,
                               mixed=1, check_every=64))
    print(st.as_dict(), flush=True)
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    # and over the stencil storage (FP64 vectors, the benchmark's solver)
    St = ctx.stencil27_from_bsr(idx.bsr(1, S.F32X2, hi, lo), (N1, N1, N1))
    x2 = torch.zeros_like(rhs)
    s2 = ctx.cg_stencil27(St, rhs, x2, ctx.cg_options(maxiters=40000, tol=1e-8, rel=1,
                                                      precond=S.PRECOND_JACOBI, mixed=2,
                                                      check_every=64))
    print(s2.as_dict(), flush=True)
    assert s2.status == 0 and s2.true_rel <= 1e-8, s2.as_dict()
