"""The device CG solver against the reference's program fixtures, the oracle
and a certified FP64 residual."""
import ctypes as C

import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def test_program_cg_fixture(ctx, oracle, fixtures):
    """test/program-cg-test.cpp: A = I - 0.01*(map f to springs), cg.sim's loop."""
    from test_oracle import cg_program_matrix
    fx = fixtures["program_cg"]
    rp, ci, vals = cg_program_matrix(oracle, fx)
    idx = ctx.build_index(3, [to_dev(np.array(fx["springs"], np.int32))])
    grp, gci = idx.to_host()
    np.testing.assert_array_equal(grp, rp)
    b = to_dev(np.array(fx["b"]))
    x = torch.zeros(3, dtype=torch.float64, device="cuda")
    o = ctx.cg_options(maxiters=fx["maxiters"], tol=fx["tol"], rel=0, norm_form=1)
    st = ctx.cg(idx.bsr(1, S.F64, to_dev(vals)), b, x, o)
    np.testing.assert_allclose(x.cpu().numpy(), fx["expected_c"], atol=fx["abs_tol"], rtol=0)


def test_program_isprings_fixture(ctx, oracle, fixtures):
    """test/program-isprings.cpp:91-175: ten implicit steps, assembly + CG on the GPU."""
    from test_oracle import isprings_setup
    fx = fixtures["program_isprings"]
    pts, sp = isprings_setup(fx)
    d = {k: to_dev(v) for k, v in pts.items()}
    d.update({k: to_dev(v) for k, v in sp.items()})
    idx = ctx.build_index(3, [d["endpoints"]])
    P = S.SpringsParams(fx["h"], fx["viscous"], (C.c_double * 3)(*fx["grav"]), 1.0, 0)
    vals = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(9, dtype=torch.float64, device="cuda")
    A = idx.bsr(3, S.F64, vals)
    o = ctx.cg_options(maxiters=fx["maxiters"], tol=fx["tol"], rel=0)
    for _ in range(fx["steps"]):
        ctx.assemble_springs(idx, P, d, d, S.F64, vals, None, rhs)
        ctx.cg(A, rhs, d["v"], o)                 # vguess = points.v, v = cg(...)
        ctx.axpby(fx["h"], d["v"], 1.0, d["x"])   # points.x = points.x + h*points.v
    np.testing.assert_allclose(d["v"].cpu().numpy(), fx["expected_v"], atol=fx["abs_tol"], rtol=0)
    np.testing.assert_allclose(d["x"].cpu().numpy(), fx["expected_x"], atol=fx["abs_tol"], rtol=0)


def tet_system(ctx, n, seed=0, field="iid", std=0.5):
    prob = meshes.tet_problem(n, seed=seed, field=field, std_log10=std, corr_len=3.0)
    d = {"x": to_dev(prob["x"]), "c": to_dev(prob["c"]), "mu": to_dev(prob["mu"]),
         "lam": to_dev(prob["lam"]), "endpoints": to_dev(prob["endpoints"])}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    P = S.TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
    v64 = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    rhs = torch.empty(prob["nverts"] * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_tets(idx, P, d, d, S.F64, v64, None, rhs)
    ctx.assemble_tets(idx, P, d, d, S.F32X2, hi, lo, rhs)
    return prob, idx, v64, hi, lo, rhs, d


@pytest.mark.parametrize("mode", ["plain64", "jacobi64", "block64", "mixed_block", "mixed_none",
                                  "mixed64v_block", "hilo64"])
def test_tet_cg_modes_reach_1e8(ctx, oracle, mode):
    prob, idx, v64, hi, lo, rhs, _ = tet_system(ctx, 10, seed=3)
    rp, ci = idx.to_host()
    n = prob["nverts"] * 3
    pc = {"plain64": 0, "jacobi64": 1, "block64": 2, "mixed_block": 2, "mixed_none": 0,
          "mixed64v_block": 2, "hilo64": 2}[mode]
    mixed = 2 if mode.startswith("mixed64v") else int(mode.startswith("mixed"))
    A = idx.bsr(3, S.F32X2, hi, lo) if (mixed or mode == "hilo64") else idx.bsr(3, S.F64, v64)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    o = ctx.cg_options(maxiters=20000, tol=1e-8, rel=1, precond=pc, mixed=mixed)
    st = ctx.cg(A, rhs, x, o)
    print(mode, st.as_dict(), flush=True)
    assert st.status == 0, st.as_dict()
    assert st.true_rel <= 1e-8
    # independent FP64 check of the residual on the oracle's matrix
    ov, of = oracle.tet_assemble(oracle.tet_params(), prob, rp, ci)
    xv = x.cpu().numpy()
    r = of - oracle.bsrmv(rp, ci, ov, xv, 3)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(of)
    # and the same solution as the oracle's CG
    xo, _ = oracle.cg(rp, ci, ov, of, np.zeros(n), 3, oracle.cg_opts(maxiters=20000, tol=1e-10))
    assert np.linalg.norm(xv - xo) <= 1e-5 * np.linalg.norm(xo)


def test_mixed_inner_spmv_sums_rows_in_fp64(ctx, oracle):
    """mixed = 1 keeps FP32 vectors, but its inner SpMV sums each row in FP64
    (the FP32 x FP32 products are exact there) and rounds q once.  On a stiff
    tet system (log10 E a random field with std 1) that keeps its block-Jacobi
    PCG within a few percent of mixed = 2's iteration count; FP32 row sums lose
    the small A p of the smooth modes to the cancellation of large terms (a
    numpy replay of the solver on this system: 1,572 iterations with FP32 row
    sums, 1,264 with FP64 row sums, 1,181 with FP64 vectors)."""
    prob = meshes.tet_problem(32, seed=0, field="grf", std_log10=1.0, corr_len=8.0)
    d = {k: to_dev(prob[k]) for k in ("x", "c", "mu", "lam", "endpoints")}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    P = S.TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
    v64 = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(prob["nverts"] * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_tets(idx, P, d, d, S.F64, v64, None, rhs)
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    ctx.split_hilo(v64, hi, lo)
    A = idx.bsr(3, S.F32X2, hi, lo)
    rp, ci = idx.to_host()
    ov, of = oracle.tet_assemble(oracle.tet_params(load=prob["load"]), prob, rp, ci)
    its = {}
    for mixed in (1, 2):
        x = torch.zeros_like(rhs)
        st = ctx.cg(A, rhs, x, ctx.cg_options(maxiters=20000, tol=1e-8, rel=1, mixed=mixed,
                                              precond=S.PRECOND_BLOCK_JACOBI, vals_f64=v64))
        print(mixed, st.as_dict(), flush=True)
        assert st.status == 0 and st.true_rel <= 1e-8, (mixed, st.as_dict())
        r = of - oracle.bsrmv(rp, ci, ov, x.cpu().numpy(), 3)
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(of)
        its[mixed] = st.iters
    assert its[1] <= 1.15 * its[2], its


def test_cg_repeatable_bitwise(ctx):
    _, idx, v64, hi, lo, rhs, _ = tet_system(ctx, 8, seed=1)
    outs = []
    for _ in range(2):
        x = torch.zeros_like(rhs)
        st = ctx.cg(idx.bsr(3, S.F32X2, hi, lo), rhs, x,
                    ctx.cg_options(maxiters=5000, tol=1e-8, precond=2, mixed=1))
        outs.append((x.clone(), st.iters))
    assert torch.equal(outs[0][0], outs[1][0]) and outs[0][1] == outs[1][1]


def test_replace_every_follows_oracle(ctx, oracle):
    """opts.replace_every = k: r = b - A x every k iterations, between the x/r
    update and beta, as the oracle's loop does (thermal.sim's periodic
    replacement); same iteration count and solution as the oracle."""
    prob, idx, v64, hi, lo, rhs, _ = tet_system(ctx, 8, seed=2)
    rp, ci = idx.to_host()
    n = prob["nverts"] * 3
    ov, of = oracle.tet_assemble(oracle.tet_params(), prob, rp, ci)
    for k in (0, 7, 1):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        st = ctx.cg(idx.bsr(3, S.F64, v64), rhs, x,
                    ctx.cg_options(maxiters=5000, tol=1e-8, rel=1, replace_every=k))
        xo, so = oracle.cg(rp, ci, ov, of, np.zeros(n), 3,
                           oracle.cg_opts(maxiters=5000, tol=1e-8, replace_every=k))
        print(k, st.as_dict(), so, flush=True)
        assert st.status == 0 and st.true_rel <= 1e-8
        # FP64 in both, summed in different orders: over ~370 iterations of an
        # ill-conditioned system the counts drift by a few (plain CG too)
        assert abs(st.iters - so["iters"]) <= max(2, 0.02 * so["iters"]), (k, st.iters, so)
        if k:
            assert st.reliable_updates >= st.iters // k
        assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-6 * np.linalg.norm(xo)


def test_mixed_fast_convergence(ctx, oracle):
    """A mass-dominated spring lattice (condition number ~1.06): the FP32
    iterated residual falls 1e4 per iteration, far below the FP32 floor of
    the true residual within one check_every chunk; the replacement before
    beta keeps the iteration converging to a certified 1e-8."""
    pts = meshes.springs_lattice(8, seed=11, k=123.0)
    d = {k: to_dev(v) for k, v in pts.items() if k != "npoints"}
    idx = ctx.build_index(pts["npoints"], [d["endpoints"]])
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    rhs = torch.empty(pts["npoints"] * 3, dtype=torch.float64, device="cuda")
    P = S.SpringsParams(0.01, 0.0, (C.c_double * 3)(0.0, 0.0, -9.8), 1.0, 1)
    ctx.assemble_springs(idx, P, d, d, S.F32X2, hi, lo, rhs)
    for check in (1, 4, 16):
        x = torch.zeros_like(rhs)
        st = ctx.cg(idx.bsr(3, S.F32X2, hi, lo), rhs, x,
                    ctx.cg_options(maxiters=200, tol=1e-8, rel=1, mixed=1, check_every=check))
        assert st.status == 0 and st.true_rel <= 1e-8 and st.iters <= 12, (check, st.as_dict())


def test_hex_poisson_mixed(ctx, oracle):
    prob = meshes.hex_problem(12, seed=2, corr_len=3.0)
    d = {"dirichlet": to_dev(prob["dirichlet"]), "kappa": to_dev(prob["kappa"]),
         "endpoints": to_dev(prob["endpoints"])}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    hi = torch.empty(idx.nnz, dtype=torch.float32, device="cuda")
    lo = torch.empty(idx.nnz, dtype=torch.float32, device="cuda")
    rhs = torch.empty(prob["nverts"], dtype=torch.float64, device="cuda")
    ctx.assemble_hex(idx, S.HexParams(prob["h"], 1.0, 1.0), d, d, S.F32X2, hi, lo, rhs)
    x = torch.zeros_like(rhs)
    st = ctx.cg(idx.bsr(1, S.F32X2, hi, lo), rhs, x,
                ctx.cg_options(maxiters=5000, tol=1e-8, precond=1, mixed=1))
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    rp, ci = idx.to_host()
    ov, of = oracle.hex_assemble(oracle.HexParams(prob["h"], 1.0, 1.0), prob, rp, ci)
    r = of - oracle.bsrmv(rp, ci, ov, x.cpu().numpy(), 1)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(of)


def test_maxiters_reports_not_converged(ctx):
    _, idx, v64, hi, lo, rhs, _ = tet_system(ctx, 6)
    x = torch.zeros_like(rhs)
    st = ctx.cg(idx.bsr(3, S.F64, v64), rhs, x, ctx.cg_options(maxiters=3, tol=1e-12))
    assert st.iters == 3 and st.status == S.ERR_NOT_CONVERGED


def test_indefinite_matrix_breaks_down(ctx):
    """p'Ap <= 0 is detected and reported, not iterated on."""
    n = 4
    ep = np.array([[0, 1], [1, 2], [2, 3]], np.int32)
    idx = ctx.build_index(n, [to_dev(ep)])
    rp, ci = idx.to_host()
    vals = np.array([-1.0 if ci[l] == i else 0.0 for i in range(n) for l in range(rp[i], rp[i + 1])])
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    st = ctx.cg(idx.bsr(1, S.F64, to_dev(vals)), to_dev(np.ones(n)), x,
                ctx.cg_options(maxiters=10, tol=1e-10))
    assert st.status == S.ERR_BREAKDOWN


def test_program_cg_fully_on_device(ctx, fixtures):
    """cg.sim end to end on the GPU: A = I - 0.01*(map f to springs) assembled
    by the device edge-block kernel, then the device CG."""
    fx = fixtures["program_cg"]
    ep = to_dev(np.array(fx["springs"], np.int32))
    idx = ctx.build_index(3, [ep])
    w = to_dev(np.array(fx["a"], dtype=np.float64))
    vals = torch.empty(idx.nnz, dtype=torch.float64, device="cuda")
    ctx.assemble_edge_blocks(idx, ep, w, 1, sign=(1, -1, -1, 1), alpha=1.0, beta=-0.01,
                             vals=vals)
    x = torch.zeros(3, dtype=torch.float64, device="cuda")
    ctx.cg(idx.bsr(1, S.F64, vals), to_dev(np.array(fx["b"])), x,
           ctx.cg_options(maxiters=fx["maxiters"], tol=fx["tol"], rel=0, norm_form=1))
    np.testing.assert_allclose(x.cpu().numpy(), fx["expected_c"], atol=fx["abs_tol"], rtol=0)


def test_matrix_vv_and_gemv_blocked_on_device(ctx, fixtures):
    fx = fixtures["assembly_matrix_vv"]
    ep = to_dev(np.array(fx["springs"], np.int32))
    idx = ctx.build_index(3, [ep])
    vals = torch.empty(idx.nnz, dtype=torch.float64, device="cuda")
    ctx.assemble_edge_blocks(idx, ep, to_dev(np.ones(2)), 1, vals=vals)
    y = torch.empty(3, dtype=torch.float64, device="cuda")
    ctx.bsrmv(idx.bsr(1, S.F64, vals), to_dev(np.array(fx["a"], dtype=np.float64)), y)
    np.testing.assert_array_equal(y.cpu().numpy(), fx["expected_b"])
    fx = fixtures["gemv_blocked"]
    ep = to_dev(np.array(fx["springs"], np.int32))
    idx = ctx.build_index(3, [ep])
    vals = torch.empty(idx.nnz * 4, dtype=torch.float64, device="cuda")
    ctx.assemble_edge_blocks(idx, ep, to_dev(np.array(fx["a"]).ravel()), 2, vals=vals)
    y = torch.empty(6, dtype=torch.float64, device="cuda")
    ctx.bsrmv(idx.bsr(2, S.F64, vals), to_dev(np.array(fx["b"]).ravel()), y, exact_order=True)
    np.testing.assert_array_equal(y.cpu().numpy().reshape(3, 2), fx["expected_c"])
