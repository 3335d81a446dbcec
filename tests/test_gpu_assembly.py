"""Device assembly against the oracle: bit-exact in FP64 (same element order,
no fused multiply-add), rounding-exact in FP32, ~2^-48 for the hi+lo pair,
and bitwise repeatable."""
import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def springs_dev(pts):
    d = {k: to_dev(pts[k]) for k in ("x", "v", "m", "fixed", "k", "l0")}
    d["endpoints"] = to_dev(pts["endpoints"].astype(np.int32))
    return d


def run_springs(ctx, pts, prec, params):
    d = springs_dev(pts)
    idx = ctx.build_index(pts["npoints"], [d["endpoints"]])
    dt = torch.float64 if prec == S.F64 else torch.float32
    vals = torch.empty(idx.nnz * 9, dtype=dt, device="cuda")
    lo = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda") if prec == S.F32X2 else None
    rhs = torch.empty(pts["npoints"] * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_springs(idx, params, d, d, prec, vals, lo, rhs)
    torch.cuda.synchronize()
    return idx, vals, lo, rhs


def make_sp_params(oracle, **kw):
    import ctypes as C
    po = oracle.springs_params(**kw)
    pp = S.SpringsParams(po.h, po.viscous, (C.c_double * 3)(*list(po.grav)), po.fixed_fill,
                         po.fixed_identity)
    return po, pp


@pytest.mark.parametrize("n,fixed_identity", [(6, 1), (9, 0)])
def test_springs_fp64_bit_exact(ctx, oracle, n, fixed_identity):
    pts = meshes.springs_lattice(n, seed=n)
    pts["v"] = np.random.default_rng(1).standard_normal((pts["npoints"], 3))
    po, pp = make_sp_params(oracle, h=0.01, viscous=0.1, grav=(0.0, 0.0, -9.8),
                            fixed_identity=fixed_identity)
    idx, vals, _, rhs = run_springs(ctx, pts, S.F64, pp)
    rp, ci = idx.to_host()
    ov, of = oracle.springs_assemble(po, pts, pts, rp, ci)
    np.testing.assert_array_equal(vals.cpu().numpy(), ov)
    np.testing.assert_array_equal(rhs.cpu().numpy(), of)


def test_springs_isolated_and_degenerate(ctx, oracle):
    rng = np.random.default_rng(4)
    npts = 40
    x = rng.standard_normal((npts, 3)) * 3
    ep = rng.integers(0, 30, size=(60, 2)).astype(np.int32)  # points 30..39 isolated
    ep[ep[:, 0] == ep[:, 1], 1] = (ep[ep[:, 0] == ep[:, 1], 1] + 1) % 30
    pts = {"npoints": npts, "x": x, "v": rng.standard_normal((npts, 3)), "m": rng.uniform(1, 2, npts),
           "fixed": (rng.uniform(size=npts) < 0.2).astype(np.uint8), "endpoints": ep,
           "k": rng.uniform(10, 100, 60), "l0": rng.uniform(0.5, 2, 60)}
    po, pp = make_sp_params(oracle, h=0.02, viscous=0.3)
    idx, vals, _, rhs = run_springs(ctx, pts, S.F64, pp)
    rp, ci = idx.to_host()
    ov, of = oracle.springs_assemble(po, pts, pts, rp, ci)
    np.testing.assert_array_equal(vals.cpu().numpy(), ov)
    np.testing.assert_array_equal(rhs.cpu().numpy(), of)


def test_springs_f32_and_hilo(ctx, oracle):
    pts = meshes.springs_lattice(8, seed=3)
    po, pp = make_sp_params(oracle)
    idx, _, _, _ = run_springs(ctx, pts, S.F64, pp)
    rp, ci = idx.to_host()
    ov, _ = oracle.springs_assemble(po, pts, pts, rp, ci)
    _, v32, _, _ = run_springs(ctx, pts, S.F32, pp)
    np.testing.assert_array_equal(v32.cpu().numpy(), ov.astype(np.float32))
    _, hi, lo, _ = run_springs(ctx, pts, S.F32X2, pp)
    hl = hi.cpu().numpy().astype(np.float64) + lo.cpu().numpy().astype(np.float64)
    assert np.abs(hl - ov).max() <= 2.0 ** -47 * np.abs(ov).max()


def tet_dev(prob):
    d = {"x": to_dev(prob["x"]), "c": to_dev(prob["c"].astype(np.int32)),
         "mu": to_dev(prob["mu"]), "lam": to_dev(prob["lam"]),
         "endpoints": to_dev(prob["endpoints"].astype(np.int32))}
    return d


def tet_params(oracle, **kw):
    import ctypes as C
    po = oracle.tet_params(**kw)
    pp = S.TetParams(po.stiff_scale, po.rho, (C.c_double * 3)(*list(po.load)), po.fixed_diag)
    return po, pp


def run_tets(ctx, prob, prec, pp):
    d = tet_dev(prob)
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    dt = torch.float64 if prec == S.F64 else torch.float32
    vals = torch.empty(idx.nnz * 9, dtype=dt, device="cuda")
    lo = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda") if prec == S.F32X2 else None
    rhs = torch.empty(prob["nverts"] * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_tets(idx, pp, d, d, prec, vals, lo, rhs)
    torch.cuda.synchronize()
    return idx, vals, lo, rhs


@pytest.mark.parametrize("kw", [dict(), dict(stiff_scale=0.005 ** 2, rho=1e3, fixed_diag=0.0)])
def test_tets_fp64_bit_exact(ctx, oracle, kw):
    prob = meshes.tet_problem(5, seed=2)
    prob["x"] = prob["x"] + np.random.default_rng(0).uniform(-0.02, 0.02, prob["x"].shape)
    po, pp = tet_params(oracle, **kw)
    idx, vals, _, rhs = run_tets(ctx, prob, S.F64, pp)
    rp, ci = idx.to_host()
    ov, of = oracle.tet_assemble(po, prob, rp, ci)
    np.testing.assert_array_equal(vals.cpu().numpy(), ov)
    np.testing.assert_array_equal(rhs.cpu().numpy(), of)


def test_tets_full_size_bit_exact(ctx, oracle):
    # the 96^3 x 5 workload: 182,750 waves of blocks, so every persistent
    # workgroup walks many of them
    prob = meshes.tet_problem(96, seed=0, field="grf", mean_log10=5.0, std_log10=1.0,
                              corr_len=8.0)
    po, pp = tet_params(oracle, load=prob["load"])
    idx, vals, _, rhs = run_tets(ctx, prob, S.F64, pp)
    rp, ci = idx.to_host()
    assert idx.nnz == 11695969
    ov, of = oracle.tet_assemble(po, prob, rp, ci)
    got = vals.cpu().numpy()
    bad = np.nonzero(got != ov)[0]
    assert bad.size == 0, f"{bad.size} values differ, first at {bad[:8]}"
    np.testing.assert_array_equal(rhs.cpu().numpy(), of)


def test_tets_hilo_and_determinism(ctx, oracle):
    prob = meshes.tet_problem(12, seed=5, field="grf", std_log10=1.0, corr_len=3.0)
    po, pp = tet_params(oracle)
    idx, hi, lo, rhs = run_tets(ctx, prob, S.F32X2, pp)
    rp, ci = idx.to_host()
    ov, of = oracle.tet_assemble(po, prob, rp, ci)
    hl = hi.cpu().numpy().astype(np.float64) + lo.cpu().numpy().astype(np.float64)
    assert np.abs(hl - ov).max() <= 2.0 ** -47 * np.abs(ov).max()
    np.testing.assert_array_equal(hi.cpu().numpy(), ov.astype(np.float32))
    _, hi2, lo2, rhs2 = run_tets(ctx, prob, S.F32X2, pp)
    assert torch.equal(hi, hi2) and torch.equal(lo, lo2) and torch.equal(rhs, rhs2)


def test_hex_fp64_bit_exact(ctx, oracle):
    import ctypes as C
    prob = meshes.hex_problem(6, seed=1, corr_len=2.0)
    d = {"dirichlet": to_dev(prob["dirichlet"]), "kappa": to_dev(prob["kappa"]),
         "endpoints": to_dev(prob["endpoints"])}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    vals = torch.empty(idx.nnz, dtype=torch.float64, device="cuda")
    rhs = torch.empty(prob["nverts"], dtype=torch.float64, device="cuda")
    ctx.assemble_hex(idx, S.HexParams(prob["h"], 2.0, 1.0), d, d, S.F64, vals, None, rhs)
    rp, ci = idx.to_host()
    ov, of = oracle.hex_assemble(oracle.HexParams(prob["h"], 2.0, 1.0), prob, rp, ci)
    np.testing.assert_array_equal(vals.cpu().numpy(), ov)
    np.testing.assert_array_equal(rhs.cpu().numpy(), of)


def test_assembly_rejects_foreign_set(ctx):
    prob = meshes.tet_problem(2)
    d = tet_dev(prob)
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    other = to_dev(prob["endpoints"].astype(np.int32))
    d2 = dict(d, endpoints=other)
    vals = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(prob["nverts"] * 3, dtype=torch.float64, device="cuda")
    import ctypes as C
    pp = S.TetParams(1.0, 0.0, (C.c_double * 3)(0, 0, 0), 1.0)
    with pytest.raises(S.SimitError):
        ctx.assemble_tets(idx, pp, d2, d2, S.F64, vals, None, rhs)
