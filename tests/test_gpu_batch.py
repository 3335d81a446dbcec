"""Batched CG (smh_cg_solve_batched): many independent spring lattices in one
block-diagonal system, each solved to its own 1e-8 relative residual.

Checked against the same systems solved one by one (smh_cg_solve, and the
oracle's CG on the CPU), the certified FP64 residual of every system, bitwise
repeatability of the graph-replayed iteration, and the rejection of row
ranges that a block couples.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def params():
    return S.SpringsParams(0.01, 0.0, (C.c_double * 3)(0.0, 0.0, -9.8), 1.0, 1)


def assemble(ctx, pts, prec):
    d = {k: to_dev(pts[k]) for k in ("x", "v", "m", "fixed", "k", "l0")}
    d["endpoints"] = to_dev(pts["endpoints"].astype(np.int32))
    idx = ctx.build_index(pts["npoints"], [d["endpoints"]])
    dt = torch.float64 if prec == S.F64 else torch.float32
    vals = torch.empty(idx.nnz * 9, dtype=dt, device="cuda")
    lo = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda") if prec == S.F32X2 else None
    rhs = torch.empty(pts["npoints"] * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_springs(idx, params(), d, d, prec, vals, lo, rhs)
    torch.cuda.synchronize()
    return {"idx": idx, "vals": vals, "lo": lo, "rhs": rhs, "d": d}


def split(batch, m):
    r0, r1 = int(batch["rows"][m]), int(batch["rows"][m + 1])
    sel = (batch["endpoints"][:, 0] >= r0) & (batch["endpoints"][:, 0] < r1)
    out = {"npoints": r1 - r0, "endpoints": batch["endpoints"][sel] - r0}
    for k in ("x", "v", "m", "fixed"):
        out[k] = batch[k][r0:r1]
    for k in ("k", "l0"):
        out[k] = batch[k][sel]
    return out


@pytest.mark.parametrize("mode", ["mixed_none", "f64_block", "mixed_block", "f64_none_replace3"])
def test_batched_matches_one_by_one(ctx, oracle, mode):
    batch = meshes.springs_batch(7, 3, 7, seed=3)
    prec = S.F64 if mode.startswith("f64") else S.F32X2
    pc = S.PRECOND_BLOCK_JACOBI if "block" in mode else S.PRECOND_NONE
    mixed = int(mode.startswith("mixed"))
    sy = assemble(ctx, batch, prec)
    A = sy["idx"].bsr(3, prec, sy["vals"], sy["lo"])
    x = torch.zeros_like(sy["rhs"])
    opts = ctx.cg_options(maxiters=2000, tol=1e-8, rel=1, precond=pc, mixed=mixed, check_every=8,
                          replace_every=3 if mode.endswith("replace3") else 0)
    stats, _ = ctx.cg_batched(A, batch["rows"], sy["rhs"], x, opts)
    xb = x.cpu().numpy()
    for m, st in enumerate(stats):
        assert st["status"] == 0 and st["true_rel"] <= 1e-8, (m, st)
        one = split(batch, m)
        s1 = assemble(ctx, one, prec)
        A1 = s1["idx"].bsr(3, prec, s1["vals"], s1["lo"])
        x1 = torch.zeros_like(s1["rhs"])
        st1 = ctx.cg(A1, s1["rhs"], x1, opts)
        assert st1.status == 0
        r0, r1 = 3 * int(batch["rows"][m]), 3 * int(batch["rows"][m + 1])
        xm = xb[r0:r1]
        ref = x1.cpu().numpy()
        assert np.linalg.norm(xm - ref) <= 1e-6 * np.linalg.norm(ref), m
        if prec == S.F64:
            assert abs(st["iters"] - st1.iters) <= 2, (m, st["iters"], st1.iters)
        # the oracle's FP64 residual of this system's solution
        rp, ci = s1["idx"].to_host()
        ov, of = oracle.springs_assemble(oracle.springs_params(), one, one, rp, ci)
        res = of - oracle.bsrmv(rp, ci, ov, xm, 3)
        assert np.linalg.norm(res) <= 1.01e-8 * np.linalg.norm(of), m


def test_batched_repeatable_and_uneven(ctx):
    """Lattices of 8 to 729 points; lattice 1 is mass-dominated (condition
    number ~1.06) and converges 1e4 per iteration, so its reliable updates
    must happen inside the iteration, before beta."""
    batch = meshes.springs_batch(12, 2, 9, seed=5)
    sy = assemble(ctx, batch, S.F32X2)
    A = sy["idx"].bsr(3, S.F32X2, sy["vals"], sy["lo"])
    outs = []
    for _ in range(2):
        x = torch.zeros_like(sy["rhs"])
        stats, _ = ctx.cg_batched(A, batch["rows"], sy["rhs"], x,
                                  ctx.cg_options(maxiters=3000, tol=1e-8, rel=1, mixed=1,
                                                 check_every=4))
        assert all(s["status"] == 0 and s["true_rel"] <= 1e-8 for s in stats), stats
        outs.append((x.clone(), [s["iters"] for s in stats]))
        torch.cuda.synchronize()
    assert torch.equal(outs[0][0], outs[1][0]) and outs[0][1] == outs[1][1]
    # systems of different size converge after different numbers of iterations
    assert len(set(outs[0][1])) > 1


def test_batched_rejects_coupled_rows(ctx):
    batch = meshes.springs_batch(3, 3, 4, seed=1)
    sy = assemble(ctx, batch, S.F64)
    A = sy["idx"].bsr(3, S.F64, sy["vals"])
    rows = batch["rows"].copy()
    rows[1] += 1  # cuts lattice 0 short: its last point's blocks now couple systems
    with pytest.raises(S.SimitError):
        ctx.cg_batched(A, rows, sy["rhs"], torch.zeros_like(sy["rhs"]), ctx.cg_options())
    bad = batch["rows"].copy()
    bad[-1] -= 1
    with pytest.raises(S.SimitError):
        ctx.cg_batched(A, bad, sy["rhs"], torch.zeros_like(sy["rhs"]), ctx.cg_options())


def test_springs_solve_batched_equals_assembly_then_batched_cg(ctx):
    """The one-call step (assembly captured in the solver's first graph) gives
    the same x and statistics, bit for bit, as the assembly followed by the
    batched solve."""
    import ctypes as Cc
    batch = meshes.springs_batch(24, 5, 14, seed=4)
    npts = batch["npoints"]
    d = {k: to_dev(batch[k]) for k in ("x", "v", "m", "fixed", "k", "l0", "endpoints")}
    idx = ctx.build_index(npts, [d["endpoints"]])
    P = S.SpringsParams(0.01, 0.0, (Cc.c_double * 3)(0.0, 0.0, -9.8), 1.0, 1)
    o = ctx.cg_options(maxiters=2000, tol=1e-8, rel=1, mixed=1, check_every=4,
                       reliable_delta=0.01)
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    rhs = torch.empty(3 * npts, dtype=torch.float64, device="cuda")
    ctx.assemble_springs(idx, P, d, d, S.F32X2, hi, lo, rhs)
    x1 = torch.zeros_like(rhs)
    st1, _ = ctx.cg_batched(idx.bsr(3, S.F32X2, hi, lo), batch["rows"], rhs, x1, o)
    hi2, lo2, rhs2 = torch.zeros_like(hi), torch.zeros_like(lo), torch.zeros_like(rhs)
    x2 = torch.zeros_like(rhs)
    st2, _ = ctx.springs_solve_batched(idx, P, d, d, hi2, lo2, rhs2, batch["rows"], x2, o)
    assert torch.equal(hi, hi2) and torch.equal(lo, lo2) and torch.equal(rhs, rhs2)
    assert torch.equal(x1, x2)
    assert st1 == st2
    assert all(s["status"] == 0 and s["true_rel"] <= 1e-8 for s in st2)


def test_springs_solve_batched_reuses_its_graphs_with_new_values(ctx):
    """A second call with the same buffers replays the cached graphs: new
    stiffnesses written between the calls are assembled and solved (the
    result equals a fresh context's), and an unchanged call repeats bit for
    bit."""
    import ctypes as Cc
    batch = meshes.springs_batch(12, 5, 11, seed=8)
    npts = batch["npoints"]
    P = S.SpringsParams(0.01, 0.0, (Cc.c_double * 3)(0.0, 0.0, -9.8), 1.0, 1)
    o = ctx.cg_options(maxiters=2000, tol=1e-8, rel=1, mixed=1, check_every=4)

    def run(c, d, idx, hi, lo, rhs, x):
        x.zero_()
        st, _ = c.springs_solve_batched(idx, P, d, d, hi, lo, rhs, batch["rows"], x, o)
        return [s["iters"] for s in st], x.clone()

    d = {k: to_dev(batch[k]) for k in ("x", "v", "m", "fixed", "k", "l0", "endpoints")}
    idx = ctx.build_index(npts, [d["endpoints"]])
    bufs = [torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda") for _ in range(2)]
    rhs = torch.empty(3 * npts, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(rhs)
    it1, x1 = run(ctx, d, idx, bufs[0], bufs[1], rhs, x)
    it1b, x1b = run(ctx, d, idx, bufs[0], bufs[1], rhs, x)
    assert it1 == it1b and torch.equal(x1, x1b)
    d["k"].mul_(3.0)  # stiffer springs, same buffers
    it2, x2 = run(ctx, d, idx, bufs[0], bufs[1], rhs, x)
    assert not torch.equal(x1, x2)
    fresh = S.Context(0)
    idx2 = fresh.build_index(npts, [d["endpoints"]])
    b2 = [torch.empty_like(bufs[0]) for _ in range(2)]
    it3, x3 = run(fresh, d, idx2, b2[0], b2[1], torch.empty_like(rhs), torch.zeros_like(rhs))
    assert it2 == it3 and torch.equal(x2, x3)
    fresh.close()
