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
