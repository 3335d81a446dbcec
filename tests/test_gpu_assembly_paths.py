"""The two assembly kernels agree bit for bit.

Every assembly runs waves of 64 blocks through a fast kernel, which uses a
per-pattern plan (pair slots, per-block contribution masks), and leaves the
waves it cannot take (rows outside its window, rows with more pairs than mask
bits, too many diagonal entries) to the generic kernel.  SMH_ASSEMBLY_GENERIC=1
sends every wave through the generic kernel, which the oracle tests pin; here
both routes must give identical bits on meshes that exercise each class of
wave: regular tets, springs and hexes, self-loops and repeated elements,
high-degree vertices (64-bit masks, generic waves), interleaved empty rows and
rows of two or three blocks.
"""
import os

import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def both_routes(run):
    """run() -> list of tensors; returns (fast, generic) results."""
    old = os.environ.pop("SMH_ASSEMBLY_GENERIC", None)
    try:
        fast = run()
        os.environ["SMH_ASSEMBLY_GENERIC"] = "1"
        gen = run()
    finally:
        os.environ.pop("SMH_ASSEMBLY_GENERIC", None)
        if old is not None:
            os.environ["SMH_ASSEMBLY_GENERIC"] = old
    torch.cuda.synchronize()
    return fast, gen


def assert_same(fast, gen):
    for a, b in zip(fast, gen):
        if a is None:
            continue
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert not np.isnan(a).any(), "unwritten values"
        iv = np.int64 if a.itemsize == 8 else np.int32
        bad = np.nonzero(a.view(iv) != b.view(iv))[0]
        assert bad.size == 0, f"{bad.size} values differ, first at {bad[:8]}"


def nan_like(n, dt):
    return torch.full((n,), float("nan"), dtype=dt, device="cuda")


@pytest.mark.parametrize("prec", [S.F64, S.F32X2])
def test_tets_routes(ctx, prec):
    prob = meshes.tet_problem(10, seed=3, field="grf", std_log10=1.0, corr_len=3.0)
    prob["x"] = prob["x"] + np.random.default_rng(0).uniform(-0.02, 0.02, prob["x"].shape)
    d = {"x": to_dev(prob["x"]), "c": to_dev(prob["c"].astype(np.int32)),
         "mu": to_dev(prob["mu"]), "lam": to_dev(prob["lam"]),
         "endpoints": to_dev(prob["endpoints"].astype(np.int32))}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    import ctypes as C
    pp = S.TetParams(0.7, 1e3, (C.c_double * 3)(0.0, -1.0, -9.81), 2.5)

    def run():
        dt = torch.float64 if prec == S.F64 else torch.float32
        vals = nan_like(idx.nnz * 9, dt)
        lo = nan_like(idx.nnz * 9, torch.float32) if prec == S.F32X2 else None
        rhs = nan_like(prob["nverts"] * 3, torch.float64)
        ctx.assemble_tets(idx, pp, d, d, prec, vals, lo, rhs)
        return [vals, lo, rhs]
    assert_same(*both_routes(run))


def test_springs_routes(ctx):
    import ctypes as C
    pts = meshes.springs_lattice(7, seed=2)
    pts["v"] = np.random.default_rng(1).standard_normal((pts["npoints"], 3))
    pts["fixed"][::5] = 1
    d = {k: to_dev(pts[k]) for k in ("x", "v", "m", "fixed", "k", "l0")}
    d["endpoints"] = to_dev(pts["endpoints"].astype(np.int32))
    idx = ctx.build_index(pts["npoints"], [d["endpoints"]])
    pp = S.SpringsParams(0.01, 0.1, (C.c_double * 3)(0.0, 0.0, -9.8), 1.0, 0)

    def run():
        vals = nan_like(idx.nnz * 9, torch.float64)
        rhs = nan_like(pts["npoints"] * 3, torch.float64)
        ctx.assemble_springs(idx, pp, d, d, S.F64, vals, None, rhs)
        return [vals, rhs]
    assert_same(*both_routes(run))


def test_hex_routes(ctx):
    prob = meshes.hex_problem(9, seed=1, corr_len=2.0)
    d = {"dirichlet": to_dev(prob["dirichlet"]), "kappa": to_dev(prob["kappa"]),
         "endpoints": to_dev(prob["endpoints"])}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])

    def run():
        vals = nan_like(idx.nnz, torch.float64)
        rhs = nan_like(prob["nverts"], torch.float64)
        ctx.assemble_hex(idx, S.HexParams(prob["h"], 2.0, 1.0), d, d, S.F64, vals, None, rhs)
        return [vals, rhs]
    assert_same(*both_routes(run))


def edge_graphs():
    rng = np.random.default_rng(7)
    nv = 700
    # random graph with self-loops and repeated edges
    e1 = rng.integers(0, nv, size=(4000, 2))
    e1[:40, 1] = e1[:40, 0]
    e1 = np.concatenate([e1, e1[100:160]])
    # a star (vertex 0 of degree 300: 64-bit masks and generic waves) on a path
    path = np.stack([np.arange(nv - 1), np.arange(1, nv)], axis=1)
    star = np.stack([np.zeros(300, np.int64), rng.choice(np.arange(1, nv), 300, replace=False)],
                    axis=1)
    e2 = np.concatenate([path, star])
    # edges between even vertices only: every other row empty
    ev = np.arange(0, nv - 2, 2)
    e3 = np.stack([ev, ev + 2], axis=1)
    # degree 40 everywhere (64-bit masks, all waves fast)
    e4 = np.stack([np.repeat(np.arange(nv), 20), (np.repeat(np.arange(nv), 20) +
                                                  np.tile(np.arange(1, 21), nv)) % nv], axis=1)
    return nv, {"random": e1, "star_path": e2, "even": e3, "deg40": e4}


@pytest.mark.parametrize("graph", ["random", "star_path", "even", "deg40"])
@pytest.mark.parametrize("bdim", [1, 2, 3])
def test_edge_block_routes(ctx, graph, bdim):
    nv, graphs = edge_graphs()
    ep = to_dev(graphs[graph].astype(np.int32))
    w = to_dev(np.random.default_rng(bdim).standard_normal((ep.shape[0], bdim * bdim)))
    idx = ctx.build_index(nv, [ep])
    for sign, alpha, beta in [((1, -1, -1, 1), 0.0, 1.0), ((0.5, 2.0, -1.0, 3.0), 1.5, -0.25)]:
        def run():
            vals = nan_like(idx.nnz * bdim * bdim, torch.float64)
            ctx.assemble_edge_blocks(idx, ep, w, bdim, sign=sign, alpha=alpha, beta=beta,
                                     precision=S.F64, vals=vals)
            return [vals]
        assert_same(*both_routes(run))
