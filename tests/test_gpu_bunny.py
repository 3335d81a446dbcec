"""An irregular mesh through the hot path: the reference's tetgen bunny
(apps/data/tet-bunny; 3,607 vertices, 12,687 tets, 19,233 edges).  Index
bit-exact against the reference's own builder (test_gpu_index runs the
ref_index_bunny_* fixtures), FP64 tet and spring assembly bit-exact against
the oracle, certified solves, and the learned correction on a mesh of a kind
it was never trained on."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest
import torch

import gnn as G
import simit_amd as S
from conftest import to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu
BUNNY = Path(__file__).resolve().parent / "golden" / "tet-bunny"


def bunny_problem(E=1e5):
    m = meshes.load_tetgen(BUNNY / "bunny.1.node", BUNNY / "bunny.1.ele", BUNNY / "bunny.1.edge")
    rng = np.random.default_rng(5)
    Ev = E * 10.0 ** rng.normal(0.0, 0.5, size=m["elements"].shape[0])
    return m, meshes.tet_mesh_problem(m["x"], m["elements"], E=Ev)


def tet_system(ctx, prob):
    d = {k: to_dev(prob[k]) for k in ("x", "c", "mu", "lam", "endpoints")}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    P = S.TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
    return d, idx, P


def test_bunny_tet_assembly_bit_exact_and_solve(ctx, oracle):
    _, prob = bunny_problem()
    d, idx, P = tet_system(ctx, prob)
    rp, ci = idx.to_host()
    orp, oci = oracle.index_vev(prob["nverts"], [prob["endpoints"]])
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(ci, oci)
    v64 = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(prob["nverts"] * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_tets(idx, P, d, d, S.F64, v64, None, rhs)
    ov, of = oracle.tet_assemble(oracle.tet_params(load=prob["load"]), prob, rp, ci)
    np.testing.assert_array_equal(v64.cpu().numpy(), ov)
    np.testing.assert_array_equal(rhs.cpu().numpy(), of)
    hi = torch.empty_like(v64, dtype=torch.float32)
    lo = torch.empty_like(hi)
    ctx.assemble_tets(idx, P, d, d, S.F32X2, hi, lo, rhs)
    for mixed in (0, 1):
        A = idx.bsr(3, S.F32X2, hi, lo) if mixed else idx.bsr(3, S.F64, v64)
        x = torch.zeros_like(rhs)
        st = ctx.cg(A, rhs, x, ctx.cg_options(maxiters=20000, tol=1e-8, rel=1, mixed=mixed,
                                              precond=S.PRECOND_BLOCK_JACOBI))
        assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
        r = of - oracle.bsrmv(rp, ci, ov, x.cpu().numpy(), 3)
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(of)


def test_bunny_springs_on_the_edge_set(ctx, oracle):
    m, prob = bunny_problem()
    x, ed = m["x"], m["edges"]
    nv = x.shape[0]
    rng = np.random.default_rng(2)
    # springs at rest length (the edges' lengths): each spring's Jacobian is
    # k n n^T, so A = M + h^2 K is SPD (a jitter on edges as short as 0.006
    # compresses some by 24% and makes A indefinite)
    pts = {"x": x.copy(), "v": np.zeros_like(x),
           "m": np.ones(nv), "fixed": prob["c"].astype(np.uint8),
           "k": 10.0 ** rng.uniform(2, 4, ed.shape[0]),
           "l0": np.linalg.norm(x[ed[:, 0]] - x[ed[:, 1]], axis=1), "endpoints": ed}
    d = {k: to_dev(pts[k]) for k in pts}
    idx = ctx.build_index(nv, [d["endpoints"]])
    rp, ci = idx.to_host()
    P = S.SpringsParams(0.01, 0.0, (C.c_double * 3)(0.0, 0.0, -9.8), 1.0, 1)
    vals = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(nv * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_springs(idx, P, d, d, S.F64, vals, None, rhs)
    ov, of = oracle.springs_assemble(oracle.springs_params(), pts, pts, rp, ci)
    np.testing.assert_array_equal(vals.cpu().numpy(), ov)
    np.testing.assert_array_equal(rhs.cpu().numpy(), of)
    xs = torch.zeros_like(rhs)
    st = ctx.cg(idx.bsr(3, S.F64, vals), rhs, xs, ctx.cg_options(maxiters=2000, tol=1e-8, rel=1))
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()


def test_bunny_learned_correction(ctx, oracle):
    # a network trained on structured 16^3 x 5 tet grids, applied to the
    # bunny: the spectral guard keeps it SPD and the solve certifies
    from simit_amd import learned
    _, prob = bunny_problem()
    d, idx, P = tet_system(ctx, prob)
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    rhs = torch.empty(prob["nverts"] * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_tets(idx, P, d, d, S.F32X2, hi, lo, rhs)
    A = idx.bsr(3, S.F32X2, hi, lo)
    net = ctx.gnn(seed=0, kappa=0.9)
    learned.train(ctx, net, n=16, nsamples=150)
    linv, corr = net.precond(A)
    info = net.precond_info()
    assert 0 < info["scale"] <= 1.0
    x = torch.zeros_like(rhs)
    st = ctx.cg(A, rhs, x, ctx.cg_options(maxiters=20000, tol=1e-8, rel=1, mixed=1,
                                          precond=S.PRECOND_LEARNED, learned=linv,
                                          learned_corr=corr))
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    x2 = torch.zeros_like(rhs)
    sb = ctx.cg(A, rhs, x2, ctx.cg_options(maxiters=20000, tol=1e-8, rel=1, mixed=1,
                                           precond=S.PRECOND_BLOCK_JACOBI))
    print(f"bunny: block-Jacobi {sb.iters}, learned {st.iters}, guard {info}")
    net.close()
