"""Scalar 27-point stencil storage (the reference's TensorStorage::Stencil,
src/storage.h:48, lowered at lower_indexexprs.cpp:669-711) against the
blocked-CSR path it replaces: the product bit for bit in every precision,
the CG to the same certified residual, the learned correction through the
converted storage."""
import numpy as np
import pytest
import torch

import gnn as G
import simit_amd as S
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def hex_system(ctx, n, seed):
    prob = meshes.hex_problem(n, seed=seed, std_log10=1.5, corr_len=3.0)
    d = {k: to_dev(prob[k]) for k in ("dirichlet", "kappa", "endpoints")}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    P = S.HexParams(prob["h"], prob["source"], 1.0)
    v64 = torch.empty(idx.nnz, dtype=torch.float64, device="cuda")
    hi = torch.empty(idx.nnz, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    rhs = torch.empty(prob["nverts"], dtype=torch.float64, device="cuda")
    ctx.assemble_hex(idx, P, d, d, S.F32X2, hi, lo, rhs)
    ctx.assemble_hex(idx, P, d, d, S.F64, v64, None, rhs)
    torch.cuda.synchronize()
    return {"idx": idx, "v64": v64, "hi": hi, "lo": lo, "rhs": rhs, "dims": (n + 1,) * 3}


def host_bsr(rp, ci, vals, precision, lo=None):
    """A 1x1 Bsr over host-built CSR arrays (kept alive by the struct)."""
    rpd, cid = to_dev(rp.astype(np.uint32)), to_dev(ci.astype(np.uint32))
    v = to_dev(vals)
    vl = to_dev(lo) if lo is not None else None
    A = S.Bsr(rp.size - 1, 1, ci.size, rpd.data_ptr(), cid.data_ptr(), precision, v.data_ptr(),
              vl.data_ptr() if vl is not None else None)
    A._keep = (rpd, cid, v, vl)
    return A


def offsets(n0, n1):
    return [d0 + n0 * d1 + n0 * n1 * d2 for d2 in (-1, 0, 1) for d1 in (-1, 0, 1)
            for d0 in (-1, 0, 1)]


@pytest.mark.parametrize("prec", ["F64", "F32", "F32D", "F32X2"])
def test_stencil_product_equals_csr_bitwise(ctx, prec):
    sy = hex_system(ctx, 11, seed=1)
    idx = sy["idx"]
    p = getattr(S, prec)
    if prec == "F64":
        A = idx.bsr(1, p, sy["v64"])
    elif prec == "F32X2":
        A = idx.bsr(1, p, sy["hi"], sy["lo"])
    else:
        A = idx.bsr(1, p, sy["hi"])
    St = ctx.stencil27_from_bsr(A, sy["dims"])
    g = torch.Generator(device="cpu").manual_seed(3)
    dt = torch.float32 if prec == "F32" else torch.float64
    x = torch.randn(idx.nrows, generator=g, dtype=torch.float64).to(dt).cuda()
    y1, y2 = torch.empty_like(x), torch.empty_like(x)
    ctx.bsrmv(A, x, y1)
    ctx.stencil27_mv(St, x, y2)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2), (y1 - y2).abs().max().item()


def test_stencil_product_random_pattern_odd_grid(ctx, oracle):
    # any matrix whose columns stay within the 27 offsets: random subsets of
    # them (wrapping across grid lines included) on a 5 x 4 x 3 grid
    n0, n1, n2 = 5, 4, 3
    n = n0 * n1 * n2
    rng = np.random.default_rng(7)
    offs = offsets(n0, n1)
    rp, ci = [0], []
    for i in range(n):
        cols = sorted({i + o for o in offs if 0 <= i + o < n and rng.random() < 0.6})
        ci += cols
        rp.append(len(ci))
    rp, ci = np.array(rp), np.array(ci)
    vals = rng.normal(size=ci.size)
    A = host_bsr(rp, ci, vals, S.F64)
    St = ctx.stencil27_from_bsr(A, (n0, n1, n2))
    x = torch.from_numpy(rng.normal(size=n)).cuda()
    y1, y2 = torch.empty_like(x), torch.empty_like(x)
    ctx.bsrmv(A, x, y1)
    ctx.stencil27_mv(St, x, y2)
    assert torch.equal(y1, y2)
    ye = oracle.bsrmv(rp.astype(np.uint32), ci.astype(np.uint32), vals, x.cpu().numpy(), 1)
    np.testing.assert_allclose(y2.cpu().numpy(), ye, rtol=1e-13, atol=1e-13)


def test_stencil_rejects_a_column_outside(ctx):
    n0, n1, n2 = 4, 4, 4
    rp = np.array([0] + [2] * 1 + [2 + 1 * k for k in range(1, 64)])
    ci = np.array([0, 2] + list(range(1, 64)))  # row 0 couples with column 2 (offset +2)
    A = host_bsr(rp, ci, np.ones(ci.size), S.F64)
    with pytest.raises(S.SimitError, match="outside the 27-point stencil"):
        ctx.stencil27_from_bsr(A, (n0, n1, n2))


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_stencil_cg_matches_csr(ctx, oracle, mixed):
    sy = hex_system(ctx, 15, seed=2)
    idx = sy["idx"]
    A = idx.bsr(1, S.F32X2, sy["hi"], sy["lo"]) if mixed else idx.bsr(1, S.F64, sy["v64"])
    St = ctx.stencil27_from_bsr(A, sy["dims"])
    opts = ctx.cg_options(maxiters=5000, tol=1e-8, rel=1, precond=S.PRECOND_JACOBI, mixed=mixed)
    x1, x2 = torch.zeros_like(sy["rhs"]), torch.zeros_like(sy["rhs"])
    s1 = ctx.cg(A, sy["rhs"], x1, opts)
    s2 = ctx.cg_stencil27(St, sy["rhs"], x2, opts)
    assert s1.status == 0 and s2.status == 0 and s2.true_rel <= 1e-8, (s1.as_dict(), s2.as_dict())
    # the same recurrences; only the grouping of the dot-product partial sums differs
    assert abs(s1.iters - s2.iters) <= max(2, s1.iters // 50), (s1.iters, s2.iters)
    rp, ci = idx.to_host()
    f = sy["rhs"].cpu().numpy()
    r = f - oracle.bsrmv(rp, ci, sy["v64"].cpu().numpy(), x2.cpu().numpy(), 1)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(f)


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_stencil_learned_correction(ctx, mixed):
    # the correction in FP16 stencil slots: one iteration from 0 gives
    # x1 = alpha B b with B = Linv (I + C16) Linv, C16 the half-rounded C
    sy = hex_system(ctx, 9, seed=3)
    idx = sy["idx"]
    A = idx.bsr(1, S.F32X2, sy["hi"], sy["lo"]) if mixed else idx.bsr(1, S.F64, sy["v64"])
    net = ctx.gnn(seed=1, kappa=0.9)
    p = G.init_params(5, dec_scale=0.5)
    net.set_params(G.flatten(p).astype(np.float32))
    linv, corr = net.precond(A)
    c16 = ctx.stencil27_correction(idx, corr, sy["dims"])
    St = ctx.stencil27_from_bsr(A, sy["dims"])
    o = ctx.cg_options(maxiters=1, tol=1e-30, rel=1, fallback=0, precond=S.PRECOND_LEARNED,
                       mixed=mixed, learned=linv, learned_corr=c16)
    x = torch.zeros_like(sy["rhs"])
    ctx.cg_stencil27(St, sy["rhs"], x, o)
    rp, ci = idx.to_host()
    n = rp.size - 1
    cv = corr.view(torch.float32).cpu().numpy().astype(np.float16).astype(np.float64)
    Li = linv.cpu().numpy().astype(np.float64)
    brow = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    Bk = Li[brow] * (cv + (brow == ci)) * Li[ci]
    import scipy.sparse as sp
    Bm = sp.csr_matrix((Bk, ci.astype(np.int64), rp.astype(np.int64)), shape=(n, n))
    z = Bm @ sy["rhs"].cpu().numpy()
    xg = x.cpu().numpy()
    err = np.linalg.norm(xg / np.linalg.norm(xg) - z / np.linalg.norm(z))
    # mixed 1 and 2 gather u = Linv r in FP32 (mixed 2 keeps FP64 r and z)
    assert err <= (1e-6 if mixed else 1e-12), err
    # the slots hold the half-rounded correction, exactly
    slots = c16.view(torch.float16).cpu().numpy().astype(np.float64).reshape(27, n)
    offs = offsets(*sy["dims"][:2])
    for k in range(0, ci.size, max(1, ci.size // 500)):
        i, j = brow[k], ci[k]
        assert slots[offs.index(int(j) - int(i)), i] == cv[k]
    # and the full solve
    o.maxiters = 5000
    o.tol = 1e-8
    x.zero_()
    st = ctx.cg_stencil27(St, sy["rhs"], x, o)
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    net.close()


@pytest.mark.parametrize("dims", [(3, 1, 1), (3, 2, 1), (5, 4, 3), (2, 2, 2), (1, 1, 1), (17, 9, 6)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_periodic_stencil_product(ctx, dims, prec):
    """Stencil storage with SMH_STENCIL_PERIODIC (the matrices of maps
    through a grid edge set): neighbour coordinates wrap around, every slot
    is its own term (slots that reach the same point on a grid of size 1 or
    2 are not merged), the 27 products summed in slot order from 0."""
    import ctypes as C
    n0, n1, n2 = dims
    n = n0 * n1 * n2
    rng = np.random.default_rng(sum(dims))
    dt = np.float64 if prec == "f64" else np.float32
    vals = rng.standard_normal((27, n)).astype(dt)
    x = rng.standard_normal(n).astype(dt)
    i = np.arange(n)
    c = np.stack([i % n0, (i // n0) % n1, i // (n0 * n1)])
    want = np.zeros(n, dtype=dt)
    for s in range(27):
        o = np.array([s % 3 - 1, s // 3 % 3 - 1, s // 9 - 1])
        w = (c + o[:, None]) % np.array(dims)[:, None]
        j = w[0] + n0 * (w[1] + n1 * w[2])
        want = (want + vals[s] * x[j]).astype(dt)
    dv, dx = to_dev(vals.ravel()), to_dev(x)
    y = torch.empty_like(dx)
    S_ = S.Stencil27(n0, n1, n2, S.F64 if prec == "f64" else S.F32, dv.data_ptr(), None,
                     S.STENCIL_PERIODIC)
    ctx.stencil27_mv(S_, dx, y)
    np.testing.assert_array_equal(y.cpu().numpy(), want)
    # the solver refuses a periodic stencil
    with pytest.raises(S.SimitError):
        ctx.cg_stencil27(S_, to_dev(x.astype(np.float64)), torch.zeros(n, dtype=torch.float64,
                                                                        device="cuda"),
                         ctx.cg_options(maxiters=5))
