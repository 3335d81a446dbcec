"""Blocked SpMV, dot and axpby against the oracle and the reference fixtures."""
import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def random_bsr(ctx, oracle, kind, n, seed=0):
    if kind == "tet":
        x, ep = meshes.tet5_grid(n)
        nv, bd = x.shape[0], 3
    elif kind == "box":
        nv, ep = meshes.create_box(n, n, n)
        bd = 3
    elif kind.startswith("ragged"):
        # random 4-vertex elements over part of the vertices: empty rows,
        # rows from 1 to ~100 blocks
        rng0 = np.random.default_rng(n)
        nv = 40 * n
        ep = rng0.integers(0, nv - 7, size=(30 * n, 4)).astype(np.int32)
        ep[:, 0] = np.where(rng0.uniform(size=len(ep)) < 0.1, 3, ep[:, 0])  # a hub row
        bd = 1 if kind == "ragged1" else 3
    else:
        x, ep = meshes.hex_grid(n)
        nv, bd = x.shape[0], 1
    idx = ctx.build_index(nv, [to_dev(ep)])
    rp, ci = idx.to_host()
    rng = np.random.default_rng(seed)
    vals = rng.standard_normal(len(ci) * bd * bd)
    xv = rng.standard_normal(nv * bd)
    return idx, rp, ci, vals, xv, bd


@pytest.mark.parametrize("kind,n", [("box", 10), ("tet", 8), ("hex", 10)])
def test_exact_order_fp64_bit_exact(ctx, oracle, kind, n):
    idx, rp, ci, vals, xv, bd = random_bsr(ctx, oracle, kind, n)
    dv, dx = to_dev(vals), to_dev(xv)
    y = torch.empty(len(xv), dtype=torch.float64, device="cuda")
    ctx.bsrmv(idx.bsr(bd, S.F64, dv), dx, y, exact_order=True)
    np.testing.assert_array_equal(y.cpu().numpy(), oracle.bsrmv(rp, ci, vals, xv, bd))


@pytest.mark.parametrize("kind,n", [("box", 10), ("tet", 8), ("hex", 10), ("tet", 40),
                                    ("ragged3", 20), ("ragged1", 20)])
def test_fast_spmv_all_precisions(ctx, oracle, kind, n):
    idx, rp, ci, vals, xv, bd = random_bsr(ctx, oracle, kind, n)
    ref = oracle.bsrmv(rp, ci, vals, xv, bd)
    absrow = oracle.bsrmv(rp, ci, np.abs(vals), np.abs(xv), bd)  # error scale per row
    y = torch.empty(len(xv), dtype=torch.float64, device="cuda")
    if np.diff(rp).max() > (128 if bd == 3 else 512):
        # rows longer than half a tile are outside the streaming kernels:
        # a clean error, and the exact-order kernel still serves FP64
        with pytest.raises(S.SimitError) as e:
            ctx.bsrmv(idx.bsr(bd, S.F64, to_dev(vals)), to_dev(xv), y)
        assert e.value.code == S.ERR_UNSUPPORTED
        ctx.bsrmv(idx.bsr(bd, S.F64, to_dev(vals)), to_dev(xv), y, exact_order=True)
        np.testing.assert_array_equal(y.cpu().numpy(), ref)
        return
    # FP64
    ctx.bsrmv(idx.bsr(bd, S.F64, to_dev(vals)), to_dev(xv), y)
    assert (np.abs(y.cpu().numpy() - ref) <= 1e-14 * absrow).all()
    # FP32 values and vectors, FP32 accumulation: within n*u32 of the row
    v32, x32 = vals.astype(np.float32), xv.astype(np.float32)
    y32 = torch.empty(len(xv), dtype=torch.float32, device="cuda")
    ctx.bsrmv(idx.bsr(bd, S.F32, to_dev(v32)), to_dev(x32), y32)
    ref32 = oracle.bsrmv(rp, ci, v32.astype(np.float64), x32.astype(np.float64), bd)
    assert (np.abs(y32.cpu().numpy() - ref32) <= 128 * 2.0 ** -24 * absrow + 1e-30).all()
    # FP32 values, FP64 vectors and accumulation (SMH_F32D)
    yd = torch.empty(len(xv), dtype=torch.float64, device="cuda")
    ctx.bsrmv(idx.bsr(bd, S.F32D, to_dev(v32)), to_dev(xv), yd)
    refd = oracle.bsrmv(rp, ci, v32.astype(np.float64), xv, bd)
    assert (np.abs(yd.cpu().numpy() - refd) <= 1e-14 * absrow + 1e-300).all()
    # hi + lo split of the FP64 values, FP64 vectors
    hi = vals.astype(np.float32)
    lo = (vals - hi.astype(np.float64)).astype(np.float32)
    yh = torch.empty(len(xv), dtype=torch.float64, device="cuda")
    ctx.bsrmv(idx.bsr(bd, S.F32X2, to_dev(hi), to_dev(lo)), to_dev(xv), yh)
    assert (np.abs(yh.cpu().numpy() - ref) <= 2.0 ** -45 * absrow).all()


def planar_host(v32, nnz):
    """The lane-planar layout of FP32 3x3 values restated on the host: groups
    of 64 blocks, entries 0..3 of the group's blocks, then 4..7, then 8."""
    G = (nnz + 63) // 64
    A = np.zeros((G * 64, 9), dtype=np.float32)
    A[:nnz] = v32.reshape(nnz, 9)
    A = A.reshape(G, 64, 9)
    return np.concatenate([A[:, :, 0:4].reshape(G, 256), A[:, :, 4:8].reshape(G, 256),
                           A[:, :, 8]], axis=1).ravel()


@pytest.mark.parametrize("kind,n", [("box", 10), ("tet", 8), ("tet", 40), ("ragged3", 20),
                                    ("ragged3", 2), ("box", 2)])
def test_planar_spmv(ctx, oracle, kind, n):
    """SMH_F32P (the layout the mixed CG's inner SpMV streams) gives the same
    products as the row-major FP32 tiles, bit for bit: each block row is
    summed in the same order and each row's blocks in ascending order."""
    idx, rp, ci, vals, xv, bd = random_bsr(ctx, oracle, kind, n)
    assert bd == 3
    nnz = len(ci)
    v32, x32 = vals.astype(np.float32), xv.astype(np.float32)
    A = idx.bsr(3, S.F32, to_dev(v32))
    P = ctx.bsr_planar(A)
    pv = P._keep[0].cpu().numpy()
    np.testing.assert_array_equal(pv, planar_host(v32, nnz))
    y32 = torch.empty(len(xv), dtype=torch.float32, device="cuda")
    yp = torch.full_like(y32, float("nan"))
    ctx.bsrmv(P, to_dev(x32), yp)
    if np.diff(rp).max() <= 168:
        ctx.bsrmv(A, to_dev(x32), y32)
        assert torch.equal(yp, y32)
    else:
        # planar tiles hold 512 blocks (rows up to 256); row-major ones 336
        with pytest.raises(S.SimitError) as e:
            ctx.bsrmv(A, to_dev(x32), y32)
        assert e.value.code == S.ERR_UNSUPPORTED
    ref32 = oracle.bsrmv(rp, ci, v32.astype(np.float64), x32.astype(np.float64), 3)
    absrow = oracle.bsrmv(rp, ci, np.abs(v32).astype(np.float64),
                          np.abs(x32).astype(np.float64), 3)
    assert (np.abs(yp.cpu().numpy() - ref32) <= 128 * 2.0 ** -24 * absrow + 1e-30).all()
    # the planar layout holds 3x3 blocks and has no exact-order form
    with pytest.raises(S.SimitError) as e:
        ctx.bsrmv(P, to_dev(x32), yp, exact_order=True)
    assert e.value.code == S.ERR_INVALID


def test_planar_rejects_other_values(ctx, oracle):
    idx, rp, ci, vals, xv, bd = random_bsr(ctx, oracle, "hex", 4)
    with pytest.raises(S.SimitError) as e:
        ctx.bsr_planar(idx.bsr(1, S.F32, to_dev(vals.astype(np.float32))))
    assert e.value.code == S.ERR_INVALID
    idx, rp, ci, vals, xv, bd = random_bsr(ctx, oracle, "tet", 3)
    with pytest.raises(S.SimitError) as e:
        ctx.bsr_planar(idx.bsr(3, S.F64, to_dev(vals)))
    assert e.value.code == S.ERR_INVALID


def test_gemv_blocked_fixture(ctx, oracle, fixtures):
    fx = fixtures["gemv_blocked"]
    ep = np.array(fx["springs"], np.int32)
    idx = ctx.build_index(3, [to_dev(ep)])
    rp, ci = idx.to_host()
    vals = np.zeros((len(ci), 4))
    for s, (a, b) in enumerate(ep):
        for i, j in ((a, a), (a, b), (b, a), (b, b)):
            vals[oracle.loc(rp, ci, int(i), int(j))] += fx["a"][s]
    y = torch.empty(6, dtype=torch.float64, device="cuda")
    ctx.bsrmv(idx.bsr(2, S.F64, to_dev(vals.ravel())), to_dev(np.array(fx["b"]).ravel()), y,
              exact_order=True)
    np.testing.assert_array_equal(y.cpu().numpy().reshape(3, 2), fx["expected_c"])


def test_dot_fixtures_and_determinism(ctx, fixtures):
    for k in ("vector_dot", "vector_dot_blocked", "vector_dot_intrinsic"):
        x = to_dev(np.array(fixtures[k]["x"]))
        assert ctx.dot(x, x) == fixtures[k]["expected"]
    rng = np.random.default_rng(0)
    a = rng.standard_normal(3_000_001)
    b = rng.standard_normal(3_000_001)
    da, db = to_dev(a), to_dev(b)
    d1, d2 = ctx.dot(da, db), ctx.dot(da, db)
    assert d1 == d2
    assert d1 == pytest.approx(float(np.dot(a, b)), rel=1e-12)
    f = ctx.dot(da.float(), db.float())
    assert f == pytest.approx(float(np.dot(a.astype(np.float32).astype(np.float64),
                                           b.astype(np.float32).astype(np.float64))), rel=1e-12)


def test_axpby(ctx):
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(1001), rng.standard_normal(1001)
    dx, dy = to_dev(x), to_dev(y)
    ctx.axpby(0.5, dx, -2.0, dy)
    np.testing.assert_allclose(dy.cpu().numpy(), 0.5 * x - 2.0 * y, rtol=1e-15, atol=1e-15)
