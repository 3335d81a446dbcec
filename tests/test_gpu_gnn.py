"""The learned block-Jacobi correction on the GPU against the CPU oracle.

The oracle (oracle/gnn.py, pinned by finite differences and torch autograd in
test_gnn_oracle.py) runs with mimic=True: it rounds to bfloat16 where the
kernels store or feed BF16.  The kernels accumulate in FP32 where the oracle
uses float64, so a value next to a BF16 rounding boundary can land one BF16
step (2^-8 relative) apart; the tolerances below allow for that and nothing
more.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import gnn as G
import simit_amd as S
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def tet_system(ctx, n, seed, corr=3.0):
    prob = meshes.tet_problem(n, seed=seed, field="grf", std_log10=1.0, corr_len=corr)
    d = {"x": to_dev(prob["x"]), "c": to_dev(prob["c"].astype(np.int32)),
         "mu": to_dev(prob["mu"]), "lam": to_dev(prob["lam"]),
         "endpoints": to_dev(prob["endpoints"].astype(np.int32))}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    P = S.TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
    v64 = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    rhs = torch.empty(prob["nverts"] * 3, dtype=torch.float64, device="cuda")
    ctx.assemble_tets(idx, P, d, d, S.F32X2, hi, lo, rhs)
    ctx.assemble_tets(idx, P, d, d, S.F64, v64, None, rhs)
    torch.cuda.synchronize()
    return {"idx": idx, "v64": v64, "hi": hi, "lo": lo, "rhs": rhs, "dev": d, "prob": prob}


def fp8_blocks(t):
    """Decode the [nnz*12] FP8 correction (e4m3 of 32 C, 3 values + pad per
    block row) into [nnz, 3, 3] float64."""
    b = t.cpu().numpy().reshape(-1, 3, 4)[:, :, :3].astype(np.int64)
    e, m = (b >> 3) & 15, b & 7
    v = np.where(e > 0, (1.0 + m / 8.0) * np.exp2(e - 7.0), m * 2.0 ** -9)
    return np.where(b & 0x80, -v, v) / 32.0


def fp8_round(x):
    """x rounded to the FP8 grid the kernels store (e4m3 of 32 x) / 32."""
    y = np.abs(32.0 * x)
    ex = np.floor(np.log2(np.maximum(y, 2.0 ** -6)))
    step = np.exp2(ex - 3)
    q = np.round(y / step) * step          # numpy rounds half to even
    return np.sign(x) * np.minimum(q, 448.0) / 32.0


def random_params(seed, scale=1.0):
    p = G.init_params(seed, dec_scale=1.0)
    rng = np.random.default_rng(seed + 100)
    for k in p:
        p[k] = rng.normal(0, 0.1, p[k].shape) if k.startswith("b") else p[k] * scale
    return G.flatten(p).astype(np.float32)


# n = 22: ~148k blocks, so the edge kernels run the full-chip grid (2,048
# workgroups) and take their tiles by XCD; n = 4 runs a small grid of plain
# strided tiles.  The dense spectral check is for n = 4.
@pytest.mark.parametrize("n", [4, 22])
def test_forward_matches_oracle(ctx, n):
    sy = tet_system(ctx, n, seed=1)
    idx = sy["idx"]
    rp, ci = idx.to_host()
    vals = sy["v64"].cpu().numpy()
    net = ctx.gnn(seed=3, kappa=0.9)
    flat = random_params(5)
    net.set_params(flat)
    np.testing.assert_array_equal(net.get_params(), flat)
    linv, corr = net.precond(idx.bsr(3, S.F64, sy["v64"]))
    torch.cuda.synchronize()
    B, aux = G.precond_blocks(G.unflatten(flat.astype(np.float64)), rp, ci, vals, kappa=0.9,
                              mimic=True)
    Lo = aux["ft"]["Linv"].reshape(-1, 9)
    np.testing.assert_allclose(linv.cpu().numpy().reshape(-1, 9), Lo, rtol=2e-6,
                               atol=2e-6 * np.abs(Lo).max())
    info = net.precond_info()
    Cg = fp8_blocks(corr)
    Co = info["scale"] * aux["C"]  # the spectral guard's global scale (1: none applied)
    g = aux["g"]
    # exact symmetry of the stored correction, zero diagonal blocks
    assert np.array_equal(Cg, Cg[g["t"]].transpose(0, 2, 1))
    assert np.all(Cg[g["diag"]] == 0.0)
    # FP8 storage: each value within half an e4m3 step of the oracle's value
    # rounded the same way, up to the BF16-chain noise of the network
    Cr = fp8_round(Co)
    scale = np.abs(Co).max()
    assert scale > 1e-3
    close = np.abs(Cg - Cr) <= 1e-6 * scale
    assert close.mean() > 0.95, close.mean()
    assert np.linalg.norm(Cg - Co) <= 5e-2 * np.linalg.norm(Co), np.linalg.norm(Cg - Co)
    assert np.abs(Cg - Co).max() <= 0.07 * scale
    if n <= 8:
        check_guard(info, dense_blocks(rp, ci, fp8_round(aux["C"])), dense_blocks(rp, ci, Cg),
                    0.9)
    net.close()


def dense_blocks(rp, ci, blocks):
    n = rp.size - 1
    bd = int(round(np.sqrt(blocks[0].size)))
    return sp.bsr_matrix((np.asarray(blocks).reshape(-1, bd, bd), ci.astype(np.int64),
                          rp.astype(np.int64)), shape=(n * bd, n * bd)).toarray()


def check_guard(info, C0, Cs, kappa):
    """The spectral guard: its Lanczos estimate of lambda_min against the exact
    spectrum of the network's unscaled C (C0, the oracle's on the storage
    grid: the GPU's differs by BF16-chain noise), and the stored C (Cs) keeps
    lambda_min(I + C) >= 1 - kappa."""
    ev = np.linalg.eigvalsh(C0)
    width = ev[-1] - ev[0]
    assert abs(info["lambda_min"] - ev[0]) <= 0.03 * width + 0.03 * abs(ev[0]), (info, ev[0])
    assert abs(info["spectral_width"] - width) <= 0.1 * width, (info, width)
    assert 0 < info["scale"] <= 1.0 and info["steps"] >= 1
    if ev[0] - 0.1 * width > -kappa:
        assert info["scale"] == 1.0, info
    lmin = np.linalg.eigvalsh(np.eye(Cs.shape[0]) + Cs)[0]
    assert lmin >= 1.0 - kappa - 0.02, (lmin, info)


@pytest.mark.parametrize("kappa", [0.9, 0.3])
def test_spectral_guard_scales_an_indefinite_correction(ctx, kappa):
    # large random weights saturate the blocks, so lambda_min(C) lies well
    # below -kappa and the guard must scale C down
    sy = tet_system(ctx, 4, seed=3)
    idx = sy["idx"]
    rp, ci = idx.to_host()
    vals = sy["v64"].cpu().numpy()
    net = ctx.gnn(seed=1, kappa=kappa)
    flat = random_params(8, scale=4.0)
    net.set_params(flat)
    linv, corr = net.precond(idx.bsr(3, S.F64, sy["v64"]))
    info = net.precond_info()
    _, aux = G.precond_blocks(G.unflatten(flat.astype(np.float64)), rp, ci, vals, mimic=True)
    C0 = dense_blocks(rp, ci, fp8_round(aux["C"]))
    assert np.linalg.eigvalsh(C0)[0] < -kappa
    assert info["scale"] < 1.0, info
    Cs = dense_blocks(rp, ci, fp8_blocks(corr))
    check_guard(info, C0, Cs, kappa)
    # the guarded preconditioner is SPD: learned PCG converges
    x = torch.zeros_like(sy["rhs"])
    st = ctx.cg(idx.bsr(3, S.F64, sy["v64"]), sy["rhs"], x,
                ctx.cg_options(maxiters=5000, tol=1e-8, rel=1, precond=S.PRECOND_LEARNED,
                               learned=linv, learned_corr=corr))
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    net.close()


def test_gradients_match_oracle(ctx):
    sy = tet_system(ctx, 3, seed=2)
    idx = sy["idx"]
    rp, ci = idx.to_host()
    vals = sy["v64"].cpu().numpy()
    n = rp.size - 1
    net = ctx.gnn(seed=0, kappa=0.9)
    flat = random_params(7, scale=0.8)
    net.set_params(flat)
    rng = np.random.default_rng(3)
    V = rng.normal(size=(3 * n, 4)).astype(np.float32)
    L, g = net.loss_grad(idx.bsr(3, S.F64, sy["v64"]), nprobes=4, probes=to_dev(V))
    Lo, go, _ = G.loss_and_grads(G.unflatten(flat.astype(np.float64)), rp, ci, vals,
                                 V.astype(np.float64), kappa=0.9, mimic=True)
    assert abs(L - Lo) <= 2e-3 * Lo, (L, Lo)
    gof = G.flatten(go)
    o = 0
    for name, shape in G.param_layout():
        k = int(np.prod(shape))
        a, b = g[o:o + k].astype(np.float64), gof[o:o + k]
        o += k
        nb = np.linalg.norm(b)
        if nb < 1e-7 * np.linalg.norm(gof):
            assert np.linalg.norm(a) <= 1e-5 * np.linalg.norm(gof), name
            continue
        assert np.linalg.norm(a - b) <= 3e-2 * nb, (name, np.linalg.norm(a - b) / nb)
    # the loss is what the probes say: same probes, same bits
    L2, g2 = net.loss_grad(idx.bsr(3, S.F64, sy["v64"]), nprobes=4, probes=to_dev(V))
    assert L2 == L and np.array_equal(g, g2)
    net.close()


def test_training_reduces_loss_and_is_deterministic(ctx):
    systems = [tet_system(ctx, 6, seed=s) for s in range(3)]
    mats = [s["idx"].bsr(3, S.F32X2, s["hi"], s["lo"]) for s in systems]
    losses = []
    for run in range(2):
        net = ctx.gnn(seed=11, kappa=0.9)
        first = [net.loss_grad(A, nprobes=8, seed=99, grads=False)[0] for A in mats]
        for step in range(120):
            net.train_step(mats[step % 3], nprobes=4, seed=1000 + step, lr=3e-3,
                           want_loss=False)
        last = [net.loss_grad(A, nprobes=8, seed=99, grads=False)[0] for A in mats]
        losses.append((first, last, net.get_params()))
        net.close()
    (f0, l0, p0), (f1, l1, p1) = losses
    assert np.mean(l0) < 0.9 * np.mean(f0), (f0, l0)
    assert f0 == f1 and l0 == l1 and np.array_equal(p0, p1)


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_learned_pcg_converges(ctx, oracle, mixed):
    sy = tet_system(ctx, 8, seed=4)
    idx = sy["idx"]
    net = ctx.gnn(seed=2, kappa=0.9)
    net.set_params(random_params(9))
    A = idx.bsr(3, S.F32X2, sy["hi"], sy["lo"]) if mixed else idx.bsr(3, S.F64, sy["v64"])
    linv, corr = net.precond(A)
    n = sy["rhs"].numel()
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    o = ctx.cg_options(maxiters=5000, tol=1e-8, rel=1, precond=S.PRECOND_LEARNED, mixed=mixed,
                       learned=linv, learned_corr=corr)
    st = ctx.cg(A, sy["rhs"], x, o)
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    rp, ci = idx.to_host()
    vals = sy["v64"].cpu().numpy()
    f = sy["rhs"].cpu().numpy()
    r = f - oracle.bsrmv(rp, ci, vals, x.cpu().numpy(), 3)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(f)
    if mixed == 0:
        # the same PCG on the CPU with the preconditioner decoded from the
        # stored Linv and FP8 blocks takes the same number of iterations
        import scipy.sparse as sp
        nr = rp.size - 1
        Li = linv.cpu().numpy().astype(np.float64).reshape(-1, 3, 3)
        Cb = fp8_blocks(corr)
        brow = np.repeat(np.arange(nr), np.diff(rp.astype(np.int64)))
        Bh = Cb + np.where((brow == ci)[:, None, None], np.eye(3), 0.0)
        Bk = Li[brow].transpose(0, 2, 1) @ Bh @ Li[ci]
        shape = (3 * nr, 3 * nr)
        Bm = sp.bsr_matrix((Bk, ci.astype(np.int64), rp.astype(np.int64)), shape=shape)
        Am = sp.bsr_matrix((vals.reshape(-1, 3, 3), ci.astype(np.int64), rp.astype(np.int64)),
                           shape=shape)
        xx = np.zeros_like(f)
        rr = f.copy()
        zz = Bm @ rr
        pp = zz.copy()
        rz = rr @ zz
        it = 0
        while np.linalg.norm(rr) > 1e-8 * np.linalg.norm(f) and it < 5000:
            qq = Am @ pp
            al = rz / (pp @ qq)
            xx += al * pp
            rr -= al * qq
            zz = Bm @ rr
            rzn = rr @ zz
            pp = zz + (rzn / rz) * pp
            rz = rzn
            it += 1
        assert abs(it - st.iters) <= max(2, st.iters // 50), (it, st.iters)
    net.close()


def learned_blocks(rp, ci, linv, Cb, bd):
    """The preconditioner B = Linv^T (I + C) Linv as a dense matrix from the
    stored buffers (linv float32 [n, bd, bd], C [nnz, bd, bd])."""
    n = rp.size - 1
    Li = linv.reshape(n, bd, bd).astype(np.float64)
    Lm = sp.bsr_matrix((Li, np.arange(n), np.arange(n + 1)), shape=(n * bd, n * bd)).toarray()
    return Lm.T @ (np.eye(n * bd) + dense_blocks(rp, ci, Cb)) @ Lm


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_learned_apply_matches_stored_preconditioner(ctx, mixed):
    # one CG iteration from x = 0 gives x1 = alpha B b: its direction pins the
    # correction pass (one FP8 block per lane) against B built on the CPU
    sy = tet_system(ctx, 5, seed=6)
    idx = sy["idx"]
    rp, ci = idx.to_host()
    net = ctx.gnn(seed=4, kappa=0.9)
    net.set_params(random_params(12, scale=0.5))
    A = idx.bsr(3, S.F32X2, sy["hi"], sy["lo"]) if mixed else idx.bsr(3, S.F64, sy["v64"])
    linv, corr = net.precond(A)
    x = torch.zeros_like(sy["rhs"])
    ctx.cg(A, sy["rhs"], x, ctx.cg_options(maxiters=1, tol=1e-30, rel=1, fallback=0,
                                           precond=S.PRECOND_LEARNED, mixed=mixed,
                                           learned=linv, learned_corr=corr))
    Bm = learned_blocks(rp, ci, linv.cpu().numpy(), fp8_blocks(corr), 3)
    z = Bm @ sy["rhs"].cpu().numpy()
    xg = x.cpu().numpy()
    cos = (xg @ z) / (np.linalg.norm(xg) * np.linalg.norm(z))
    err = np.linalg.norm(xg / np.linalg.norm(xg) - z / np.linalg.norm(z))
    # FP64 vectors: rounding only; FP32 vectors: the float32 r, u and z
    assert err <= (1e-12 if mixed != 1 else 1e-6), (err, cos)
    net.close()


def test_learned_pcg_rejects_missing_buffers(ctx):
    sy = tet_system(ctx, 3, seed=0)
    A = sy["idx"].bsr(3, S.F64, sy["v64"])
    x = torch.zeros_like(sy["rhs"])
    with pytest.raises(S.SimitError):
        ctx.cg(A, sy["rhs"], x, ctx.cg_options(precond=S.PRECOND_LEARNED))


def test_trained_network_beats_block_jacobi(ctx):
    # ablation: the same PCG with the trained correction against block Jacobi
    # alone (the network with its output zeroed) on a held-out system
    train = [tet_system(ctx, 10, seed=100 + s, corr=4.0) for s in range(8)]
    mats = [s["idx"].bsr(3, S.F32X2, s["hi"], s["lo"]) for s in train]
    net = ctx.gnn(seed=0, kappa=0.9)
    for step in range(400):
        net.train_step(mats[step % len(mats)], nprobes=4, seed=step, lr=3e-3, want_loss=False)
    test = tet_system(ctx, 16, seed=7, corr=4.0)
    A = test["idx"].bsr(3, S.F64, test["v64"])
    n = test["rhs"].numel()

    def iters(linv, corr):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        st = ctx.cg(A, test["rhs"], x, ctx.cg_options(
            maxiters=20000, tol=1e-8, rel=1, precond=S.PRECOND_LEARNED, learned=linv,
            learned_corr=corr))
        assert st.status == 0 and st.true_rel <= 1e-8
        return st.iters

    linv, corr = net.precond(A)
    it_learned = iters(linv, corr)
    zero = torch.zeros_like(corr)
    it_bj = iters(linv, zero)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    st = ctx.cg(A, test["rhs"], x, ctx.cg_options(maxiters=20000, tol=1e-8, rel=1,
                                                  precond=S.PRECOND_BLOCK_JACOBI))
    print(f"block Jacobi {st.iters} / zero correction {it_bj} / learned {it_learned}")
    assert abs(it_bj - st.iters) <= max(3, st.iters // 50)
    assert it_learned < 0.92 * it_bj, (it_learned, it_bj)
    net.close()
