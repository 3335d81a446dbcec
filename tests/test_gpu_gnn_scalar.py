"""The learned correction on scalar (1x1) systems — the hex Poisson matrix —
on the GPU against the CPU oracle (oracle/gnn.py with bd = 1, pinned by the
scalar finite-difference and dense-formula tests in test_gnn_oracle.py).

As for 3x3 blocks, the oracle runs with mimic=True (BF16 where the kernels
store or feed BF16), and the tolerances allow the one-BF16-step differences
FP32 accumulation can cause.  For 1x1 blocks the correction is stored in
FP32, so the stored C is compared directly.
"""
import numpy as np
import pytest
import torch

import gnn as G
import simit_amd as S
from simit_amd import meshes
from test_gpu_gnn import check_guard, dense_blocks, learned_blocks

pytestmark = pytest.mark.gpu


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def hex_system(ctx, n, seed, corr=3.0, std=1.5):
    prob = meshes.hex_problem(n, seed=seed, std_log10=std, corr_len=corr)
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
    return {"idx": idx, "v64": v64, "hi": hi, "lo": lo, "rhs": rhs, "dev": d}


def random_params(seed, scale=1.0):
    p = G.init_params(seed, dec_scale=1.0)
    rng = np.random.default_rng(seed + 100)
    for k in p:
        p[k] = rng.normal(0, 0.1, p[k].shape) if k.startswith("b") else p[k] * scale
    return G.flatten(p).astype(np.float32)


def corr_values(corr):
    """The [nnz*4] byte buffer of a scalar system's correction as float64."""
    return corr.view(torch.float32).cpu().numpy().astype(np.float64)


# n = 17: 140,608 entries, so the edge kernels run the full-chip grid (2,048
# workgroups, a multiple of 8) and take their tiles by XCD; n = 6 runs the
# small grid of plain strided tiles.  The dense spectral check is for n = 6.
@pytest.mark.parametrize("n", [6, 17])
def test_scalar_forward_matches_oracle(ctx, n):
    sy = hex_system(ctx, n, seed=1)
    idx = sy["idx"]
    rp, ci = idx.to_host()
    vals = sy["v64"].cpu().numpy()
    net = ctx.gnn(seed=3, kappa=0.9)
    flat = random_params(5)
    net.set_params(flat)
    linv, corr = net.precond(idx.bsr(1, S.F64, sy["v64"]))
    torch.cuda.synchronize()
    assert linv.numel() == rp.size - 1 and corr.numel() == 4 * len(ci)
    B, aux = G.precond_blocks(G.unflatten(flat.astype(np.float64)), rp, ci, vals, kappa=0.9,
                              mimic=True, bd=1)
    Lo = aux["ft"]["Linv"].ravel()
    np.testing.assert_allclose(linv.cpu().numpy(), Lo, rtol=2e-6)
    info = net.precond_info()
    Cg, Co, g = corr_values(corr), info["scale"] * aux["C"].ravel(), aux["g"]
    assert np.array_equal(Cg, Cg[g["t"]])           # exactly symmetric
    assert np.all(Cg[g["diag"]] == 0.0)
    scale = np.abs(Co).max()
    assert scale > 1e-3
    assert np.linalg.norm(Cg - Co) <= 5e-2 * np.linalg.norm(Co), np.linalg.norm(Cg - Co)
    assert np.abs(Cg - Co).max() <= 0.07 * scale
    if n <= 8:
        check_guard(info, dense_blocks(rp, ci, aux["C"].ravel()), dense_blocks(rp, ci, Cg), 0.9)
    net.close()


def test_scalar_gradients_match_oracle(ctx):
    sy = hex_system(ctx, 4, seed=2)
    idx = sy["idx"]
    rp, ci = idx.to_host()
    vals = sy["v64"].cpu().numpy()
    n = rp.size - 1
    net = ctx.gnn(seed=0, kappa=0.9)
    flat = random_params(7, scale=0.8)
    net.set_params(flat)
    rng = np.random.default_rng(3)
    V = rng.normal(size=(n, 4)).astype(np.float32)
    L, g = net.loss_grad(idx.bsr(1, S.F64, sy["v64"]), nprobes=4, probes=to_dev(V))
    Lo, go, _ = G.loss_and_grads(G.unflatten(flat.astype(np.float64)), rp, ci, vals,
                                 V.astype(np.float64), kappa=0.9, mimic=True, bd=1)
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
    net.close()


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_scalar_learned_pcg_converges(ctx, oracle, mixed):
    sy = hex_system(ctx, 12, seed=4)
    idx = sy["idx"]
    net = ctx.gnn(seed=2, kappa=0.9)
    net.set_params(random_params(9))
    A = idx.bsr(1, S.F32X2, sy["hi"], sy["lo"]) if mixed else idx.bsr(1, S.F64, sy["v64"])
    linv, corr = net.precond(A)
    x = torch.zeros_like(sy["rhs"])
    st = ctx.cg(A, sy["rhs"], x, ctx.cg_options(maxiters=5000, tol=1e-8, rel=1,
                                                precond=S.PRECOND_LEARNED, mixed=mixed,
                                                learned=linv, learned_corr=corr))
    assert st.status == 0 and st.true_rel <= 1e-8, st.as_dict()
    rp, ci = idx.to_host()
    vals = sy["v64"].cpu().numpy()
    f = sy["rhs"].cpu().numpy()
    r = f - oracle.bsrmv(rp, ci, vals, x.cpu().numpy(), 1)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(f)
    if mixed == 0:
        # the same PCG on the CPU with B = L (I + C) L from the stored buffers
        import scipy.sparse as sp
        nr = rp.size - 1
        Li = linv.cpu().numpy().astype(np.float64)
        Cv = corr_values(corr)
        brow = np.repeat(np.arange(nr), np.diff(rp.astype(np.int64)))
        Bk = Li[brow] * (Cv + (brow == ci)) * Li[ci]
        Bm = sp.csr_matrix((Bk, ci.astype(np.int64), rp.astype(np.int64)), shape=(nr, nr))
        Am = sp.csr_matrix((vals, ci.astype(np.int64), rp.astype(np.int64)), shape=(nr, nr))
        xx, rr = np.zeros_like(f), f.copy()
        zz = Bm @ rr
        pp, rz, it = zz.copy(), rr @ zz, 0
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


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_scalar_learned_apply_matches_stored_preconditioner(ctx, mixed):
    sy = hex_system(ctx, 10, seed=6)
    idx = sy["idx"]
    rp, ci = idx.to_host()
    net = ctx.gnn(seed=4, kappa=0.9)
    net.set_params(random_params(12, scale=0.5))
    A = idx.bsr(1, S.F32X2, sy["hi"], sy["lo"]) if mixed else idx.bsr(1, S.F64, sy["v64"])
    linv, corr = net.precond(A)
    x = torch.zeros_like(sy["rhs"])
    ctx.cg(A, sy["rhs"], x, ctx.cg_options(maxiters=1, tol=1e-30, rel=1, fallback=0,
                                           precond=S.PRECOND_LEARNED, mixed=mixed,
                                           learned=linv, learned_corr=corr))
    Bm = learned_blocks(rp, ci, linv.cpu().numpy(), corr_values(corr).reshape(-1, 1, 1), 1)
    z = Bm @ sy["rhs"].cpu().numpy()
    xg = x.cpu().numpy()
    err = np.linalg.norm(xg / np.linalg.norm(xg) - z / np.linalg.norm(z))
    assert err <= (1e-12 if mixed != 1 else 1e-6), err
    net.close()


def test_scalar_trained_network_beats_jacobi(ctx):
    train = [hex_system(ctx, 10, seed=100 + s) for s in range(8)]
    mats = [s["idx"].bsr(1, S.F32X2, s["hi"], s["lo"]) for s in train]
    net = ctx.gnn(seed=0, kappa=0.9)
    for step in range(400):
        net.train_step(mats[step % len(mats)], nprobes=4, seed=step, lr=3e-3, want_loss=False)
    test = hex_system(ctx, 16, seed=7)
    A = test["idx"].bsr(1, S.F64, test["v64"])

    def iters(linv, corr):
        x = torch.zeros_like(test["rhs"])
        st = ctx.cg(A, test["rhs"], x, ctx.cg_options(
            maxiters=20000, tol=1e-8, rel=1, precond=S.PRECOND_LEARNED, learned=linv,
            learned_corr=corr))
        assert st.status == 0 and st.true_rel <= 1e-8
        return st.iters

    linv, corr = net.precond(A)
    it_learned = iters(linv, corr)
    it_zero = iters(linv, torch.zeros_like(corr))
    x = torch.zeros_like(test["rhs"])
    st = ctx.cg(A, test["rhs"], x, ctx.cg_options(maxiters=20000, tol=1e-8, rel=1,
                                                  precond=S.PRECOND_JACOBI))
    print(f"Jacobi {st.iters} / zero correction {it_zero} / learned {it_learned}")
    assert abs(it_zero - st.iters) <= max(3, st.iters // 50)
    assert it_learned < 0.92 * it_zero, (it_learned, it_zero)
    net.close()
