"""Pins the CPU oracle of the learned preconditioner (oracle/gnn.py).

The reference has no learned part, so these checks stand on mathematics:
central finite differences and torch-CPU autograd for the gradients, dense
linear algebra for the loss and the preconditioner, and the structural
promises (symmetry, positive definiteness, block Jacobi at zero output).
"""
import numpy as np
import pytest

import gnn as G
from simit_amd import meshes


def small_system(oracle, n=3, seed=0):
    prob = meshes.tet_problem(n, seed=seed, field="grf", std_log10=1.0, corr_len=2.0)
    rp, ci = oracle.index_vev(prob["nverts"], [prob["endpoints"]])
    vals, _ = oracle.tet_assemble(oracle.tet_params(), prob, rp, ci)
    return rp, ci, vals


def dense(rp, ci, blocks, n):
    M = np.zeros((3 * n, 3 * n))
    b = np.asarray(blocks).reshape(-1, 3, 3)
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            j = ci[k]
            M[3 * i:3 * i + 3, 3 * j:3 * j + 3] = b[k]
    return M


def random_params(seed, scale=1.0):
    p = G.init_params(seed, dec_scale=1.0)
    rng = np.random.default_rng(seed + 100)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.normal(0, 0.1, p[k].shape)
        else:
            p[k] = p[k] * scale
    return p


def test_normalisation_and_block_jacobi_at_zero_output(oracle):
    rp, ci, vals = small_system(oracle)
    n = rp.size - 1
    g = G.graph(rp, ci)
    ft = G.features(g, vals)
    Vb = vals.reshape(-1, 3, 3)
    D = Vb[g["diag"]]
    Li = ft["Linv"]
    np.testing.assert_allclose(Li @ D @ Li.transpose(0, 2, 1), np.tile(np.eye(3), (n, 1, 1)),
                               atol=1e-12)
    assert np.all(np.triu(Li, 1) == 0.0)
    p = random_params(1)
    p["W_d"][:] = 0.0
    p["b_d"][:] = 0.0
    B, aux = G.precond_blocks(p, rp, ci, vals)
    assert np.all(aux["C"] == 0.0)
    Bd = dense(rp, ci, B, n)
    Dinv = np.zeros_like(Bd)
    for i in range(n):
        Dinv[3 * i:3 * i + 3, 3 * i:3 * i + 3] = np.linalg.inv(D[i])
    np.testing.assert_allclose(Bd, Dinv, rtol=1e-10, atol=1e-12 * np.abs(Dinv).max())


@pytest.mark.parametrize("scale", [1.0, 30.0])
def test_preconditioner_symmetric_and_congruent(oracle, scale):
    # large weights push the raw output far out: every block stays inside the
    # unit Frobenius ball, C and B are exactly symmetric, and B = L^T (I + C) L
    # is congruent to I + C (same inertia), so B is SPD exactly when I + C is,
    # which the library's spectral guard enforces
    rp, ci, vals = small_system(oracle, n=3, seed=4)
    n = rp.size - 1
    B, aux = G.precond_blocks(random_params(2, scale), rp, ci, vals, kappa=0.9)
    g = aux["g"]
    Bm = B.reshape(-1, 3, 3)
    assert np.array_equal(Bm, Bm[g["t"]].transpose(0, 2, 1))
    C = aux["C"]
    assert np.sqrt((C * C).sum(axis=(1, 2))).max() <= 1.0 + 1e-12
    Cd = dense(rp, ci, C, n)
    np.testing.assert_allclose(Cd, Cd.T, rtol=0, atol=1e-15)
    Bd = dense(rp, ci, B, n)
    assert np.array_equal(Bd, Bd.T)
    ec = np.linalg.eigvalsh(np.eye(3 * n) + Cd)
    eb = np.linalg.eigvalsh(Bd)
    tol = 1e-12 * np.abs(eb).max()
    assert (ec < -1e-12).sum() == (eb < -tol).sum()
    assert (ec > 1e-12).sum() == (eb > tol).sum()


def test_loss_matches_dense_formula(oracle):
    rp, ci, vals = small_system(oracle)
    n = rp.size - 1
    rng = np.random.default_rng(0)
    V = rng.normal(size=(3 * n, 4))
    p = random_params(3)
    L, _, aux = G.loss_and_grads(p, rp, ci, vals, V, kappa=0.9)
    Li = dense(rp, ci, [np.zeros((3, 3))] * len(ci), n)
    for i in range(n):
        Li[3 * i:3 * i + 3, 3 * i:3 * i + 3] = aux["ft"]["Linv"][i]
    A = dense(rp, ci, vals, n)
    Ah = Li @ A @ Li.T
    C = dense(rp, ci, aux["C"], n)
    R = V - (np.eye(3 * n) + C) @ Ah @ V
    assert abs(L - (R * R).sum() / V.size) <= 1e-12 * L
    # and B = Linv^T (I + C) Linv
    Bd = dense(rp, ci, aux["B"], n)
    np.testing.assert_allclose(Bd, Li.T @ (np.eye(3 * n) + C) @ Li, rtol=1e-9,
                               atol=1e-12 * np.abs(Bd).max())


def test_gradients_match_finite_differences(oracle):
    rp, ci, vals = small_system(oracle, n=2, seed=1)
    n = rp.size - 1
    rng = np.random.default_rng(5)
    V = rng.normal(size=(3 * n, 3))
    p = random_params(7, scale=0.7)
    L, grads, _ = G.loss_and_grads(p, rp, ci, vals, V)
    flat = G.flatten(p)
    gflat = G.flatten(grads)
    idx = rng.choice(flat.size, 60, replace=False)
    idx = np.concatenate([idx, [flat.size - 1 - k for k in range(16)]])  # decoder
    eps = 1e-6
    for k in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[k] += eps
        fm[k] -= eps
        Lp = G.loss_and_grads(G.unflatten(fp), rp, ci, vals, V)[0]
        Lm = G.loss_and_grads(G.unflatten(fm), rp, ci, vals, V)[0]
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - gflat[k]) <= 1e-6 * max(1.0, abs(gflat).max()) + 1e-5 * abs(fd), \
            (k, fd, gflat[k])


def test_gradients_match_torch_autograd(oracle):
    torch = pytest.importorskip("torch")
    rp, ci, vals = small_system(oracle, n=3, seed=2)
    n = rp.size - 1
    rng = np.random.default_rng(9)
    V = rng.normal(size=(3 * n, 4))
    p = random_params(11, scale=0.8)
    kappa = 0.9
    L, grads, aux = G.loss_and_grads(p, rp, ci, vals, V, kappa=kappa)
    g, ft = aux["g"], aux["ft"]
    T = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in p.items()}
    brow = torch.tensor(g["brow"])
    col = torch.tensor(g["ci"])
    t = torch.tensor(g["t"])
    rowlen = torch.tensor(np.maximum(g["rowlen"], 1), dtype=torch.float64)[:, None]
    fv = torch.tensor(ft["fv"])
    fe = torch.tensor(ft["fe"])
    h = torch.relu(fv @ T["W_ev"] + T["b_ev"])
    e = torch.relu(fe @ T["W_ee"] + T["b_ee"])
    for l in range(G.NLAYERS):
        m = torch.relu(torch.cat([h[brow], h[col], e], 1) @ T[f"W_m{l}"] + T[f"b_m{l}"])
        a = torch.zeros(n, G.H, dtype=torch.float64).index_add(0, brow, m) / rowlen
        h = h + torch.relu(torch.cat([h, a], 1) @ T[f"W_u{l}"] + T[f"b_u{l}"])
    R = (torch.cat([h[brow], h[col], e], 1) @ T["W_d"] + T["b_d"])[:, :9].reshape(-1, 3, 3)
    S = 0.5 * (R + R[t].transpose(1, 2))
    C = (1.0 / torch.sqrt(1 + (S * S).sum((1, 2))))[:, None, None] * S
    C = C * torch.tensor(~g["diag"], dtype=torch.float64)[:, None, None]
    Ah = torch.tensor(ft["Ah"])
    V3 = torch.tensor(V.reshape(n, 3, -1))

    def apply(Bk, x):
        return torch.zeros_like(x).index_add(0, brow, Bk @ x[col])

    W3 = apply(Ah, V3)
    res = V3 - W3 - apply(C, W3)
    Lt = (res * res).sum() / V.size
    Lt.backward()
    assert abs(Lt.item() - L) <= 1e-12 * L
    for k in p:
        np.testing.assert_allclose(grads[k], T[k].grad.numpy(), rtol=1e-9,
                                   atol=1e-12 * max(1.0, np.abs(grads[k]).max()), err_msg=k)


def test_bf16_rounding_matches_definition():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-39, 65504.0], np.float32)
    got = G.bf16(x)
    assert got[0] == 1.0
    assert got[1] == 1.0                 # tie to even
    assert got[2] == 1.0 + 2 ** -7       # tie to even (odd lower neighbour)
    assert got[3] == -2.5
    assert got[5] == 65536.0
    u = np.asarray(got, np.float32).view(np.uint32)
    assert np.all(u & 0xFFFF == 0)


# ---- scalar systems (1x1 blocks): the hex Poisson matrix ------------------------------
def small_hex(oracle, n=3, seed=0):
    prob = meshes.hex_problem(n, seed=seed, std_log10=1.0, corr_len=2.0)
    rp, ci = oracle.index_vev(prob["nverts"], [prob["endpoints"]])
    vals, _ = oracle.hex_assemble(oracle.HexParams(prob["h"], 1.0, 1.0), prob, rp, ci)
    return rp, ci, vals


def dense1(rp, ci, vals, n):
    M = np.zeros((n, n))
    for i in range(n):
        M[i, ci[rp[i]:rp[i + 1]]] = np.asarray(vals).ravel()[rp[i]:rp[i + 1]]
    return M


def test_scalar_normalisation_and_jacobi_at_zero_output(oracle):
    rp, ci, vals = small_hex(oracle)
    n = rp.size - 1
    g = G.graph(rp, ci)
    ft = G.features(g, vals, bd=1)
    d = vals[g["diag"]]
    np.testing.assert_allclose(ft["Linv"].ravel() ** 2 * d, 1.0, rtol=1e-14)
    assert np.all(ft["fe"][:, 1:9] == 0.0) and np.all(ft["fv"][:, 1:9] == 0.0)
    p = random_params(1)
    p["W_d"][:] = 0.0
    p["b_d"][:] = 0.0
    B, aux = G.precond_blocks(p, rp, ci, vals, bd=1)
    assert np.all(aux["C"] == 0.0)
    np.testing.assert_allclose(dense1(rp, ci, B, n), np.diag(1.0 / d), rtol=1e-13)


def test_scalar_preconditioner_symmetric_and_congruent(oracle):
    rp, ci, vals = small_hex(oracle, n=3, seed=4)
    n = rp.size - 1
    B, aux = G.precond_blocks(random_params(2, 30.0), rp, ci, vals, kappa=0.9, bd=1)
    g = aux["g"]
    C = aux["C"].ravel()
    assert np.abs(C).max() <= 1.0 + 1e-12
    assert np.array_equal(C, C[g["t"]])
    Bd = dense1(rp, ci, B, n)
    assert np.array_equal(Bd, Bd.T)
    Cd = dense1(rp, ci, aux["C"], n)
    ec = np.linalg.eigvalsh(np.eye(n) + Cd)
    eb = np.linalg.eigvalsh(Bd)
    tol = 1e-12 * np.abs(eb).max()
    assert (ec < -1e-12).sum() == (eb < -tol).sum()
    assert (ec > 1e-12).sum() == (eb > tol).sum()


def test_scalar_loss_and_gradients(oracle):
    rp, ci, vals = small_hex(oracle, n=2, seed=1)
    n = rp.size - 1
    rng = np.random.default_rng(5)
    V = rng.normal(size=(n, 3))
    p = random_params(7, scale=0.7)
    L, grads, aux = G.loss_and_grads(p, rp, ci, vals, V, bd=1)
    # dense formula
    Li = np.diag(aux["ft"]["Linv"].ravel())
    Ah = Li @ dense1(rp, ci, vals, n) @ Li
    C = dense1(rp, ci, aux["C"], n)
    R = V - (np.eye(n) + C) @ Ah @ V
    assert abs(L - (R * R).sum() / V.size) <= 1e-12 * L
    # central finite differences
    flat, gflat = G.flatten(p), G.flatten(grads)
    idx = np.concatenate([rng.choice(flat.size, 40, replace=False),
                          [flat.size - 16, flat.size - 16 - 1536]])  # b_d[0], W_d[0][0]
    eps = 1e-6
    for k in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[k] += eps
        fm[k] -= eps
        Lp = G.loss_and_grads(G.unflatten(fp), rp, ci, vals, V, bd=1)[0]
        Lm = G.loss_and_grads(G.unflatten(fm), rp, ci, vals, V, bd=1)[0]
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - gflat[k]) <= 1e-6 * max(1.0, abs(gflat).max()) + 1e-5 * abs(fd), \
            (k, fd, gflat[k])
