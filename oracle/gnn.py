"""CPU oracle of the learned block-Jacobi correction (numpy, float64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
kernels in simit_amd/csrc/smh_gnn.hip.  The product never imports it.

The reference has no learned preconditioner, so nothing in it pins this
part: the oracle is pinned by central finite differences of its own loss
(float64, no BF16 rounding) and by torch-CPU autograd of the same forward
pass (tests/test_gnn_oracle.py).

Model (all sizes fixed: F = 16 input features, H = 32 hidden):
  normalisation   D_i = A_ii = L_i L_i^T (3x3 Cholesky), Linv_i = L_i^-1,
                  Ahat_k = Linv_i A_k Linv_j^T for block k = (i, j)
  edge features   [Ahat_k (9), ||Ahat_k||_F, [i == j], 0 x 5]   (diagonal: Ahat = I)
  vertex features [sum_{j != i} Ahat_ij (9), sum_{j != i} ||Ahat_ij||_F, deg_i / 16, 0 x 5]
  encoders        h = relu(f_v W_ev + b_ev),  e = relu(f_e W_ee + b_ee)
  3 layers        m_k = relu([h_i, h_j, e_k] W_m + b_m)
                  a_i = mean over the blocks k of row i of m_k
                  h_i = h_i + relu([h_i, a_i] W_u + b_u)
  decoder         R_k = [h_i, h_j, e_k] W_d + b_d  (first 9 of 16 outputs, a 3x3 block)
  output          S_k = (R_k + R_t(k)^T) / 2  (t(k): the block (j, i))
                  C_k = S_k / sqrt(1 + ||S_k||_F^2),  C_ii = 0
                  B_k = Linv_i^T (I [k diagonal] + C_k) Linv_j
  C is exactly symmetric.  Positive definiteness of I + C is not built into
  the output: the library's spectral guard (a Lanczos estimate of lambda_min
  of the stored C, kept at or above -kappa by a global scale) enforces it at
  inference; this oracle returns the unscaled C, which the tests compare with
  the library's when its guard applied no scale, and against whose exact
  lambda_min (scipy) the guard's estimate is checked.
  loss            L = sum (V - (I + C) Ahat V)^2 / (K * ndof) over K probe
                  vectors V: the stochastic Frobenius norm of I - (I + C) Ahat.

Scalar systems (1x1 blocks, bd = 1): Linv_i = 1 / sqrt(a_ii); Ahat_k and the
sums take feature slot 0 (slots 1-8 zero); the decoder's first output is R_k.

With mimic=True the oracle rounds to bfloat16 where the GPU kernels store or
feed BF16 (weights, MFMA inputs h / a / features, stored e and m); gradients
pass those roundings straight through, as the GPU's do.
"""
from __future__ import annotations

import numpy as np

F = 16
H = 32
NLAYERS = 3


def param_layout():
    """(name, shape) in the order of the flat float32 parameter vector."""
    lay = [("W_ev", (F, H)), ("b_ev", (H,)), ("W_ee", (F, H)), ("b_ee", (H,))]
    for l in range(NLAYERS):
        lay += [(f"W_m{l}", (3 * H, H)), (f"b_m{l}", (H,)),
                (f"W_u{l}", (2 * H, H)), (f"b_u{l}", (H,))]
    lay += [("W_d", (3 * H, 16)), ("b_d", (16,))]
    return lay


def num_params():
    return sum(int(np.prod(s)) for _, s in param_layout())


def unflatten(flat):
    out, o = {}, 0
    for name, shape in param_layout():
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def flatten(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel()
                           for name, _ in param_layout()])


def init_params(seed=0, dec_scale=0.1):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in param_layout():
        if name.startswith("W"):
            a = np.sqrt(6.0 / (shape[0] + shape[1]))
            p[name] = rng.uniform(-a, a, shape) * (dec_scale if name == "W_d" else 1.0)
        else:
            p[name] = np.zeros(shape)
    return p


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(x):
    return np.asarray(x, dtype=np.float64)


# ---- graph and features -------------------------------------------------------------
def chol_inv3(D):
    """Linv = L^-1 for D = L L^T (lower triangle of D used), batched [n,3,3]."""
    d = D
    l00 = np.sqrt(d[:, 0, 0])
    l10 = d[:, 1, 0] / l00
    l20 = d[:, 2, 0] / l00
    l11 = np.sqrt(d[:, 1, 1] - l10 * l10)
    l21 = (d[:, 2, 1] - l20 * l10) / l11
    l22 = np.sqrt(d[:, 2, 2] - l20 * l20 - l21 * l21)
    i00, i11, i22 = 1.0 / l00, 1.0 / l11, 1.0 / l22
    i10 = -l10 * i00 * i11
    i21 = -l21 * i11 * i22
    i20 = -(l20 * i00 + l21 * i10) * i22
    Li = np.zeros_like(D)
    Li[:, 0, 0], Li[:, 1, 1], Li[:, 2, 2] = i00, i11, i22
    Li[:, 1, 0], Li[:, 2, 1], Li[:, 2, 0] = i10, i21, i20
    return Li


def graph(rowptr, colidx):
    rp = np.asarray(rowptr, dtype=np.int64)
    ci = np.asarray(colidx, dtype=np.int64)
    n = rp.size - 1
    rowlen = np.diff(rp)
    brow = np.repeat(np.arange(n), rowlen)
    # t(k): position of block (j, i) (the pattern is symmetric)
    key = brow * n + ci
    tkey = ci * n + brow
    order = np.argsort(key)
    t = order[np.searchsorted(key[order], tkey)]
    assert np.array_equal(key[t], tkey), "the sparsity pattern is not symmetric"
    diag = ci == brow
    deg = rowlen - np.bincount(brow[diag], minlength=n)
    return {"n": n, "rp": rp, "ci": ci, "brow": brow, "t": t, "diag": diag,
            "rowlen": rowlen, "deg": deg}


def features(g, vals, bd=3):
    """Linv, normalised blocks and the input features of the network.  For
    1x1 blocks (bd = 1) the scalar Ahat_k takes feature slot 0 and slots 1-8
    stay zero; Linv_i = 1 / sqrt(a_ii)."""
    bb = bd * bd
    V = np.asarray(vals, dtype=np.float64).reshape(-1, bd, bd)
    n, brow, ci, diag = g["n"], g["brow"], g["ci"], g["diag"]
    D = np.tile(np.eye(bd), (n, 1, 1))
    D[brow[diag]] = V[diag]
    Linv = chol_inv3(D) if bd == 3 else 1.0 / np.sqrt(D)
    Ah = Linv[brow] @ V @ Linv[ci].transpose(0, 2, 1)
    Ah[diag] = np.eye(bd)
    nrm = np.sqrt((Ah * Ah).sum(axis=(1, 2)))
    fe = np.zeros((V.shape[0], F))
    fe[:, :bb] = Ah.reshape(-1, bb)
    fe[:, 9] = nrm
    fe[:, 10] = diag.astype(np.float64)
    off = ~diag
    fv = np.zeros((n, F))
    np.add.at(fv[:, :bb], brow[off], Ah[off].reshape(-1, bb))
    np.add.at(fv[:, 9], brow[off], nrm[off])
    fv[:, 10] = g["deg"] / 16.0
    return {"Linv": Linv, "Ah": Ah, "fe": fe, "fv": fv, "bd": bd}


# ---- network ------------------------------------------------------------------------------
def _seg_mean(m, g):
    s = np.zeros((g["n"], m.shape[1]))
    np.add.at(s, g["brow"], m)
    return s / np.maximum(g["rowlen"], 1)[:, None]


def forward(params, g, ft, mimic=False):
    q = bf16 if mimic else _ident
    W = {k: (q(v) if k.startswith("W") else v) for k, v in params.items()}
    brow, ci = g["brow"], g["ci"]
    cache = {"mimic": mimic}
    fvq, feq = q(ft["fv"]), q(ft["fe"])
    zv = fvq @ W["W_ev"] + params["b_ev"]
    h = np.maximum(zv, 0.0)
    ze = feq @ W["W_ee"] + params["b_ee"]
    e = q(np.maximum(ze, 0.0))
    cache.update(fvq=fvq, feq=feq, zv=zv, ze=ze, e=e, layers=[])
    for l in range(NLAYERS):
        hq = q(h)
        X = np.concatenate([hq[brow], hq[ci], e], axis=1)
        z = X @ W[f"W_m{l}"] + params[f"b_m{l}"]
        m = q(np.maximum(z, 0.0))
        a = _seg_mean(m, g)
        aq = q(a)
        Xu = np.concatenate([hq, aq], axis=1)
        U = Xu @ W[f"W_u{l}"] + params[f"b_u{l}"]
        cache["layers"].append({"h": h, "hq": hq, "X": X, "m": m, "aq": aq, "Xu": Xu, "U": U})
        h = h + np.maximum(U, 0.0)
    hq = q(h)
    Xd = np.concatenate([hq[brow], hq[ci], e], axis=1)
    R = Xd @ W["W_d"] + params["b_d"]
    cache.update(h_out=h, Xd=Xd, W=W)
    bb = ft.get("bd", 3) ** 2
    return R[:, :bb].copy(), cache


def output(R, g, ft, kappa):
    """C (normalised correction) and B (the preconditioner blocks)."""
    bd = ft.get("bd", 3)
    Rm = R.reshape(-1, bd, bd)
    S = 0.5 * (Rm + Rm[g["t"]].transpose(0, 2, 1))
    nS2 = (S * S).sum(axis=(1, 2))
    phi = 1.0 / np.sqrt(1.0 + nS2)
    c = np.ones_like(phi)
    C = (c * phi)[:, None, None] * S
    C[g["diag"]] = 0.0
    Bh = C.copy()
    Bh[g["diag"]] = np.eye(bd)
    Linv = ft["Linv"]
    B = Linv[g["brow"]].transpose(0, 2, 1) @ Bh @ Linv[g["ci"]]
    # exact symmetry: the block of the lower position of each pair is the
    # transpose of the upper one
    lower = g["t"] < np.arange(g["t"].size)
    B[lower] = B[g["t"][lower]].transpose(0, 2, 1)
    return C, B, {"S": S, "phi": phi, "c": c}


def apply_blocks(g, blocks, x3):
    """y_i = sum_k blocks_k x_j over row i (x3: [n, 3, K])."""
    y = np.zeros_like(x3)
    np.add.at(y, g["brow"], blocks @ x3[g["ci"]])
    return y


def loss(C, g, ft, V):
    """L, dL/dC and W = Ahat V for probes V [ndof, K]."""
    n, bd = g["n"], ft.get("bd", 3)
    K = V.shape[1]
    V3 = V.reshape(n, bd, K)
    W3 = apply_blocks(g, ft["Ah"], V3)
    Y3 = W3 + apply_blocks(g, C, W3)
    res = V3 - Y3
    N = K * bd * n
    L = float((res * res).sum() / N)
    G = -2.0 * res / N
    dC = G[g["brow"]] @ W3[g["ci"]].transpose(0, 2, 1)
    dC[g["diag"]] = 0.0
    return L, dC, W3


def output_backward(dC, g, oc):
    S, phi, c = oc["S"], oc["phi"], oc["c"]
    inner = (dC * S).sum(axis=(1, 2))
    dS = (c * phi)[:, None, None] * dC - (c * phi ** 3 * inner)[:, None, None] * S
    dS[g["diag"]] = 0.0
    dR = 0.5 * dS + 0.5 * dS[g["t"]].transpose(0, 2, 1)
    return dR.reshape(dR.shape[0], -1)


def backward(params, g, cache, dR9):
    W = cache["W"]
    brow, ci, n = g["brow"], g["ci"], g["n"]
    grads = {k: np.zeros_like(v) for k, v in params.items()}
    dR = np.zeros((dR9.shape[0], 16))
    dR[:, :dR9.shape[1]] = dR9
    grads["W_d"] = cache["Xd"].T @ dR
    grads["b_d"] = dR.sum(axis=0)
    dX = dR @ W["W_d"].T
    dh = np.zeros((n, H))
    np.add.at(dh, brow, dX[:, :H])
    np.add.at(dh, ci, dX[:, H:2 * H])
    de = dX[:, 2 * H:].copy()
    for l in reversed(range(NLAYERS)):
        c = cache["layers"][l]
        dU = dh * (c["U"] > 0)
        grads[f"W_u{l}"] = c["Xu"].T @ dU
        grads[f"b_u{l}"] = dU.sum(axis=0)
        dXu = dU @ W[f"W_u{l}"].T
        da = dXu[:, H:]
        dh = dh + dXu[:, :H]
        dm = (da / np.maximum(g["rowlen"], 1)[:, None])[brow]
        dz = dm * (c["m"] > 0)
        grads[f"W_m{l}"] = c["X"].T @ dz
        grads[f"b_m{l}"] = dz.sum(axis=0)
        dX = dz @ W[f"W_m{l}"].T
        np.add.at(dh, brow, dX[:, :H])
        np.add.at(dh, ci, dX[:, H:2 * H])
        de += dX[:, 2 * H:]
    dze = de * (cache["ze"] > 0)
    grads["W_ee"] = cache["feq"].T @ dze
    grads["b_ee"] = dze.sum(axis=0)
    dzv = dh * (cache["zv"] > 0)
    grads["W_ev"] = cache["fvq"].T @ dzv
    grads["b_ev"] = dzv.sum(axis=0)
    return grads


def loss_and_grads(params, rowptr, colidx, vals, V, kappa=0.9, mimic=False, bd=3):
    g = graph(rowptr, colidx)
    ft = features(g, vals, bd)
    R, cache = forward(params, g, ft, mimic)
    C, B, oc = output(R, g, ft, kappa)
    L, dC, _ = loss(C, g, ft, V)
    dR = output_backward(dC, g, oc)
    return L, backward(params, g, cache, dR), {"R": R, "C": C, "B": B, "g": g, "ft": ft}


def precond_blocks(params, rowptr, colidx, vals, kappa=0.9, mimic=False, bd=3):
    g = graph(rowptr, colidx)
    ft = features(g, vals, bd)
    R, _ = forward(params, g, ft, mimic)
    C, B, _ = output(R, g, ft, kappa)
    return B, {"R": R, "C": C, "g": g, "ft": ft}
