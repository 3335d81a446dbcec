"""The oracle pinned against the reference's own fixtures, the reference's
own index builder (golden .npz from oracle/_ref), closed forms, and
independent numpy / scipy computations.  CPU only."""
import itertools

import numpy as np
import pytest

from conftest import ref_index_files
from simit_amd import meshes


def rows(rp, ci):
    return [list(ci[rp[i]:rp[i + 1]]) for i in range(len(rp) - 1)]


# ---- sparsity ---------------------------------------------------------------------------
def test_index_reference_fixtures(oracle, fixtures):
    for name, fx in fixtures.items():
        if not name.startswith("pathindex"):
            continue
        sets = [np.array(s, dtype=np.int32).reshape(len(s), -1) if s else
                np.zeros((0, 2), np.int32) for s in fx["sets"]]
        if fx["kind"] == "ve":
            rp, ci = oracle.index_ve(fx["nverts"], sets[0])
        else:
            rp, ci = oracle.index_vev(fx["nverts"], sets)
        assert rows(rp, ci) == fx["expected"], name


@pytest.mark.parametrize("path", ref_index_files(), ids=lambda p: p.stem)
def test_index_matches_reference_builder(oracle, path):
    z = np.load(path)
    sets = [z[k] for k in sorted(z.files) if k.startswith("set")]
    rp, ci = oracle.index_vev(int(z["nverts"]), sets)
    np.testing.assert_array_equal(rp, z["rowptr"])
    np.testing.assert_array_equal(ci, z["colidx"])


def test_index_brute_force_random(oracle):
    rng = np.random.default_rng(3)
    for card in (1, 2, 3, 4, 8):
        nv = 50
        ep = rng.integers(0, nv - 5, size=(40, card)).astype(np.int32)
        rp, ci = oracle.index_vev(nv, [ep])
        want = [set() for _ in range(nv)]
        for e in ep:
            for a in e:
                want[a].update(int(b) for b in e)
        assert rows(rp, ci) == [sorted(w) for w in want]
        rp2, ci2 = oracle.index_vev(nv, [ep], force_diag=True)
        assert rows(rp2, ci2) == [sorted(w | {i}) for i, w in enumerate(want)]


def test_index_closed_forms(oracle):
    n = 6
    npts, ep = meshes.create_box(n, n, n)
    rp, _ = oracle.index_vev(npts, [ep])
    assert rp[-1] == npts + 2 * ep.shape[0]
    assert (np.diff(rp) <= 7).all() and (np.diff(rp) == 7).sum() == (n - 2) ** 3
    x, tets = meshes.tet5_grid(n)
    rp, _ = oracle.index_vev(x.shape[0], [tets])
    axis, face = 3 * n * (n + 1) ** 2, 3 * n * n * (n + 1)
    assert rp[-1] == x.shape[0] + 2 * (axis + face)
    deg = np.diff(rp).reshape(n + 1, n + 1, n + 1)
    inner = deg[1:-1, 1:-1, 1:-1]
    i, j, k = np.meshgrid(*[np.arange(1, n)] * 3, indexing="ij")
    assert (inner[(i + j + k) % 2 == 0] == 19).all()
    assert (inner[(i + j + k) % 2 == 1] == 7).all()
    x, hexes = meshes.hex_grid(n)
    rp, _ = oracle.index_vev(x.shape[0], [hexes])
    assert rp[-1] == (3 * (n + 1) - 2) ** 3


def test_create_box_order_matches_reference():
    p = [f for f in ref_index_files() if "box435" in f.name]
    z = np.load(p[0].parent / "ref_create_box_435.npz")
    _, ep = meshes.create_box(4, 3, 5)
    np.testing.assert_array_equal(ep, z["endpoints"])


def test_loc_is_position_in_row(oracle):
    npts, ep = meshes.create_box(3, 3, 3)
    rp, ci = oracle.index_vev(npts, [ep])
    for (a, b) in ep:
        for i, j in ((a, a), (a, b), (b, a), (b, b)):
            l = oracle.loc(rp, ci, int(i), int(j))
            assert rp[i] <= l < rp[i + 1] and ci[l] == j


# ---- element math ----------------------------------------------------------------------
def test_det_inv_textbook(oracle):
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = rng.standard_normal((3, 3))
        assert oracle.det3(a) == pytest.approx(np.linalg.det(a), rel=1e-12, abs=1e-14)
        np.testing.assert_allclose(oracle.inv3(a) @ a, np.eye(3), atol=1e-10)


def _literal_tet_Ke(xs, mu, lam):
    """fem_linear.sim:95-156 written out literally: Ke = W * B' * E * B."""
    M = np.stack([xs[i] - xs[3] for i in range(3)], axis=1)
    Binv = np.linalg.inv(M)
    W = -np.linalg.det(M) / 6.0
    E = np.zeros((6, 6))
    E[:3, :3] = lam
    for i in range(3):
        E[i, i] = lam + 2 * mu
    for i in range(3, 6):
        E[i, i] = mu
    dN = np.zeros((4, 3))
    dN[:3] = Binv
    dN[3] = -Binv.sum(axis=0)
    B = np.zeros((6, 12))
    for vi in range(4):
        for di in range(3):
            B[di, 3 * vi + di] = dN[vi, di]
        B[3, 3 * vi], B[3, 3 * vi + 1] = dN[vi, 1], dN[vi, 0]
        B[4, 3 * vi + 1], B[4, 3 * vi + 2] = dN[vi, 2], dN[vi, 1]
        B[5, 3 * vi], B[5, 3 * vi + 2] = dN[vi, 2], dN[vi, 0]
    return W * B.T @ E @ B, W


def test_tet_block_matches_literal_product(oracle):
    rng = np.random.default_rng(1)
    x, tets = meshes.tet5_grid(2)
    for t in tets[:10]:
        xs = x[t] + rng.uniform(-0.05, 0.05, size=(4, 3))
        mu, lam = 3.0e4 * rng.uniform(0.5, 2), 5.0e4 * rng.uniform(0.5, 2)
        Ke, W = _literal_tet_Ke(xs, mu, lam)
        for a, b in itertools.product(range(4), range(4)):
            blk, W2 = oracle.tet_block(xs, mu, lam, a, b)
            np.testing.assert_allclose(blk, Ke[3 * a:3 * a + 3, 3 * b:3 * b + 3],
                                       rtol=1e-11, atol=1e-9 * np.abs(Ke).max())
        assert W2 == pytest.approx(W, rel=1e-13)
        # rigid translations are in the null space; Ke is symmetric
        assert np.abs(Ke @ np.tile([1.0, 0, 0], 4)).max() < 1e-8 * np.abs(Ke).max()
        np.testing.assert_allclose(Ke, Ke.T, atol=1e-9 * np.abs(Ke).max())


def test_hex_kref_matches_quadrature(oracle):
    g = np.array([-1, 1]) / np.sqrt(3) * 0.5 + 0.5
    K = np.zeros((8, 8))
    for qx, qy, qz in itertools.product(g, g, g):
        grads = []
        for a in range(8):
            d = [(a >> s) & 1 for s in range(3)]
            q = [qx, qy, qz]
            f = [q[s] if d[s] else 1 - q[s] for s in range(3)]
            s_ = [1 if d[s] else -1 for s in range(3)]
            grads.append([s_[0] * f[1] * f[2], f[0] * s_[1] * f[2], f[0] * f[1] * s_[2]])
        G = np.array(grads)
        K += G @ G.T / 8.0
    for a, b in itertools.product(range(8), range(8)):
        assert oracle.hex_kref(a, b) == pytest.approx(K[a, b], abs=1e-14)


# ---- assembly -------------------------------------------------------------------------
def test_springs_assembly_properties(oracle):
    pts = meshes.springs_lattice(4, seed=2)
    rp, ci = oracle.index_vev(pts["npoints"], [pts["endpoints"]])
    P = oracle.springs_params(h=0.01, viscous=0.0)
    pts["fixed"][:] = 0
    vals, f = oracle.springs_assemble(P, pts, pts, rp, ci)
    n = pts["npoints"]
    A = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for l in range(rp[i], rp[i + 1]):
            A[3 * i:3 * i + 3, 3 * ci[l]:3 * ci[l] + 3] = vals[9 * l:9 * l + 9].reshape(3, 3)
    np.testing.assert_allclose(A, A.T, atol=1e-12)
    # A = M + h^2 K with K a Laplacian-like stiffness: A @ (translation) = m * translation
    np.testing.assert_allclose(A @ np.tile([0.0, 1.0, 0.0], n), np.tile([0.0, 1.0, 0.0], n),
                               atol=1e-10)
    assert np.linalg.eigvalsh(A).min() > 0


def test_tet_assembly_rigid_modes(oracle):
    prob = meshes.tet_problem(3, seed=0)
    prob["c"][:] = 0
    rp, ci = oracle.index_vev(prob["nverts"], [prob["endpoints"]])
    vals, f = oracle.tet_assemble(oracle.tet_params(fixed_diag=0.0), prob, rp, ci)
    n = prob["nverts"]
    y = oracle.bsrmv(rp, ci, vals, np.tile([1.0, -2.0, 0.5], n), 3)
    assert np.abs(y).max() < 1e-9 * np.abs(vals).max()
    # the body load sums to total volume times the load density
    np.testing.assert_allclose(f.reshape(-1, 3).sum(0), np.array(prob["load"]), rtol=1e-12)


def test_hex_assembly_row_sums(oracle):
    prob = meshes.hex_problem(3, seed=0)
    prob["dirichlet"][:] = 0
    rp, ci = oracle.index_vev(prob["nverts"], [prob["endpoints"]])
    vals, f = oracle.hex_assemble(oracle.HexParams(prob["h"], 1.0, 1.0), prob, rp, ci)
    y = oracle.bsrmv(rp, ci, vals, np.ones(prob["nverts"]), 1)
    assert np.abs(y).max() < 1e-12 * np.abs(vals).max()
    assert f.sum() == pytest.approx(1.0, rel=1e-12)


# ---- linear algebra and programs -------------------------------------------------------
def test_gemv_blocked_fixture(oracle, fixtures):
    fx = fixtures["gemv_blocked"]
    ep = np.array(fx["springs"], np.int32)
    rp, ci = oracle.index_vev(3, [ep])
    vals = np.zeros((len(ci), 4))
    for s, (a, b) in enumerate(ep):
        for i, j in ((a, a), (a, b), (b, a), (b, b)):
            vals[oracle.loc(rp, ci, int(i), int(j))] += fx["a"][s]
    c = oracle.bsrmv(rp, ci, vals.ravel(), np.array(fx["b"]).ravel(), 2)
    np.testing.assert_array_equal(c.reshape(3, 2), fx["expected_c"])


def test_matrix_vv_fixture(oracle, fixtures):
    fx = fixtures["assembly_matrix_vv"]
    ep = np.array(fx["springs"], np.int32)
    rp, ci = oracle.index_vev(3, [ep])
    vals = np.zeros(len(ci))
    for (a, b) in ep:
        for i, j in ((a, a), (a, b), (b, a), (b, b)):
            vals[oracle.loc(rp, ci, int(i), int(j))] += 1
    np.testing.assert_array_equal(oracle.bsrmv(rp, ci, vals, np.array(fx["a"], float), 1),
                                  fx["expected_b"])


def test_dot_fixtures(oracle, fixtures):
    for k in ("vector_dot", "vector_dot_blocked", "vector_dot_intrinsic"):
        x = np.array(fixtures[k]["x"])
        assert oracle.dot(x, x) == fixtures[k]["expected"]


def cg_program_matrix(oracle, fx):
    """A = I - 0.01 * (map f to springs reduce +) of cg.sim:7-59."""
    ep = np.array(fx["springs"], np.int32)
    rp, ci = oracle.index_vev(3, [ep])
    M = np.zeros(len(ci))
    for s, (a, b) in enumerate(ep):
        av = fx["a"][s]
        for (i, j), sg in zip(((a, a), (a, b), (b, a), (b, b)), (1, -1, -1, 1)):
            M[oracle.loc(rp, ci, int(i), int(j))] += sg * av
    vals = np.array([(1.0 if ci[l] == i else 0.0) - 0.01 * M[l]
                     for i in range(3) for l in range(rp[i], rp[i + 1])])
    return rp, ci, vals


def test_program_cg_fixture(oracle, fixtures):
    fx = fixtures["program_cg"]
    rp, ci, vals = cg_program_matrix(oracle, fx)
    o = oracle.cg_opts(maxiters=fx["maxiters"], tol=fx["tol"], rel=0, norm_form=1)
    x, st = oracle.cg(rp, ci, vals, np.array(fx["b"]), np.zeros(3), 1, o)
    np.testing.assert_allclose(x, fx["expected_c"], atol=fx["abs_tol"], rtol=0)


def isprings_setup(fx):
    m = fx["pi"] * fx["radius"] ** 2 * fx["l0"] * fx["density"]
    x = np.array(fx["x"], dtype=np.float64)
    pts = {"x": x.copy(), "v": np.zeros((3, 3)), "m": np.full(3, m),
           "fixed": (x[:, 2] < fx["zfloor"]).astype(np.uint8)}
    sp = {"endpoints": np.array(fx["springs"], np.int32), "k": np.full(2, fx["stiffness"]),
          "l0": np.full(2, fx["l0"])}
    return pts, sp


def test_program_isprings_fixture(oracle, fixtures):
    fx = fixtures["program_isprings"]
    pts, sp = isprings_setup(fx)
    rp, ci = oracle.index_vev(3, [sp["endpoints"]])
    P = oracle.springs_params(h=fx["h"], viscous=fx["viscous"], grav=fx["grav"],
                              fixed_fill=1.0, fixed_identity=0)
    o = oracle.cg_opts(maxiters=fx["maxiters"], tol=fx["tol"], rel=0)
    for _ in range(fx["steps"]):
        oracle.springs_step(P, o, pts, sp, rp, ci)
    np.testing.assert_allclose(pts["v"], fx["expected_v"], atol=fx["abs_tol"], rtol=0)
    np.testing.assert_allclose(pts["x"], fx["expected_x"], atol=fx["abs_tol"], rtol=0)


def test_planted_mistake_is_caught(oracle, fixtures):
    """A sign error in the spring force must break the isprings fixture."""
    fx = dict(fixtures["program_isprings"])
    fx["grav"] = [0.0, 0.0, 9.81]
    pts, sp = isprings_setup(fx)
    rp, ci = oracle.index_vev(3, [sp["endpoints"]])
    P = oracle.springs_params(h=fx["h"], viscous=fx["viscous"], grav=fx["grav"],
                              fixed_fill=1.0, fixed_identity=0)
    o = oracle.cg_opts(maxiters=fx["maxiters"], tol=fx["tol"], rel=0)
    for _ in range(fx["steps"]):
        oracle.springs_step(P, o, pts, sp, rp, ci)
    assert np.abs(pts["v"] - np.array(fx["expected_v"])).max() > 1e-3


def test_cg_matches_direct_solve(oracle):
    prob = meshes.tet_problem(3, seed=1)
    rp, ci = oracle.index_vev(prob["nverts"], [prob["endpoints"]])
    vals, f = oracle.tet_assemble(oracle.tet_params(), prob, rp, ci)
    n = prob["nverts"]
    A = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for l in range(rp[i], rp[i + 1]):
            A[3 * i:3 * i + 3, 3 * ci[l]:3 * ci[l] + 3] = vals[9 * l:9 * l + 9].reshape(3, 3)
    xd = np.linalg.solve(A, f)
    # thermal.sim's PCG also stops on z'r <= tol^2 b'b, which scales with the
    # preconditioner: it needs a tighter tol for the same residual
    for pc, tol in ((0, 1e-10), (1, 1e-13), (2, 1e-13)):
        x, st = oracle.cg(rp, ci, vals, f, np.zeros(3 * n), 3,
                          oracle.cg_opts(maxiters=5000, tol=tol, precond=pc))
        assert st["true_rel"] < 1e-9
        np.testing.assert_allclose(x, xd, rtol=1e-6, atol=1e-8 * np.abs(xd).max())


def test_cg_residual_replacement(oracle):
    """or_cg's periodic r = b - A x (thermal.sim:75-99): with replace_every = k
    the iterates still solve the system (direct solve), the iteration count
    stays that of the plain recurrence (they agree in exact arithmetic), and
    after a replacing iteration the iterated r'r is the true residual's."""
    prob = meshes.tet_problem(3, seed=1)
    rp, ci = oracle.index_vev(prob["nverts"], [prob["endpoints"]])
    vals, f = oracle.tet_assemble(oracle.tet_params(), prob, rp, ci)
    n = prob["nverts"]
    A = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for l in range(rp[i], rp[i + 1]):
            A[3 * i:3 * i + 3, 3 * ci[l]:3 * ci[l] + 3] = vals[9 * l:9 * l + 9].reshape(3, 3)
    xd = np.linalg.solve(A, f)
    bn = np.linalg.norm(f)
    for pc, tol in ((0, 1e-10), (2, 1e-13)):
        _, st0 = oracle.cg(rp, ci, vals, f, np.zeros(3 * n), 3,
                           oracle.cg_opts(maxiters=5000, tol=tol, precond=pc))
        for every in (1, 2, 7):
            x, st = oracle.cg(rp, ci, vals, f, np.zeros(3 * n), 3,
                              oracle.cg_opts(maxiters=5000, tol=tol, precond=pc,
                                             replace_every=every))
            np.testing.assert_allclose(x, xd, rtol=1e-6, atol=1e-8 * np.abs(xd).max())
            assert abs(st["iters"] - st0["iters"]) <= 0.1 * st0["iters"] + 2, (pc, every)
            if st["iters"] % every == 0:
                true_rr = (st["true_rel"] * bn) ** 2
                assert abs(st["rr"] - true_rr) <= 1e-10 * true_rr, (pc, every, st)

