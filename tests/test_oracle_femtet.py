"""Pins of the neo-Hookean tet oracle (test/input/program/femTet.sim), on the
CPU: the reference's own program test values, the log it uses, the
stiffness as the derivative of the force, and rigid-motion identities."""
from pathlib import Path

import numpy as np
import pytest

from simit_amd import meshes

GOLDEN = Path(__file__).resolve().parent / "golden" / "fem-bar2k"


def bar2k():
    return meshes.femtet_problem(GOLDEN / "bar2k.node", GOLDEN / "bar2k.ele")


def test_log_within_two_ulp(oracle):
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 2000)), 1 + rng.uniform(-1e-3, 1e-3, 2000),
                         rng.uniform(0.5, 2.0, 2000), [1.0, 2.0, 0.5, 1e-300, 1e300, 0.7071067811865476,
                                                       1.4142135623730951, np.nextafter(1, 0)]])
    worst = 0.0
    for x in xs:
        got, want = oracle.log(x), float(np.log(x))
        if want == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - want) / np.spacing(abs(want)))
    assert worst <= 2.0, worst
    assert oracle.log(0.0) == -np.inf and np.isnan(oracle.log(-1.0))


def test_ten_steps_reproduce_the_reference_program_test(oracle):
    """test/program-fem-test.cpp:97-105: positions of vertices 100, 200 and
    300 after ten femTet.sim steps of the bar2k mesh, to 1e-13."""
    prob = bar2k()
    assert prob["nverts"] == 439 and prob["endpoints"].shape == (1151, 4)
    prob["W"], prob["B"] = oracle.femtet_precompute(prob)
    rp, ci = oracle.index_vev(prob["nverts"], [prob["endpoints"]])
    for _ in range(10):
        oracle.femtet_step(prob, rp, ci)
    x = prob["x"]
    want = {100: (0.010771915616785779, 0.058853573999788439, 0.030899457015375883),
            200: (0.0028221631202928516, 0.017969982607667911, 0.012885386063393013),
            300: (0.02411959295647129, 0.052036155669135678, 0.030173075240629205)}
    for v, w in want.items():
        np.testing.assert_allclose(x[v], w, rtol=0, atol=1e-13)


def one_tet(seed):
    rng = np.random.default_rng(seed)
    rest = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) + \
        0.1 * rng.standard_normal((4, 3))
    xs = rest + 0.05 * rng.standard_normal((4, 3))  # deformed, not inverted
    return rest, xs


@pytest.mark.parametrize("seed", [1, 2])
def test_stiffness_is_the_derivative_of_the_force(oracle, seed):
    rest, xs = one_tet(seed)
    B, W = np.zeros(9), 0.0
    import ctypes as C
    Wc = C.c_double()
    oracle.lib().or_femtet_geom(rest.ravel().ctypes.data, B.ctypes.data, C.addressof(Wc))
    W = Wc.value
    u, l = 2.7e3, 1.9e4

    def forces(x):
        H = oracle.femtet_H(x, u, l, W, B)
        f = np.zeros((4, 3))
        f[:3] = H.T  # f_ii = H(:, ii)
        f[3] = -H.sum(axis=1)
        return f
    eps = 1e-7
    for row in range(4):
        for kk in range(3):
            xp, xm = xs.copy(), xs.copy()
            xp[row, kk] += eps
            xm[row, kk] -= eps
            df = (forces(xp) - forces(xm)) / (2 * eps)
            for jj in range(4):
                kb = oracle.femtet_kb(xs, u, l, W, B, jj, row)
                np.testing.assert_allclose(kb[:, kk], df[jj], rtol=1e-5,
                                           atol=1e-6 * np.abs(df).max())


def test_rigid_motions(oracle):
    """Element forces sum to zero, and the stiffness blocks of a row sum to
    zero over the columns (a translation changes nothing) and over the rows
    (the forces stay balanced); a rotated tet has rotated forces."""
    rest, xs = one_tet(3)
    B = np.zeros(9)
    import ctypes as C
    Wc = C.c_double()
    oracle.lib().or_femtet_geom(rest.ravel().ctypes.data, B.ctypes.data, C.addressof(Wc))
    W, u, l = Wc.value, 2.7e3, 1.9e4
    K = np.array([[oracle.femtet_kb(xs, u, l, W, B, jj, row) for row in range(4)]
                  for jj in range(4)])
    scale = np.abs(K).max()
    assert np.abs(K.sum(axis=0)).max() <= 1e-11 * scale
    assert np.abs(K.sum(axis=1)).max() <= 1e-11 * scale
    H = oracle.femtet_H(xs, u, l, W, B)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    Hr = oracle.femtet_H(xs @ R.T, u, l, W, B)
    np.testing.assert_allclose(Hr, R @ H, rtol=0, atol=1e-10 * np.abs(H).max())
