"""The neo-Hookean tet program (test/input/program/femTet.sim) on the GPU,
on the reference's bar2k mesh (its tetgen files are a fixture):
initializeTet and main's assembly bit-exact against the oracle in FP64, and
ten implicit steps through the C ABI (assembly, CG with the program's
absolute 1e-6 on r'r and 100 iterations, the position update) reproduce
test/program-fem-test.cpp:97-105 to 1e-13."""
from pathlib import Path

import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden" / "fem-bar2k"
WANT = {100: (0.010771915616785779, 0.058853573999788439, 0.030899457015375883),
        200: (0.0028221631202928516, 0.017969982607667911, 0.012885386063393013),
        300: (0.02411959295647129, 0.052036155669135678, 0.030173075240629205)}


def params(h=0.01):
    import ctypes as C
    return S.FemTetParams(h, 1000.0, (C.c_double * 3)(0.0, -10.0, 0.0))


def setup(ctx):
    prob = meshes.femtet_problem(GOLDEN / "bar2k.node", GOLDEN / "bar2k.ele")
    d = {k: to_dev(prob[k]) for k in ("x", "v", "c", "m", "endpoints", "u", "l")}
    idx = ctx.build_index(prob["nverts"], [d["endpoints"]])
    nt = prob["endpoints"].shape[0]
    d["W"] = torch.empty(nt, dtype=torch.float64, device="cuda")
    d["B"] = torch.empty(nt * 9, dtype=torch.float64, device="cuda")
    ctx.femtet_precompute(idx, params(), d, d, d["W"], d["B"], d["m"])
    return prob, d, idx


def test_femtet_precompute_and_assembly_bit_exact(ctx, oracle):
    prob, d, idx = setup(ctx)
    W, B = oracle.femtet_precompute(prob)
    np.testing.assert_array_equal(d["W"].cpu().numpy(), W)
    np.testing.assert_array_equal(d["B"].cpu().numpy(), B)
    np.testing.assert_array_equal(d["m"].cpu().numpy(), prob["m"])
    prob["W"], prob["B"] = W, B
    rp, ci = idx.to_host()
    orp, oci = oracle.index_vev(prob["nverts"], [prob["endpoints"]])
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    vals = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(3 * prob["nverts"], dtype=torch.float64, device="cuda")
    ctx.assemble_femtet(idx, params(), d, d, S.F64, vals, None, rhs)
    ov, of = oracle.femtet_assemble(prob, rp, ci)
    np.testing.assert_array_equal(vals.cpu().numpy(), ov)
    np.testing.assert_array_equal(rhs.cpu().numpy(), of)
    # the FP32 hi/lo split of the same sums
    hi = torch.empty(idx.nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    ctx.assemble_femtet(idx, params(), d, d, S.F32X2, hi, lo, rhs)
    h64 = hi.double() + lo.double()
    assert torch.allclose(h64, vals, rtol=2 ** -46, atol=0)


def test_femtet_ten_steps_match_the_reference(ctx):
    prob, d, idx = setup(ctx)
    vals = torch.empty(idx.nnz * 9, dtype=torch.float64, device="cuda")
    rhs = torch.empty(3 * prob["nverts"], dtype=torch.float64, device="cuda")
    A = idx.bsr(3, S.F64, vals)
    h = 0.01
    for step in range(10):
        ctx.assemble_femtet(idx, params(h), d, d, S.F64, vals, None, rhs)
        st = ctx.cg(A, rhs, d["v"], ctx.cg_options(maxiters=100, tol=1e-6, rel=0))
        print(step, st.as_dict(), flush=True)
        ctx.axpby(h, d["v"], 1.0, d["x"])
    x = d["x"].cpu().numpy().reshape(-1, 3)
    for v, w in WANT.items():
        np.testing.assert_allclose(x[v], w, rtol=0, atol=1e-13)
