"""The fixed kernels under compiled Simit programs, through the C ABI.

smh_vec_op (the elementwise statements of an exported function) rounds once
per operation, like numpy; smh_stage_reduce sums each entry's staged writes
in the listed order from 0 (`reduce +`) or keeps the last write taken (a map
without a reduction); smh_hiprtc_compile / smh_module_load /
smh_module_launch run a generated kernel with its argument struct.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import simit_amd as S
from conftest import to_dev

pytestmark = pytest.mark.gpu

VEC = {"lincomb": 0, "divs": 1, "emul": 2, "ediv": 3, "fill": 4, "adds": 5, "sdiv": 6}


def vec_op(ctx, op, a, x, b, y, out):
    ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    ctx._check(ctx.lib.smh_vec_op(ctx.h, VEC[op], out.numel(), a, ptr(x), b, ptr(y), ptr(out)))
    ctx.synchronize()


def test_vec_op_rounds_like_numpy(ctx):
    rng = np.random.default_rng(5)
    n = 100_003
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    y = rng.standard_normal(n)
    a, b = 0.1 + 1e-9, -3.7
    dx, dy = to_dev(x), to_dev(y)
    out = torch.empty_like(dx)
    expect = {
        "lincomb": a * x + b * y,
        "divs": x / a,
        "emul": x * y,
        "ediv": x / y,
        "fill": np.full(n, a),
        "adds": x + a,
        "sdiv": a / x,
    }
    for op, want in expect.items():
        vec_op(ctx, op, a, dx, b, dy, out)
        got = out.cpu().numpy()
        assert np.array_equal(got, want), op
    vec_op(ctx, "lincomb", a, dx, 0.0, None, out)  # a*x alone
    assert np.array_equal(out.cpu().numpy(), a * x)
    # r - alpha*q as one launch equals the two-step evaluation
    vec_op(ctx, "lincomb", 1.0, dx, -a, dy, out)
    assert np.array_equal(out.cpu().numpy(), x - a * y)
    # in place (out aliases x)
    dz = dx.clone()
    vec_op(ctx, "lincomb", 1.0, dz, b, dy, dz)
    assert np.array_equal(dz.cpu().numpy(), x + b * y)


def stage_reduce(ctx, ptr, src, stage, taken, out, bs):
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    ctx._check(ctx.lib.smh_stage_reduce(ctx.h, len(ptr) - 1, bs, p(ptr), p(src), p(stage),
                                        p(taken), p(out)))
    ctx.synchronize()


@pytest.mark.parametrize("bs", [1, 3, 9])
def test_stage_reduce_order_and_assignment(ctx, bs):
    rng = np.random.default_rng(bs)
    nslots, nout = 5000, 700
    stage = rng.standard_normal((nslots, bs)) * 10.0 ** rng.integers(-6, 6, (nslots, 1))
    counts = rng.integers(0, 12, nout)
    counts[::97] = 0  # entries nothing writes
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    src = rng.integers(0, nslots, int(ptr[-1])).astype(np.uint32)
    # reduce +: ((0 + w0) + w1) + ... in list order
    want = np.zeros((nout, bs))
    for i in range(nout):
        acc = np.zeros(bs)
        for k in range(int(ptr[i]), int(ptr[i + 1])):
            acc = acc + stage[src[k]]
        want[i] = acc
    out = torch.full((nout * bs,), 7.0, dtype=torch.float64, device="cuda")
    stage_reduce(ctx, to_dev(ptr.view(np.int64)), to_dev(src.view(np.int32)),
                 to_dev(stage.ravel()), None, out, bs)
    assert np.array_equal(out.cpu().numpy().reshape(nout, bs), want)
    # plain writes: the last taken one; untouched entries keep their values
    taken = (rng.random(nslots) < 0.6).astype(np.uint8)
    init = rng.standard_normal((nout, bs))
    want = init.copy()
    for i in range(nout):
        for k in range(int(ptr[i]), int(ptr[i + 1])):
            if taken[src[k]]:
                want[i] = stage[src[k]]
    out = to_dev(init.ravel())
    stage_reduce(ctx, to_dev(ptr.view(np.int64)), to_dev(src.view(np.int32)),
                 to_dev(stage.ravel()), to_dev(taken), out, bs)
    assert np.array_equal(out.cpu().numpy().reshape(nout, bs), want)


SOURCE = b"""
struct Args { unsigned long long n; double *y; const double *x; double s; };
extern "C" __global__ void __launch_bounds__(256) scale_shift(Args a) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  a.y[i] = a.s * a.x[i] + (double)i;
}
"""


def test_hiprtc_module_launch(ctx):
    lib = ctx.lib
    img, nbytes = C.c_void_p(), C.c_size_t()
    log = C.create_string_buffer(1 << 14)
    rc = lib.smh_hiprtc_compile(SOURCE, b"t.hip", b"gfx950", C.byref(img), C.byref(nbytes), log,
                                len(log))
    assert rc == 0, log.value.decode()
    mod = C.c_void_p()
    try:
        ctx._check(lib.smh_module_load(ctx.h, img, nbytes, C.byref(mod)))
    finally:
        lib.smh_image_free(img)
    n = 10_000
    x = np.random.default_rng(1).standard_normal(n)
    dx = to_dev(x)
    dy = torch.zeros(n, dtype=torch.float64, device="cuda")

    class Args(C.Structure):
        _fields_ = [("n", C.c_uint64), ("y", C.c_void_p), ("x", C.c_void_p), ("s", C.c_double)]

    a = Args(n, dy.data_ptr(), dx.data_ptr(), 2.5)
    ctx._check(lib.smh_module_launch(ctx.h, mod, b"scale_shift", n, C.byref(a), C.sizeof(a)))
    ctx.synchronize()
    assert np.array_equal(dy.cpu().numpy(), 2.5 * x + np.arange(n))
    ctx._check(lib.smh_module_destroy(ctx.h, mod))
    # a compile error comes back with the compiler's message
    rc = lib.smh_hiprtc_compile(b"this is not C++", b"bad.hip", b"gfx950", C.byref(img),
                                C.byref(nbytes), log, len(log))
    assert rc != 0 and b"error" in log.value
