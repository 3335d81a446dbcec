"""Benchmark of the Simit hot path on one MI355X.

Workload (default): linear-elastic tets on a 96^3 cube grid split 5 tets per
cube (912,673 vertices, 2,738,019 DOF, 4,423,680 tets, 11,695,969 3x3
blocks), log10 E = 5 + 1 * Gaussian random field (correlation length 8
cells), nu = 0.3, clamped z = 0 face, uniform body load.  Synthetic, seeded.

One step = assemble the FP32 hi/lo blocked-CSR matrix and the right-hand
side from the resident mesh, then solve to a 1e-8 relative FP64 residual with
the mixed-precision block-Jacobi PCG (FP32 SpMV, FP64 dots and reliable
updates), from x = 0.  The metric is the time of that step; plain FP64 CG on
the same GPU is timed beside it.

The roofline object is for the dominant kernel, the FP32 blocked SpMV:
algorithmic bytes per launch = 36*nnz (values) + 4*nnz (colidx) + 4*(N+1)
(rowptr) + 12*N (x) + 12*N (y), timed with HIP events on the stream the
kernel runs on.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=96, help="cubes per edge of the tet grid")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--plain-maxiters", type=int, default=40000)
    ap.add_argument("--cg-maxiters", type=int, default=20000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)
    import torch
    import simit_amd as S
    from simit_amd import meshes

    torch.cuda.init()
    ctx = S.Context(0)
    stream = torch.cuda.current_stream()
    n = args.n
    t0 = time.time()
    prob = meshes.tet_problem(n, seed=args.seed, field="grf", mean_log10=5.0, std_log10=1.0,
                              corr_len=8.0)
    log(f"[bench] generated mesh n={n} nverts={prob['nverts']} ntets={prob['endpoints'].shape[0]}"
        f" in {time.time() - t0:.1f}s")
    dev = {k: torch.from_numpy(np.ascontiguousarray(prob[k])).cuda()
           for k in ("x", "c", "mu", "lam", "endpoints")}
    nv, nt = prob["nverts"], prob["endpoints"].shape[0]
    ndof = 3 * nv

    def ev():
        return torch.cuda.Event(enable_timing=True)

    # ---- sparsity index (device) ---------------------------------------------------
    torch.cuda.synchronize()
    e0, e1 = ev(), ev()
    e0.record(stream)
    idx = ctx.build_index(nv, [dev["endpoints"]])
    e1.record(stream)
    torch.cuda.synchronize()
    index_ms = e0.elapsed_time(e1)
    nnz = idx.nnz
    log(f"[bench] index nnz={nnz} in {index_ms:.2f} ms")

    P = S.TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
    hi = torch.empty(nnz * 9, dtype=torch.float32, device="cuda")
    lo = torch.empty_like(hi)
    rhs = torch.empty(ndof, dtype=torch.float64, device="cuda")
    A = idx.bsr(3, S.F32X2, hi, lo)
    opts = ctx.cg_options(maxiters=args.cg_maxiters, tol=1e-8, rel=1, precond=S.PRECOND_BLOCK_JACOBI,
                          mixed=1, check_every=16)

    def assemble():
        ctx.assemble_tets(idx, P, dev, dev, S.F32X2, hi, lo, rhs)

    # ---- assembly kernel time ----------------------------------------------------------
    assemble()
    torch.cuda.synchronize()
    reps = 5
    e0.record(stream)
    for _ in range(reps):
        assemble()
    e1.record(stream)
    torch.cuda.synchronize()
    asm_ms = e0.elapsed_time(e1) / reps
    asm_bytes = 8 * 9 * nnz + 8 * ndof + 16 * nt + 16 * nt + 24 * nv
    log(f"[bench] assembly {asm_ms:.3f} ms  {asm_bytes / asm_ms / 1e6:.0f} GB/s")

    # ---- SpMV kernel time (FP32 values, the inner-iteration operator) ------------------
    A32 = idx.bsr(3, S.F32, hi)
    xs = torch.randn(ndof, dtype=torch.float32, device="cuda")
    ys = torch.empty_like(xs)
    ctx.bsrmv(A32, xs, ys)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.spmv_reps):
        ctx.bsrmv(A32, xs, ys)
    e1.record(stream)
    torch.cuda.synchronize()
    spmv_ms = e0.elapsed_time(e1) / args.spmv_reps
    spmv_bytes = 40 * nnz + 4 * (nv + 1) + 24 * nv
    spmv_gbs = spmv_bytes / spmv_ms / 1e6
    log(f"[bench] spmv f32 {spmv_ms * 1e3:.1f} us  {spmv_gbs:.0f} GB/s")

    # ---- the step: assemble + mixed PCG to 1e-8 ---------------------------------------------
    def step():
        assemble()
        x = torch.zeros(ndof, dtype=torch.float64, device="cuda")
        st = ctx.cg(A, rhs, x, opts)
        return st, x

    for w in range(args.warmup):
        st, _ = step()
        log(f"[bench] warmup {w}: iters={st.iters} ru={st.reliable_updates} "
            f"true_rel={st.true_rel:.2e} {st.seconds:.3f}s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for s in range(args.steps):
        st, x = step()
        stats.append(st.as_dict())
        log(f"[bench] step {s}: iters={st.iters} ru={st.reliable_updates} "
            f"true_rel={st.true_rel:.2e} status={st.status}")
    torch.cuda.synchronize()
    step_s = (time.perf_counter() - t0) / args.steps
    iters = stats[-1]["iters"]
    converged = all(s["status"] == 0 and s["true_rel"] <= 1e-8 for s in stats)

    # ---- plain FP64 CG on the same GPU (the comparison the metric names) ----------------
    plain = None
    if not args.no_plain:
        v64 = torch.empty(nnz * 9, dtype=torch.float64, device="cuda")
        A64 = idx.bsr(3, S.F64, v64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.assemble_tets(idx, P, dev, dev, S.F64, v64, None, rhs)
        x = torch.zeros(ndof, dtype=torch.float64, device="cuda")
        st = ctx.cg(A64, rhs, x, ctx.cg_options(maxiters=args.plain_maxiters, tol=1e-8, rel=1,
                                                 precond=S.PRECOND_NONE, check_every=64))
        torch.cuda.synchronize()
        plain = {"seconds": time.perf_counter() - t0, "iters": st.iters,
                 "true_rel": st.true_rel, "converged": st.status == 0}
        log(f"[bench] plain fp64 CG: {plain}")
        del v64

    # ---- CPU baseline: the oracle on a bounded sample of the same workload --------------------
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle as O
        rp, ci = idx.to_host()
        t0 = time.perf_counter()
        ov, of = O.tet_assemble(O.tet_params(load=prob["load"]), prob, rp, ci)
        t_asm = time.perf_counter() - t0
        o2 = O.cg_opts(maxiters=2, tol=1e-8, precond=2)
        o12 = O.cg_opts(maxiters=12, tol=1e-8, precond=2)
        t0 = time.perf_counter()
        O.cg(rp, ci, ov, of, np.zeros(ndof), 3, o2)
        t2 = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.cg(rp, ci, ov, of, np.zeros(ndof), 3, o12)
        t12 = time.perf_counter() - t0
        per_iter = (t12 - t2) / 10
        cpu_total = t_asm + t2 + per_iter * (iters - 2)
        cpu = {"value": round(cpu_total, 3), "unit": "s", "cores": 1, "kind": "port",
               "sample": (f"oracle (single-thread FP64 C restatement) on the full mesh: one "
                          f"assembly ({t_asm:.2f}s) and 12 block-Jacobi PCG iterations "
                          f"({per_iter * 1e3:.0f} ms/iter), extrapolated to the {iters} "
                          f"iterations the GPU solve took"),
               "nproc": os.cpu_count()}
        log(f"[bench] cpu baseline {cpu}")

    out = {
        "metric": "time-to-1e-8 solve: learned-PCG vs plain CG; assembly+SpMV HBM GB/s",
        "value": round(step_s, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 SpMV / f64 dots+refinement",
        "data": "synthetic (seeded Gaussian random field material, generated on the host)",
        "config": {"workload": f"tet5 elasticity {n}^3 cubes, {ndof} DOF, Jacobi-PCG mixed "
                               f"precision to 1e-8", "nverts": nv, "ntets": nt, "nnz_blocks": nnz,
                   "solver": "block-Jacobi PCG, FP32 SpMV, FP64 dots + reliable updates"},
        "iterations": iters,
        "converged": converged,
        "true_rel": stats[-1]["true_rel"],
        "plain_cg_f64": plain,
        "speedup_vs_plain_cg_f64": (round(plain["seconds"] / step_s, 2) if plain else None),
        "index_build_ms": round(index_ms, 3),
        "assembly_ms": round(asm_ms, 4),
        "assembly_GBps": round(asm_bytes / asm_ms / 1e6, 1),
        "spmv_us": round(spmv_ms * 1e3, 2),
        "spmv_GBps": round(spmv_gbs, 1),
        "roofline": {"bound": "hbm", "kernel": "k_bsrmv3<SMH_F32> (FP32 3x3 BSR SpMV)",
                     "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes": spmv_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
