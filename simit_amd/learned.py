"""Training driver of the learned block-Jacobi correction (smh_gnn).

The samples are systems this library assembles (hex Poisson samples for
scalar systems: HexSampler).  For elasticity: one
n^3 x 5 tet grid (clamped z = 0 face, uniform body load) whose Young's
modulus is redrawn per sample as log10 E = mean + std * GRF (seeded,
FFT-smoothed white noise, correlation length in cells), Poisson ratio 0.3.
Each sample is assembled on the GPU (FP32 hi/lo) and fed to one Adam step of
the network's own forward and backward kernels.  Everything runs on the
device, the seeded material draw included (device_random_field: white noise
from a seeded device generator, smoothed with device FFTs).
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import meshes


def device_random_field(n: int, corr_len: float, seed: int):
    """Zero-mean, unit-std periodic field on n^3 cells, as
    meshes.gaussian_random_field (white noise smoothed in Fourier space by a
    Gaussian kernel of std corr_len cells) but drawn and filtered on the
    device from a seeded device generator: a float64 cuda tensor of n^3."""
    import torch
    gen = torch.Generator(device="cuda")
    gen.manual_seed(int(seed))
    w = torch.randn((n, n, n), generator=gen, dtype=torch.float64, device="cuda")
    f = torch.fft.fftfreq(n, dtype=torch.float64, device="cuda")
    k2 = (2 * math.pi) ** 2 * (f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2)
    g = torch.fft.ifftn(torch.fft.fftn(w) * torch.exp(-0.5 * k2 * corr_len ** 2)).real
    g = g - g.mean()
    sd = g.std(unbiased=False)
    # no host read of sd: the draw stays queued behind the device's work
    return torch.where(sd > 1e-12, g / sd, torch.zeros_like(g)).reshape(-1)


def tet_material(n: int, seed: int, mean_log10=5.0, std_log10=1.0, corr_len=8.0, nu=0.3):
    """(mu, lam) per tet of the n^3 x 5 grid for one seeded random field."""
    g = meshes.gaussian_random_field(n, corr_len, seed).ravel()
    E = 10.0 ** np.repeat(mean_log10 + std_log10 * g, 5)
    return meshes.lame(E, nu)


class TetSampler:
    """One resident n^3 x 5 tet mesh whose material is redrawn per sample."""

    def __init__(self, ctx, n: int = 32, mean_log10=5.0, std_log10=1.0, corr_len=8.0):
        import torch
        from . import TetParams, F32X2
        self.ctx, self.n = ctx, n
        self.field = (mean_log10, std_log10, corr_len)
        prob = meshes.tet_problem(n, seed=0)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        self.dev = {"x": dev(prob["x"]), "c": dev(prob["c"].astype(np.int32)),
                    "endpoints": dev(prob["endpoints"].astype(np.int32)),
                    "mu": dev(prob["mu"]), "lam": dev(prob["lam"])}
        self.idx = ctx.build_index(prob["nverts"], [self.dev["endpoints"]])
        self.params = TetParams(1.0, 0.0, (C.c_double * 3)(*prob["load"]), 1.0)
        self.hi = torch.empty(self.idx.nnz * 9, dtype=torch.float32, device="cuda")
        self.lo = torch.empty_like(self.hi)
        self.rhs = torch.empty(prob["nverts"] * 3, dtype=torch.float64, device="cuda")
        self.A = self.idx.bsr(3, F32X2, self.hi, self.lo)
        self._prec = F32X2

    def sample(self, seed: int):
        """Assemble the system of material seed (returns the resident Bsr)."""
        mean, std, corr = self.field
        g = device_random_field(self.n, corr, seed)
        E = (10.0 ** (mean + std * g)).repeat_interleave(5)
        nu = 0.3
        self.dev["mu"].copy_(E / (2.0 * (1.0 + nu)))
        self.dev["lam"].copy_(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)))
        self.ctx.assemble_tets(self.idx, self.params, self.dev, self.dev, self._prec, self.hi,
                               self.lo, self.rhs)
        return self.A


class HexSampler:
    """One resident n^3 trilinear-hex Poisson mesh (T = 0 on x = 0, uniform
    source) whose conductivity is redrawn per sample as 10^(std * GRF); the
    correlation length is in cells, so a small training mesh sees the same
    local statistics as the full-size system.  Scalar (1x1) systems."""

    def __init__(self, ctx, n: int = 48, std_log10=1.5, corr_len=12.0):
        import torch
        from . import HexParams, F32X2
        self.ctx, self.n = ctx, n
        self.field = (std_log10, corr_len)
        prob = meshes.hex_problem(n, seed=0, std_log10=std_log10, corr_len=corr_len)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        self.dev = {"dirichlet": dev(prob["dirichlet"]), "kappa": dev(prob["kappa"]),
                    "endpoints": dev(prob["endpoints"])}
        self.idx = ctx.build_index(prob["nverts"], [self.dev["endpoints"]])
        self.params = HexParams(prob["h"], prob["source"], 1.0)
        self.hi = torch.empty(self.idx.nnz, dtype=torch.float32, device="cuda")
        self.lo = torch.empty_like(self.hi)
        self.rhs = torch.empty(prob["nverts"], dtype=torch.float64, device="cuda")
        self.A = self.idx.bsr(1, F32X2, self.hi, self.lo)
        self._prec = F32X2

    def sample(self, seed: int):
        """Assemble the system of conductivity seed (returns the resident Bsr)."""
        import torch
        std, corr = self.field
        g = device_random_field(self.n, corr, seed)
        self.dev["kappa"].copy_(10.0 ** (std * g))
        self.ctx.assemble_hex(self.idx, self.params, self.dev, self.dev, self._prec, self.hi,
                              self.lo, self.rhs)
        return self.A


def train(ctx, net, n: int = 32, nsamples: int = 2000, epochs: int = 1, lr: float = 3e-3,
          lr_final: float = 3e-4, nprobes: int = 4, seed0: int = 10_000, log_every: int = 0,
          log=None, sampler=None):
    """Train net on nsamples seeded systems (n^3 x 5 tets unless a sampler,
    e.g. HexSampler, is given); returns a dict of timings and the loss curve
    (the loss is read back only every log_every steps)."""
    smp = sampler or TetSampler(ctx, n)
    steps = nsamples * epochs
    curve = []
    t_gen = t_gpu = 0.0
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        seed = seed0 + (s % nsamples)
        # the sample's time is its own: the queue is drained before and after
        # it (the step before and the draw itself run asynchronously)
        torch.cuda.synchronize()
        a = time.perf_counter()
        A = smp.sample(seed)
        torch.cuda.synchronize()
        b = time.perf_counter()
        # cosine decay from lr to lr_final
        cur = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * s / max(steps - 1, 1)))
        want = bool(log_every) and (s % log_every == 0 or s == steps - 1)
        loss = net.train_step(A, nprobes=nprobes, seed=seed * 7919 + s, lr=cur, want_loss=want)
        c = time.perf_counter()
        t_gen += b - a
        t_gpu += c - b
        if want:
            curve.append((s, loss))
            if log:
                log(f"[train] step {s} loss {loss:.4f} lr {cur:.2e}")
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    return {"steps": steps, "seconds": total, "sample_seconds": t_gen,
            "step_seconds": total - t_gen, "curve": curve}
