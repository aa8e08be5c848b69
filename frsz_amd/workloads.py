"""Seeded synthetic inputs of the benchmark workloads.

* ``smooth_signal``   float32/64 x_i = sin(i/500) + 0.3 sin(i/37) + N(0, 1e-3)
                      (the plain CPU round trip of 2^20 values, b = 16, B = 32,
                      and the per-array generator of the batched workload).
* ``gaussian_field``  a 3-D Gaussian random field with power spectrum k^-3 plus
                      a constant offset of 1e3, float64 (512^3 = 1 GiB on the
                      GPU, b = 32, B = 32).
* ``lognormal_edge``  float32 magnitudes 2^N(0,10) clipped to +-30, random signs,
                      10% all-zero 64-value runs, 1% subnormals (the edge-case
                      sweep over b in {8,12,16,21,32} x B in {16,32,64}).

numpy versions run on the host; ``*_torch`` versions generate on the GPU.
"""
from __future__ import annotations

import math

import numpy as np


def smooth_signal(n: int, seed: int = 1, dtype=np.float32) -> np.ndarray:
    rng = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.float64)
    x = np.sin(i / 500.0) + 0.3 * np.sin(i / 37.0) + rng.normal(0.0, 1e-3, n)
    return x.astype(dtype)


def lognormal_edge(n: int, seed: int = 3) -> np.ndarray:
    rng = np.random.default_rng(seed)
    e = np.clip(rng.normal(0.0, 10.0, n), -30.0, 30.0)
    x = np.exp2(e) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    x = x.astype(np.float32)
    runs = n // 64
    zero_runs = rng.random(runs) < 0.10
    x[: runs * 64].reshape(runs, 64)[zero_runs] = 0.0
    sub = rng.random(n) < 0.01
    mant = rng.integers(1, 1 << 23, int(sub.sum()), dtype=np.uint32)
    sign = (rng.random(int(sub.sum())) < 0.5).astype(np.uint32) << np.uint32(31)
    x[sub] = (mant | sign).view(np.float32)
    return x


def gaussian_field_torch(side: int, seed: int = 2, device="cuda", offset: float = 1e3):
    """float64 field of side^3 values, C-order flattened by the caller."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    noise = torch.randn((side, side, side), dtype=torch.float64, device=device, generator=g)
    spec = torch.fft.rfftn(noise)
    del noise
    k = torch.fft.fftfreq(side, device=device, dtype=torch.float64) * side
    kz = torch.fft.rfftfreq(side, device=device, dtype=torch.float64) * side
    k2 = k.view(-1, 1, 1) ** 2 + k.view(1, -1, 1) ** 2 + kz.view(1, 1, -1) ** 2
    k2[0, 0, 0] = 1.0
    amp = k2.pow_(-0.75)  # |k|^-3/2 amplitude -> power k^-3
    amp[0, 0, 0] = 0.0
    spec.mul_(amp)
    del amp, k2
    field = torch.fft.irfftn(spec, s=(side, side, side))
    del spec
    field.sub_(field.mean()).div_(field.std()).add_(offset)
    return field


def smooth_signal_torch(n: int, seed: int, dtype, device="cuda"):
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    i = torch.arange(n, dtype=torch.float64, device=device)
    x = torch.sin(i / 500.0).add_(torch.sin(i / 37.0), alpha=0.3)
    x.add_(torch.randn(n, dtype=torch.float64, device=device, generator=g), alpha=1e-3)
    return x.to(dtype)


def batched_lengths(count: int = 4096, lo: float = 1e3, hi: float = 1.6e7, seed: int = 4):
    """Array lengths log-uniform over [lo, hi], 60% float32 / 40% float64,
    bits drawn from {16, 32}."""
    rng = np.random.default_rng(seed)
    lengths = np.exp(rng.uniform(math.log(lo), math.log(hi), count)).astype(np.int64)
    is_f64 = rng.random(count) < 0.4
    bits = np.where(rng.random(count) < 0.5, 16, 32)
    return lengths, is_f64, bits
