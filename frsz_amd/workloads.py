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


def serving_trace(seconds: float, rate: float, source_values: dict, seed: int = 5, lo_bytes: int = 4096,
                  hi_bytes: int = 256 << 20, recent: int = 64):
    """Request stream of the serving workload: Poisson arrivals at ``rate``
    per second for ``seconds``; 50% compress, 30% full decompress, 20%
    random-access reads of 1-4096 values; sizes log-uniform over
    [lo_bytes, hi_bytes]; float32 / float64 half each; bits from {16, 32}.
    A compress reads ``n`` values at offset ``offset`` (a multiple of 4
    values) of a source buffer of ``source_values[is_f64]`` values; a
    decompress or gather reads the stream of one of the ``recent`` latest
    compress requests (``source``, a trace index) with its dtype, bits and n.

    Returns numpy arrays: kind (0/1/2), f64, bits, n, offset, source, m,
    arrival (seconds)."""
    rng = np.random.default_rng(seed)
    nreq = int(seconds * rate)
    arrival = np.cumsum(rng.exponential(1.0 / rate, nreq))
    kind = rng.choice(3, size=nreq, p=[0.5, 0.3, 0.2])
    if nreq:
        kind[0] = 0
    f64 = rng.random(nreq) < 0.5
    bits = np.where(rng.random(nreq) < 0.5, 16, 32)
    nbytes = np.exp(rng.uniform(np.log(lo_bytes), np.log(hi_bytes), nreq))
    n = np.zeros(nreq, dtype=np.int64)
    offset = np.zeros(nreq, dtype=np.int64)
    source = np.zeros(nreq, dtype=np.int64)
    latest = []
    for i in range(nreq):
        if kind[i] != 0 and not latest:
            kind[i] = 0
        if kind[i] == 0:
            vs = 8 if f64[i] else 4
            n[i] = max(1, int(nbytes[i]) // vs)
            total = int(source_values[bool(f64[i])])
            offset[i] = int(rng.integers(0, (total - n[i]) // 4 + 1)) * 4
            latest.append(i)
            if len(latest) > recent:
                latest.pop(0)
        else:
            s = latest[int(rng.integers(0, len(latest)))]
            source[i] = s
            n[i] = n[s]
            f64[i] = f64[s]
            bits[i] = bits[s]
    m = np.where(kind == 2, rng.integers(1, 4097, nreq), 0)
    return {"kind": kind, "f64": f64, "bits": bits, "n": n, "offset": offset, "source": source, "m": m,
            "arrival": arrival}


def serving_requests(tr: dict, sources: dict, block: int, keep, seed: int = 6):
    """Request records (frsz_amd.serve.REQUEST_DTYPE) for a ``serving_trace``
    over the device source buffers ``sources[is_f64]``.  Gather indices are
    drawn on the GPU, uniform over each request's n.  ``keep`` (bool per
    request) marks the requests whose output stays held; the rest discard it.
    Returns (records, idx_all, idx_off): request i's indices are
    idx_all[idx_off[i] : idx_off[i] + m[i]]."""
    import torch

    from .serve import REQ_DISCARD, REQUEST_DTYPE

    kind, f64, n, m = tr["kind"], tr["f64"], tr["n"], tr["m"]
    nreq = kind.size
    gi = np.nonzero(kind == 2)[0]
    idx_all = None
    idx_off = np.zeros(nreq, dtype=np.int64)
    if gi.size:
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        counts = torch.from_numpy(m[gi]).cuda()
        lim = torch.from_numpy(n[gi]).cuda()
        rep = torch.repeat_interleave(lim, counts)
        idx_all = (torch.rand(rep.numel(), dtype=torch.float64, device="cuda", generator=g) * rep).long()
        idx_all = torch.minimum(idx_all, rep - 1)
        del rep
        idx_off[gi] = np.concatenate([[0], np.cumsum(m[gi])[:-1]])
    rec = np.zeros(nreq, dtype=REQUEST_DTYPE)
    rec["op"] = kind
    rec["dtype"] = f64.astype(np.int32)
    rec["bits"] = tr["bits"]
    rec["flags"] = np.where(keep, 0, REQ_DISCARD)
    rec["block"] = block
    rec["n"] = n
    base32, base64 = sources[False].data_ptr(), sources[True].data_ptr()
    offs = tr["offset"]
    rec["values"] = np.where(kind == 0, np.where(f64, base64 + offs * 8, base32 + offs * 4), 0)
    rec["source"] = tr["source"]
    rec["idx"] = np.where(kind == 2, (idx_all.data_ptr() if idx_all is not None else 0) + idx_off * 8, 0)
    rec["m"] = m
    rec["arrival"] = tr["arrival"]
    return rec, idx_all, idx_off


def batched_plan_arrays(count: int = 4096, seed: int = 4, target_input_bytes=None):
    """The batched workload's arrays: ``batched_lengths``, optionally with
    every length divided by one factor so that the input totals
    ``target_input_bytes``.  Returns (lengths, is_f64, bits, scale); array i
    holds ``smooth_signal_torch(lengths[i], seed=10_000 + i)`` in its dtype."""
    lengths, is_f64, bits = batched_lengths(count, seed=seed)
    scale = 1.0
    if target_input_bytes:
        full = float(np.sum(lengths * np.where(is_f64, 8, 4)))
        scale = full / target_input_bytes
        lengths = np.maximum(1, np.round(lengths / scale)).astype(np.int64)
    return lengths, is_f64, bits, scale
