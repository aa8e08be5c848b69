"""Write the golden vectors of tests/golden/ from the reference's own CPU codec.

The reference codec is include/frsz2.hpp of robertu94/frsz compiled by
``make -C oracle ref`` into oracle/_ref/libfrsz2_ref.so (that build needs the
reference checkout, so this script runs only where it is present).  Output
buffers are zero-filled before the reference writes them, so the bytes it
never writes (exponent pad, unused tail words of a partial last block) are 0.

Files written:
  cases_f32.npz / cases_f64.npz  concatenated inputs, streams, decodes and
                                 gathers of many (bits, B, n) cases, with an
                                 index table
  plain_roundtrip_f32.json       hashes for the 2^20-value float32 smooth
                                 signal at bits 16, B 32
  dispatch_errors.json           the reference dispatch's messages for
                                 unsupported bits / block sizes

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from frsz_amd.workloads import lognormal_edge, smooth_signal  # noqa: E402
from oracle.oracle import Reference  # noqa: E402

REPEAT9 = np.array([1.0, 2.0, 3.0, 4.0, 0.25, -0.25, -0.125, 1 / 32.0, 0.125])


def edge_values(rng, n, dtype):
    """Mixed magnitudes with exact zeros inside non-zero blocks, all-zero
    runs, subnormals, negative zero and random signs."""
    W = 32 if dtype == np.float32 else 64
    span = 30 if W == 32 else 200
    x = np.exp2(rng.uniform(-span, span, n)) * rng.standard_normal(n)
    x[rng.random(n) < 0.08] = 0.0
    if n > 80:
        s = int(rng.integers(0, n - 70))
        x[s : s + 70] = 0.0
    x = x.astype(dtype)
    sub = rng.random(n) < 0.03
    if W == 32:
        x[sub] = rng.integers(1, 1 << 23, int(sub.sum()), dtype=np.uint32).view(np.float32)
    else:
        x[sub] = rng.integers(1, 1 << 52, int(sub.sum()), dtype=np.uint64).view(np.float64)
    neg0 = rng.random(n) < 0.02
    x[neg0] = -0.0
    flip = rng.random(n) < 0.5
    x[flip] = -x[flip]
    return x


def cases_for(dtype, ref, rng):
    W = 32 if dtype == np.float32 else 64
    out = []
    # the reference GPU test's 9-value pattern (test/test_frsz_gpu.cu:382-410)
    if W == 64:
        pat = [(32, 32), (16, 32), (16, 8), (15, 8), (9, 8), (9, 4), (9, 5)]
        n = 2049
    else:
        pat = [(16, 8), (15, 8), (9, 4), (9, 5)]
        n = 2160
    for bits, B in pat:
        out.append(("repeat9", np.resize(REPEAT9, n).astype(dtype), bits, B))
    # b = W, B = 1 random values in (-1e8, 1e8) (test/test_frsz_gpu.cu:422-441, seeded here)
    out.append(("uniform1e8", rng.uniform(-1e8, 1e8, 3000).astype(dtype), W, 1))
    bits_list = [4, 5, 7, 8, 9, 12, 15, 16, 17, 21, 24, 31, 32]
    if W == 64:
        bits_list += [33, 40, 47, 52, 53, 60, 63, 64]
    blocks = [1, 2, 3, 4, 5, 7, 8, 9, 16, 32, 64, 128, 256, 512, 1024]
    for bits in bits_list:
        for B in blocks:
            if rng.random() > 0.45 and B not in (16, 32, 64):
                continue
            n = int(rng.integers(1, 700)) if B < 256 else int(rng.integers(B - 5, 3 * B + 7))
            x = edge_values(rng, n, dtype)
            if W == 64 and bits <= 8:
                # exponents inside int8 range: the reference's int8 exponent
                # truncation (include/frsz2.hpp:510) is then well defined
                x = (np.exp2(rng.uniform(-100, 100, n)) * rng.standard_normal(n)).astype(dtype)
                x[: min(n, 2 * B)] = 0.0
            out.append(("edge", x, bits, B))
    # the smooth signal and the lognormal edge generator at config shapes (short)
    for bits, B in ((16, 32), (32, 32), (12, 16), (21, 16), (8, 64)):
        if W == 32:
            out.append(("lognormal", lognormal_edge(4099, seed=bits * 100 + B), bits, B))
        out.append(("smooth", smooth_signal(5000, seed=7, dtype=dtype), bits, B))
    if W == 32:
        x = rng.standard_normal(600).astype(np.float32)
        x[rng.random(600) < 0.1] = np.nan
        out.append(("nan", x, 16, 16))
        out.append(("nan", x, 21, 32))
    out.append(("empty", np.zeros(0, dtype), 16, 32))
    out.append(("single", np.array([3.5], dtype), 12, 64))
    return out


def write_cases(dtype, path, ref, rng):
    rows, names, xs, ss, ys, gi, gv = [], [], [], [], [], [], []
    xo = so = go = 0
    for name, x, bits, B in cases_for(dtype, ref, rng):
        n = x.size
        stream = ref.compress(x, bits, B, fill=0)
        y = ref.decompress(stream, n, dtype, bits, B)
        m = min(n, 64)
        idx = rng.integers(0, n, m).astype(np.uint64) if n else np.zeros(0, np.uint64)
        g = ref.gather(stream, n, dtype, bits, B, idx)
        rows.append((bits, B, n, xo, so, stream.size, go, m))
        names.append(name)
        xs.append(x)
        ss.append(stream)
        ys.append(y)
        gi.append(idx)
        gv.append(g)
        xo += n
        so += stream.size
        go += m
    np.savez_compressed(
        path,
        table=np.array(rows, dtype=np.int64),
        names=np.array(names),
        x=np.concatenate(xs),
        streams=np.concatenate(ss),
        y=np.concatenate(ys),
        gidx=np.concatenate(gi),
        gval=np.concatenate(gv),
    )
    return len(rows)


def main():
    ref = Reference()
    rng = np.random.default_rng(20240601)
    n32 = write_cases(np.float32, os.path.join(HERE, "cases_f32.npz"), ref, rng)
    n64 = write_cases(np.float64, os.path.join(HERE, "cases_f64.npz"), ref, rng)

    x = smooth_signal(1 << 20, seed=1)
    stream = ref.compress(x, 16, 32, fill=0)
    y = ref.decompress(stream, x.size, np.float32, 16, 32)
    meta = {
        "generator": "smooth_signal(n=2**20, seed=1, float32)",
        "bits": 16,
        "block": 32,
        "n": x.size,
        "stream_bytes": int(stream.size),
        "sha256_input": hashlib.sha256(x.tobytes()).hexdigest(),
        "sha256_stream": hashlib.sha256(stream.tobytes()).hexdigest(),
        "sha256_decoded": hashlib.sha256(y.tobytes()).hexdigest(),
        "max_abs_error": float(np.max(np.abs(y.astype(np.float64) - x.astype(np.float64)))),
    }
    with open(os.path.join(HERE, "plain_roundtrip_f32.json"), "w") as f:
        json.dump(meta, f, indent=1)

    errs = []
    xs = np.ones(40, np.float32)
    for bits, B in ((5, 32), (33, 32), (16, 3), (16, 2048), (16, 0)):
        rc, msg = ref.dispatch_compress(xs, bits, B)
        errs.append({"dtype": "float32", "bits": bits, "block": B, "rc": rc, "message": msg})
    with open(os.path.join(HERE, "dispatch_errors.json"), "w") as f:
        json.dump(errs, f, indent=1)
    print(f"wrote {n32} float32 cases, {n64} float64 cases")


if __name__ == "__main__":
    main()
