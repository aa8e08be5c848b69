"""Pins of the oracle (oracle/frsz2_oracle.c) before it is trusted as the
checker: the reference's own known-answer vectors, golden vectors written by
the reference's compiled CPU codec, the compiled reference itself (where it
is built), and closed forms / invariants of the format."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_CASES
from frsz_amd.workloads import smooth_signal
from oracle.oracle import compressed_bytes, storage_bits


def f32(bits):
    return np.array([bits], np.uint32).view(np.float32)[0]


# ---- known answers of the reference's tests

def test_decomposition_known_answers(port):
    # test/test_frsz.cc:529-552 and test/test_frsz_gpu.cu:213-229
    L = port.lib
    assert L.fo_significand_f32(32.0) == 1 << 23 and L.fo_exponent_f32(32.0) == 5 and L.fo_sign_f32(32.0) == 0
    assert L.fo_significand_f32(0.5) == 1 << 23 and L.fo_exponent_f32(0.5) == -1
    assert L.fo_significand_f64(32.0) == 1 << 52 and L.fo_exponent_f64(32.0) == 5
    assert L.fo_significand_f32(0.0) == 0 and L.fo_exponent_f32(0.0) == -127
    assert L.fo_significand_f64(0.0) == 0 and L.fo_exponent_f64(0.0) == -1023 and L.fo_sign_f64(0.0) == 0
    for v, s, e, m in ((32.0, 0, 5, 1 << 23), (0.5, 0, -1, 1 << 23), (0.0, 0, -127, 0), (-0.375, 1, -2, 3 << 22)):
        assert (L.fo_sign_f32(v), L.fo_exponent_f32(v), L.fo_significand_f32(v)) == (s, e, m)
    for v, s, e, m in ((-32.0, 1, 5, 1 << 52), (0.5, 0, -1, 1 << 52), (0.0, 0, -1023, 0),
                       (-0.375, 1, -2, 3 << 51)):
        assert (L.fo_sign_f64(v), L.fo_exponent_f64(v), L.fo_significand_f64(v)) == (s, e, m)
    # float exponent_t is int8: inf/NaN (biased 255) wraps to -128 (include/frsz2.hpp:280)
    assert L.fo_exponent_f32(float("inf")) == -128


def test_fixed_conversion_known_answers(port):
    L = port.lib
    # test/test_frsz_gpu.cu:276-301
    sub = f32(0b0_00000000_11111111111111111111111)
    inter32 = [(32.0, 5, 0x800000 << 7), (2.0, 5, 0x800000 << 3), (sub, 5, 0), (0.125, 2, 0x02000000),
               (-0.375, 1, 0x8C000000)]
    out32 = [32.0, 2.0, f32(0b0_01100101_00000000000000000000000), 0.125, -0.375]
    for (v, e, want), back in zip(inter32, out32):
        fx = L.fo_floating_to_fixed_f32(v, e)
        assert fx == want
        assert np.float32(L.fo_fixed_to_floating_f32(fx, e)).view(np.uint32) == np.float32(back).view(np.uint32)
    inter64 = [(32.0, 5, 1 << 62), (2.0, 5, 1 << 58), (-0.125, 2, 0x82 << 56), (0.375, -1, 3 << 60)]
    for v, e, want in inter64:
        fx = L.fo_floating_to_fixed_f64(v, e)
        assert fx == want
        assert L.fo_fixed_to_floating_f64(fx, e) == v
    # test/test_frsz.cc:554-563: round trips at E = 5 and the zero-code quirk 0x32800000 = 2^-26
    for v in (32.0, 2.0):
        assert L.fo_fixed_to_floating_f32(L.fo_floating_to_fixed_f32(v, 5), 5) == v
        assert L.fo_fixed_to_floating_f64(L.fo_floating_to_fixed_f64(v, 5), 5) == v
    sub2 = f32(0b0000000001111111111111111111111)
    got = np.float32(L.fo_fixed_to_floating_f32(L.fo_floating_to_fixed_f32(sub2, 5), 5)).view(np.uint32)
    assert got == 0b00110010100000000000000000000000


def test_zero_code_closed_form(port):
    # a zero magnitude code decodes to +-2^(E-31) when E + bias >= 31, for both
    # float and double (host countl_zero(0) == 32, include/frsz2.hpp:180-192)
    L = port.lib
    for E in (-90, -5, 0, 5, 100):
        assert L.fo_fixed_to_floating_f32(0, E) == 2.0 ** (E - 31)
        assert L.fo_fixed_to_floating_f32(1 << 31, E) == -(2.0 ** (E - 31))
        assert L.fo_fixed_to_floating_f64(0, E) == 2.0 ** (E - 31)
    # below that the exponent field saturates at 0: +-0
    assert L.fo_fixed_to_floating_f32(0, -127) == 0.0
    assert L.fo_fixed_to_floating_f64(0, -1023) == 0.0


# ---- layout closed form

@pytest.mark.parametrize("bits,B,n,want", [
    (16, 32, 1 << 20, 2162944),           # 2^20 float32 values at b=16, B=32
    (32, 32, 512 ** 3, 553648384),        # 512^3 float64 values at b=32, B=32
    (16, 32, 0, 256),
    (4, 1, 3, 256 + 3),
    (21, 16, 1600, 512 + 100 * 44),      # X = 400 B of int32 exponents -> R = 512
])
def test_layout_closed_form(port, bits, B, n, want):
    assert compressed_bytes(bits, B, n) == want
    assert port.lib.fo_compressed_bytes(bits, B, n) == want


# ---- golden vectors written by the compiled reference

@pytest.mark.parametrize("case", GOLDEN_CASES, ids=repr)
def test_oracle_matches_golden(port, case):
    s = port.compress(case.x, case.bits, case.B, fill=0)
    assert np.array_equal(s, case.stream)
    y = port.decompress(case.stream, case.n, case.dtype, case.bits, case.B)
    assert np.array_equal(y.view(np.uint8), case.y.view(np.uint8))
    g = port.gather(case.stream, case.n, case.dtype, case.bits, case.B, case.gidx)
    assert np.array_equal(g.view(np.uint8), case.gval.view(np.uint8))


def test_plain_roundtrip_golden_hashes(port):
    meta = json.load(open(os.path.join(GOLDEN, "plain_roundtrip_f32.json")))
    x = smooth_signal(meta["n"], seed=1)
    if hashlib.sha256(x.tobytes()).hexdigest() != meta["sha256_input"]:
        pytest.skip("this host's libm generates a different input; hashes not comparable")
    s = port.compress(x, meta["bits"], meta["block"])
    assert s.size == meta["stream_bytes"]
    assert hashlib.sha256(s.tobytes()).hexdigest() == meta["sha256_stream"]
    y = port.decompress(s, x.size, np.float32, meta["bits"], meta["block"])
    assert hashlib.sha256(y.tobytes()).hexdigest() == meta["sha256_decoded"]


def test_multithreaded_oracle_is_byte_identical(port):
    x = smooth_signal(200_003, seed=5, dtype=np.float64)
    p8 = port.with_threads(8)
    for bits, B in ((32, 32), (21, 16), (9, 5)):
        a = port.compress(x, bits, B)
        assert np.array_equal(a, p8.compress(x, bits, B))
        ya = port.decompress(a, x.size, np.float64, bits, B)
        assert np.array_equal(ya, p8.decompress(a, x.size, np.float64, bits, B))


# ---- invariants

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lossless_at_full_width_block_one(port, dtype):
    # b = W with B = 1 is lossless for finite values (test/test_frsz_gpu.cu:414-441)
    rng = np.random.default_rng(11)
    W = np.dtype(dtype).itemsize * 8
    u = rng.integers(0, 1 << (W - 1), 50000, dtype=np.uint64 if W == 64 else np.uint32)
    x = u.astype(u.dtype).view(dtype)
    x = x[np.isfinite(x)]
    s = port.compress(x, W, 1)
    y = port.decompress(s, x.size, dtype, W, 1)
    assert np.array_equal(y.view(u.dtype), x.view(u.dtype))


@pytest.mark.parametrize("dtype,bits", [(np.float32, 16), (np.float32, 12), (np.float64, 32), (np.float64, 21)])
def test_truncation_error_bound(port, dtype, bits):
    # |x - x^| < 2^(E - bits + 2) for non-zero codes (truncation toward zero)
    x = smooth_signal(40000, seed=9, dtype=dtype)
    B = 32
    s = port.compress(x, bits, B)
    y = port.decompress(s, x.size, dtype, bits, B)
    U = 8 if bits <= 8 else 16 if bits <= 16 else 32 if bits <= 32 else 64
    E = s[: (x.size // B) * U // 8].view({16: np.int16, 32: np.int32}[U]).astype(np.int64)
    err = np.abs(y[: E.size * B].astype(np.float64) - x[: E.size * B].astype(np.float64)).reshape(-1, B)
    bound = np.exp2(E - bits + 2)[:, None]
    assert np.all(err < bound)
    assert np.all(np.abs(y) <= np.abs(x) + 0)  # truncation never grows a magnitude here


def test_blocks_are_independent(port):
    # any range of whole blocks compresses to the matching slice of the stream
    x = smooth_signal(32 * 500, seed=2)
    bits, B = 12, 32
    full = port.compress(x, bits, B)
    part = port.compress(x[32 * 100: 32 * 300], bits, B)
    R_full = compressed_bytes(bits, B, x.size) - 500 * 48
    R_part = compressed_bytes(bits, B, 6400) - 200 * 48
    assert np.array_equal(full[R_full + 100 * 48: R_full + 300 * 48], part[R_part:])
    assert np.array_equal(full[200:600], part[:400])


# ---- against the compiled reference where it is built (this container)

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_matches_compiled_reference(port, reference, dtype):
    rng = np.random.default_rng(5)
    W = np.dtype(dtype).itemsize * 8
    for bits in range(4 if W == 32 else 9, W + 1):
        for B in (1, 2, 3, 5, 8, 16, 32, 64, 128, 512, 1024):
            n = int(rng.integers(1, 2500))
            x = (np.exp2(rng.uniform(-40, 40, n)) * rng.standard_normal(n)).astype(dtype)
            x[rng.random(n) < 0.1] = 0
            a = reference.compress(x, bits, B)
            assert np.array_equal(a, port.compress(x, bits, B)), (bits, B, n)
            ya = reference.decompress(a, n, dtype, bits, B)
            assert np.array_equal(ya.view(np.uint8), port.decompress(a, n, dtype, bits, B).view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_decodes_arbitrary_streams_like_compiled_reference(port, reference, dtype):
    """Streams no compress writes (random payload bytes; block exponents over
    the range where the reference's decode is defined: its shift counts stay
    below the width, E + bias >= -24 for float32 and >= -53 for float64, up to
    40000), which the GPU parity tests decode against the oracle."""
    rng = np.random.default_rng(12)
    W = np.dtype(dtype).itemsize * 8
    for bits in (4, 7, 8, 9, 12, 16, 17, 21, 24, 31, 32) + ((33, 40, 52, 53, 63, 64) if W == 64 else ()):
        for B in (1, 5, 16, 32, 64):
            n = int(rng.integers(1, 3000))
            U = storage_bits(bits)
            nb = -(-n // B)
            s = rng.integers(0, 256, compressed_bytes(bits, B, n), dtype=np.uint8)
            et = {8: np.int8, 16: np.int16, 32: np.int32, 64: np.int64}[U]
            lo = max(int(np.iinfo(et).min), -151 if W == 32 else -1076)
            hi = min(int(np.iinfo(et).max), 40_000)
            E = rng.integers(lo, hi, nb, endpoint=True).astype(et)
            E[: min(nb, 3)] = np.array([lo, hi, -1], dtype=et)[: min(nb, 3)]
            s[: nb * U // 8] = E.view(np.uint8)
            a = reference.decompress(s, n, dtype, bits, B)
            b = port.decompress(s, n, dtype, bits, B)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (bits, B, n)


def test_value_helpers_match_compiled_reference(port, reference):
    rng = np.random.default_rng(6)
    R = reference.lib
    u = rng.integers(0, 1 << 32, 20000, dtype=np.uint64).astype(np.uint32)
    for v in u.view(np.float32):
        if not np.isfinite(v):
            continue
        assert R.ref_exponent_f32a(v) == port.lib.fo_exponent_f32(v)
        assert R.ref_significand_f32a(v) == port.lib.fo_significand_f32(v)
        E = max(int(port.lib.fo_exponent_f32(v)), -127) + int(rng.integers(0, 40))
        E = min(E, 127)
        fx = R.ref_floating_to_fixed_f32a(v, E)
        assert fx == port.lib.fo_floating_to_fixed_f32(v, E)
        a = np.float32(R.ref_fixed_to_floating_f32a(fx, E)).view(np.uint32)
        b = np.float32(port.lib.fo_fixed_to_floating_f32(fx, E)).view(np.uint32)
        assert a == b


def test_block_entry_points_match_compiled_reference(port, reference):
    """compress_cpu_block(k, data, exp_block_elements) with any length up to
    the block (a partial block inside the array included) and
    decompress_cpu_block, into buffers of arbitrary prior contents."""
    rng = np.random.default_rng(11)
    for dtype in (np.float32, np.float64):
        W = np.dtype(dtype).itemsize * 8
        for bits in (4, 7, 8, 9, 12, 16, 21, 31, 32) + ((33, 40, 52, 64) if W == 64 else ()):
            for B in (1, 3, 5, 8, 9, 32, 128):
                n = int(rng.integers(1, 700))
                x = (rng.standard_normal(n) * np.exp2(rng.integers(-20, 20, n))).astype(dtype)
                x[rng.random(n) < 0.1] = 0
                nb = -(-n // B)
                for _ in range(2):
                    k = int(rng.integers(0, nb))
                    ln = int(rng.integers(1, min(B, n - k * B) + 1))
                    s1 = np.full(compressed_bytes(bits, B, n), int(rng.integers(0, 256)), np.uint8)
                    s2 = s1.copy()
                    port.compress_block(s1, n, bits, B, k, x[k * B: k * B + ln])
                    reference.compress_block(s2, n, bits, B, k, x[k * B: k * B + ln])
                    assert np.array_equal(s1, s2), (dtype, bits, B, n, k, ln)
                    a = port.decompress_block(s1, n, dtype, bits, B, k)
                    b = reference.decompress_block(s2, n, dtype, bits, B, k)
                    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))

