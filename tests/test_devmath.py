"""The value-level device arithmetic (include/frsz2_device.hpp: encode_value,
decode_value and their narrow-width forms encode_value_short and
decode_value_short, max_key, block_exponent_from_key), compiled
for the host into tests/native/libdevmath_host.so, against the oracle on the
CPU: every float32 exponent E and every bits value over a sample of bit
patterns per biased exponent (edges included), float64 on a stride of E.
The HIP kernels use exactly these functions, so this pins their arithmetic
independently of the GPU."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "tests", "native", "libdevmath_host.so")
_vp = ctypes.c_void_p


@pytest.fixture(scope="module")
def dm():
    if not os.path.exists(LIB):
        pytest.fail("tests/native/libdevmath_host.so not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def roundtrip(lib, name, u, E, bits, oracle):
    c = np.empty_like(u)
    v = np.empty_like(u)
    e = ctypes.c_int16(E) if oracle else ctypes.c_int(E)
    getattr(lib, name)(_vp(u.ctypes.data), ctypes.c_uint64(u.size), e, ctypes.c_int(bits), _vp(c.ctypes.data),
                       _vp(v.ctypes.data))
    return c, v


def apply(lib, name, codes, E, bits):
    out = np.empty_like(codes)
    getattr(lib, name)(_vp(codes.ctypes.data), ctypes.c_uint64(codes.size), ctypes.c_int(E), ctypes.c_int(bits),
                       _vp(out.ctypes.data))
    return out


def f32_patterns(rng, per_exp):
    mant = np.concatenate([np.array([0, 1, 2, 0x7FFFFF, 0x400000, 0x3FFFFF], np.uint32),
                           rng.integers(0, 1 << 23, per_exp, dtype=np.uint32)])
    be = np.arange(0, 255, dtype=np.uint32)
    u = (be[:, None] << np.uint32(23) | mant[None, :]).ravel()
    return np.concatenate([u, u | np.uint32(0x80000000)])


def test_float32_encode_decode_match_oracle(dm, port):
    u = f32_patterns(np.random.default_rng(0), 250)
    exps = ((u >> np.uint32(23)) & np.uint32(0xFF)).astype(np.int64) - 127
    for E in range(-127, 128):
        sel = np.ascontiguousarray(u[exps <= E])
        for bits in range(4, 33):
            c1, v1 = roundtrip(port.lib, "fo_roundtrip_f32", sel, E, bits, True)
            c2, v2 = roundtrip(dm, "dm_roundtrip_f32", sel, E, bits, False)
            assert np.array_equal(c1, c2), ("codes", E, bits)
            assert np.array_equal(v1, v2), ("values", E, bits)


def test_float64_encode_decode_match_oracle(dm, port):
    rng = np.random.default_rng(1)
    mant = np.concatenate([np.array([0, 1, 2, (1 << 52) - 1, 1 << 51], np.uint64),
                           rng.integers(0, 1 << 52, 60, dtype=np.uint64)])
    be = np.arange(0, 2047, 3, dtype=np.uint64)
    u = (be[:, None] << np.uint64(52) | mant[None, :]).ravel()
    u = np.concatenate([u, u | np.uint64(1 << 63)])
    exps = ((u >> np.uint64(52)) & np.uint64(0x7FF)).astype(np.int64) - 1023
    for E in list(range(-1023, 1024, 37)) + [-1022, 0, 1023]:
        sel = np.ascontiguousarray(u[exps <= E])
        for bits in range(4, 65):
            if bits <= 8 and not -128 <= E <= 127:
                continue  # int8 block exponent: the reference truncates E there
            c1, v1 = roundtrip(port.lib, "fo_roundtrip_f64", sel, E, bits, True)
            c2, v2 = roundtrip(dm, "dm_roundtrip_f64", sel, E, bits, False)
            assert np.array_equal(c1, c2), ("codes", E, bits)
            assert np.array_equal(v1, v2), ("values", E, bits)


def test_short_decode_float32_equals_generic(dm):
    """decode_value_short (bits <= 24) against decode_value: every code for
    bits <= 16, edges plus 4096 random codes above, every E of finite data."""
    rng = np.random.default_rng(2)
    for bits in range(4, 25):
        if bits <= 16:
            codes = np.arange(1 << bits, dtype=np.uint32)
        else:
            top = (1 << bits) - 1
            codes = np.concatenate([np.array([0, 1, top, top >> 1, 1 << (bits - 1), (1 << (bits - 1)) + 1],
                                             np.uint32), rng.integers(0, 1 << bits, 4096, dtype=np.uint32)])
        for E in range(-127, 128):
            a = apply(dm, "dm_decode_f32", codes, E, bits)
            b = apply(dm, "dm_decode_short_f32", codes, E, bits)
            assert np.array_equal(a, b), (bits, E)


def test_short_decode_float64_equals_generic(dm):
    rng = np.random.default_rng(3)
    for bits in range(4, 54):
        top = (1 << bits) - 1
        codes = np.concatenate([np.array([0, 1, top, top >> 1, 1 << (bits - 1), (1 << (bits - 1)) + 1], np.uint64),
                                rng.integers(0, 1 << bits, 1024, dtype=np.uint64)])
        for E in list(range(-1023, 1024, 11)) + [-1022, -1, 0, 1, 1023]:
            if bits <= 8 and not -128 <= E <= 127:
                continue
            a = apply(dm, "dm_decode_f64", codes, E, bits)
            b = apply(dm, "dm_decode_short_f64", codes, E, bits)
            assert np.array_equal(a, b), (bits, E)


@pytest.mark.parametrize("wide", [False, True])
def test_decodes_ignore_bits_above_the_code(dm, wide):
    """The packed-code extraction hands the decodes a whole word whose bits
    above `bits` belong to the next codes: both decodes must ignore them."""
    rng = np.random.default_rng(6)
    ut, W, SB, bias = (np.uint64, 64, 52, 1023) if wide else (np.uint32, 32, 23, 127)
    name = "f64" if wide else "f32"
    for bits in range(4, SB + 2):
        codes = rng.integers(0, 1 << bits, 2048, dtype=ut)
        codes[:8] = np.array([0, 1, (1 << (bits - 1)), (1 << bits) - 1, 2, 3, 1 << (bits - 2), 5], ut)
        junk = (rng.integers(0, np.iinfo(ut).max, 2048, dtype=ut, endpoint=True) >> ut(bits)) << ut(bits)
        for E in ([-bias, -bias + 1, -1, 0, 1, 30, bias] + list(rng.integers(-bias, bias + 1, 4))):
            E = int(E)
            if wide and bits <= 8 and not -128 <= E <= 127:
                continue
            want = apply(dm, "dm_decode_" + name, codes, E, bits)
            assert np.array_equal(apply(dm, "dm_decode_" + name, codes | junk, E, bits), want), (bits, E)
            assert np.array_equal(apply(dm, "dm_decode_short_" + name, codes | junk, E, bits), want), (bits, E)


def test_short_encode_float32_equals_generic(dm):
    """encode_value_short (bits <= 24) against encode_value on every biased
    exponent (NaN patterns included) for every block exponent E >= exponent(x)."""
    u = f32_patterns(np.random.default_rng(4), 120)
    nan = np.array([0x7FC00000, 0xFFC00001, 0x7F800001, 0x7FFFFFFF], np.uint32)
    exps = ((u >> np.uint32(23)) & np.uint32(0xFF)).astype(np.int64) - 127
    for E in range(-127, 128):
        sel = np.ascontiguousarray(np.concatenate([u[exps <= E], nan]))
        for bits in range(4, 25):
            a = apply(dm, "dm_encode_f32", sel, E, bits)
            b = apply(dm, "dm_encode_short_f32", sel, E, bits)
            assert np.array_equal(a, b), (bits, E)


def test_short_encode_float64_equals_generic(dm):
    rng = np.random.default_rng(5)
    mant = np.concatenate([np.array([0, 1, (1 << 52) - 1, 1 << 51], np.uint64),
                           rng.integers(0, 1 << 52, 40, dtype=np.uint64)])
    be = np.arange(0, 2047, 5, dtype=np.uint64)
    u = (be[:, None] << np.uint64(52) | mant[None, :]).ravel()
    u = np.concatenate([u, u | np.uint64(1 << 63)])
    exps = ((u >> np.uint64(52)) & np.uint64(0x7FF)).astype(np.int64) - 1023
    for E in list(range(-1023, 1024, 41)) + [-1022, 0, 1023]:
        sel = np.ascontiguousarray(u[exps <= E])
        for bits in range(4, 54):
            if bits <= 8 and not -128 <= E <= 127:
                continue
            a = apply(dm, "dm_encode_f64", sel, E, bits)
            b = apply(dm, "dm_encode_short_f64", sel, E, bits)
            assert np.array_equal(a, b), (bits, E)

