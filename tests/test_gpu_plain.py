"""Parity of the plain HIP kernels (compress / decompress / gather through the
C ABI) with the golden vectors of the reference codec and with the oracle.

Bar: bit-exact stream bytes, decoded bit patterns and gathered values."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_CASES

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def fz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frsz_amd

    return frsz_amd


def tdt(dtype):
    return torch.float32 if np.dtype(dtype) == np.float32 else torch.float64


def ubits(a):
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def gpu_compress(fz, x, bits, B):
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    s = fz.compress(xd, bits, B)
    return s


def gpu_roundtrip(fz, x, bits, B):
    s = gpu_compress(fz, x, bits, B)
    y = fz.decompress(s, x.size, tdt(x.dtype), bits, B)
    torch.cuda.synchronize()
    return s.cpu().numpy(), y.cpu().numpy()


@pytest.mark.parametrize("case", GOLDEN_CASES, ids=repr)
def test_golden_vectors(fz, case):
    print(f"golden {case!r}", flush=True)
    s, y = gpu_roundtrip(fz, case.x, case.bits, case.B)
    assert np.array_equal(s, case.stream), "stream bytes differ from the reference"
    assert np.array_equal(ubits(y), ubits(case.y)), "decoded values differ from the reference"
    if case.gidx.size:
        st = torch.from_numpy(case.stream).cuda()
        idx = torch.from_numpy(case.gidx.astype(np.int64)).cuda()
        g = fz.gather(st, case.n, tdt(case.dtype), case.bits, case.B, idx).cpu().numpy()
        assert np.array_equal(ubits(g), ubits(case.gval))


def test_plain_roundtrip_2p20_float32(fz, port):
    """2^20 smooth-signal float32 values at bits 16, B 32: golden hashes."""
    from frsz_amd.workloads import smooth_signal

    meta = json.load(open(os.path.join(GOLDEN, "plain_roundtrip_f32.json")))
    x = smooth_signal(meta["n"], seed=1)
    s, y = gpu_roundtrip(fz, x, meta["bits"], meta["block"])
    assert s.size == meta["stream_bytes"]
    if hashlib.sha256(x.tobytes()).hexdigest() == meta["sha256_input"]:
        assert hashlib.sha256(s.tobytes()).hexdigest() == meta["sha256_stream"]
        assert hashlib.sha256(y.tobytes()).hexdigest() == meta["sha256_decoded"]
    assert np.array_equal(s, port.compress(x, meta["bits"], meta["block"]))
    assert np.array_equal(ubits(y), ubits(port.decompress(s, x.size, np.float32, meta["bits"], meta["block"])))


def mixed_values(rng, n, dtype):
    W = 32 if dtype == np.float32 else 64
    span = 30 if W == 32 else 300
    x = (np.exp2(rng.uniform(-span, span, n)) * rng.standard_normal(n)).astype(dtype)
    x[rng.random(n) < 0.05] = 0
    if n > 200:
        s = int(rng.integers(0, n - 150))
        x[s:s + 150] = 0
    sub = rng.random(n) < 0.01
    if W == 32:
        x[sub] = rng.integers(1, 1 << 23, int(sub.sum()), dtype=np.uint32).view(np.float32)
    else:
        x[sub] = rng.integers(1, 1 << 52, int(sub.sum()), dtype=np.uint64).view(np.float64)
    return x


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("B", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 3, 5, 6, 7, 9, 12, 100, 2048, 24, 4096,
                               5000])
def test_random_parity_all_bits(fz, port, dtype, B):
    """Every bit width 4..W against the oracle, sizes spanning several tiles
    plus a ragged tail.  Power-of-two B up to 4096 take the tiled kernels
    (beyond a 64-thread tile, the 256-thread one); other sizes up to 4096 the
    generic tiles (block maxima by LDS atomics below B = 6, by the key array
    from 6 on); B = 1 and B = 5000 the one-thread-per-block compress and the
    block-max pass with one thread per payload word."""
    rng = np.random.default_rng(B * 7 + (dtype == np.float64))
    W = 32 if dtype == np.float32 else 64
    for bits in range(4, W + 1):
        if W == 64 and bits <= 8:
            continue
        n = int(rng.integers(1, 9000))
        x = mixed_values(rng, n, dtype)
        print(f"{np.dtype(dtype).name} bits={bits} B={B} n={n}", flush=True)
        s, y = gpu_roundtrip(fz, x, bits, B)
        want = port.compress(x, bits, B)
        assert np.array_equal(s, want), (bits, B, n)
        assert np.array_equal(ubits(y), ubits(port.decompress(want, n, dtype, bits, B))), (bits, B, n)


@pytest.mark.parametrize("bits,B", [(8, 16), (8, 64), (7, 32), (5, 1), (4, 1024)])
def test_double_low_bits_defined_inputs(fz, port, bits, B):
    """float64 with bits <= 8 stores int8 block exponents; inputs whose
    exponents fit int8 (and all-zero blocks) are well defined."""
    rng = np.random.default_rng(bits + B)
    n = 5000
    x = np.exp2(rng.uniform(-100, 100, n)) * rng.standard_normal(n)
    x[:2 * B] = 0
    s, y = gpu_roundtrip(fz, x, bits, B)
    want = port.compress(x, bits, B)
    assert np.array_equal(s, want)
    assert np.array_equal(ubits(y), ubits(port.decompress(want, n, np.float64, bits, B)))


@pytest.mark.parametrize("n", [0, 1, 31, 32, 33, 1023, 1024, 1025, 4097, 262143])
@pytest.mark.parametrize("bits,B", [(16, 32), (12, 16), (21, 64), (32, 1)])
def test_ragged_sizes(fz, port, n, bits, B):
    rng = np.random.default_rng(n)
    x = mixed_values(rng, n, np.float32)
    s, y = gpu_roundtrip(fz, x, bits, B)
    assert np.array_equal(s, port.compress(x, bits, B))
    assert np.array_equal(ubits(y), ubits(x if bits == 32 and B == 1 else port.decompress(s, n, np.float32, bits, B)))


def test_no_write_outside_stream(fz):
    """The codec writes exactly the stream's bytes: a guard band after it
    keeps its fill (the reference GPU test's 0xFF + 128 B guard,
    test/test_frsz_gpu.cu:331-359)."""
    rng = np.random.default_rng(3)
    for bits, B, n in ((16, 32, 5000), (21, 16, 4099), (9, 5, 2049), (32, 32, 1 << 16)):
        x = torch.from_numpy(mixed_values(rng, n, np.float32)).cuda()
        size = fz.compressed_size(torch.float32, bits, B, n)
        buf = torch.full((size + 4096,), 0xAB, dtype=torch.uint8, device="cuda")
        fz.compress_into(x, buf, bits, B)
        tail = buf[size:].cpu().numpy()
        assert np.all(tail == 0xAB)


@pytest.mark.parametrize("bits,B", [(16, 32), (12, 8), (21, 64), (12, 4096), (16, 5000), (32, 1)])
def test_values_off_16_byte_alignment(fz, port, bits, B):
    """Values 4 bytes off 16-byte alignment: compress loads them one by one
    in the 4096-value tile (power-of-two B up to 4096) or takes the generic
    kernels (B = 5000); decompress stores them through the strided decode."""
    rng = np.random.default_rng(4 + B)
    n = 10001
    x = mixed_values(rng, n + 1, np.float32)
    xd = torch.from_numpy(x).cuda()[1:]  # 4-byte offset: not 16-byte aligned
    s = fz.compress(xd, bits, B)
    want = port.compress(x[1:], bits, B)
    assert np.array_equal(s.cpu().numpy(), want)
    out = torch.empty(n + 1, dtype=torch.float32, device="cuda")[1:]
    fz.decompress_into(s, out, bits, B)
    assert np.array_equal(ubits(out.cpu().numpy()), ubits(port.decompress(want, n, np.float32, bits, B)))


@pytest.mark.parametrize("dtype,bits,B", [(np.float32, 16, 32), (np.float32, 12, 32), (np.float64, 32, 64),
                                          (np.float64, 21, 16)])
def test_stream_and_values_off_16_byte_alignment(fz, port, dtype, bits, B):
    """A stream 8 bytes off 16-byte alignment and values one element off it:
    unpacked codes still take the tiles (strided decode, word loads), packed
    ones the generic kernels; the bytes are the oracle's either way."""
    rng = np.random.default_rng(bits * 31 + B)
    n = 20011
    x = mixed_values(rng, n, dtype)
    want = port.compress(x, bits, B)
    sbuf = torch.zeros(want.size + 8, dtype=torch.uint8, device="cuda")
    s = sbuf[8:]
    assert s.data_ptr() % 16 == 8
    s.copy_(torch.from_numpy(want))
    out = torch.empty(n + 1, dtype=tdt(dtype), device="cuda")[1:]
    assert out.data_ptr() % 16 != 0
    fz.decompress_into(s, out, bits, B)
    assert np.array_equal(ubits(out.cpu().numpy()), ubits(port.decompress(want, n, dtype, bits, B)))
    aligned = torch.empty(n, dtype=tdt(dtype), device="cuda")
    fz.decompress_into(s, aligned, bits, B)
    assert np.array_equal(ubits(aligned.cpu().numpy()), ubits(port.decompress(want, n, dtype, bits, B)))
    xs = torch.from_numpy(x).cuda()
    s.zero_()
    fz.compress_into(xs, s, bits, B)
    assert np.array_equal(s.cpu().numpy(), want)


@pytest.mark.parametrize("dtype,bits,B", [(np.float32, 8, 3), (np.float32, 16, 100), (np.float32, 32, 7),
                                          (np.float64, 16, 5), (np.float64, 32, 12), (np.float32, 8, 1000)])
def test_generic_unpacked_decode_into_any_alignment(fz, port, dtype, bits, B):
    """Unpacked codes (bits = U) of block sizes that are not powers of two
    decode four consecutive values per lane on the generic tiles: 16-byte
    stores into aligned values, single stores into values one element off
    16-byte alignment, and partial chunks at the end of the array."""
    rng = np.random.default_rng(bits * 13 + B)
    for n in (1, 3, 4095, 20011):
        x = mixed_values(rng, n, dtype)
        want = port.compress(x, bits, B)
        ref = ubits(port.decompress(want, n, dtype, bits, B))
        s = torch.from_numpy(want).cuda()
        aligned = torch.full((n,), float("nan"), dtype=tdt(dtype), device="cuda")
        fz.decompress_into(s, aligned, bits, B)
        assert np.array_equal(ubits(aligned.cpu().numpy()), ref), (n, "aligned")
        buf = torch.full((n + 1,), float("nan"), dtype=tdt(dtype), device="cuda")
        off = buf[1:]
        assert off.data_ptr() % 16 != 0
        fz.decompress_into(s, off, bits, B)
        assert np.array_equal(ubits(off.cpu().numpy()), ref), (n, "one element off")
        assert np.isnan(buf[0].item())


def test_gather_out_of_range_is_zero(fz):
    x = torch.ones(1000, dtype=torch.float32, device="cuda")
    s = fz.compress(x, 16, 32)
    idx = torch.tensor([0, 999, 1000, 1 << 40], dtype=torch.int64, device="cuda")
    g = fz.gather(s, 1000, torch.float32, 16, 32, idx).cpu().numpy()
    assert g[0] == 1.0 and g[1] == 1.0 and g[2] == 0.0 and g[3] == 0.0


def test_deterministic_and_stream_ordered(fz):
    from frsz_amd.workloads import smooth_signal_torch

    x = smooth_signal_torch(1 << 22, seed=3, dtype=torch.float64)
    s1 = fz.compress(x, 32, 32)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        side.wait_stream(torch.cuda.current_stream())
        s2 = fz.compress(x, 32, 32)
    torch.cuda.current_stream().wait_stream(side)
    assert torch.equal(s1, s2)


def test_host_memory_entry_points(fz, port):
    """frsz2_*_host: host in, host out, same bytes as the device path."""
    import ctypes

    from frsz_amd import _capi

    rng = np.random.default_rng(8)
    x = mixed_values(rng, 7777, np.float64)
    bits, B = 21, 16
    size = port.compress(x, bits, B).size
    out = np.zeros(size, np.uint8)
    _capi.check(_capi.lib.frsz2_compress_host(1, bits, B, x.ctypes.data, x.size, out.ctypes.data, size))
    assert np.array_equal(out, port.compress(x, bits, B))
    y = np.empty_like(x)
    _capi.check(_capi.lib.frsz2_decompress_host(1, bits, B, out.ctypes.data, size, x.size, y.ctypes.data))
    assert np.array_equal(ubits(y), ubits(port.decompress(out, x.size, np.float64, bits, B)))
    idx = rng.integers(0, x.size, 300).astype(np.uint64)
    g = np.empty(300, np.float64)
    _capi.check(_capi.lib.frsz2_gather_host(1, bits, B, out.ctypes.data, size, x.size, idx.ctypes.data, 300,
                                            g.ctypes.data))
    assert np.array_equal(ubits(g), ubits(y[idx.astype(np.int64)]))
    del ctypes


ARBITRARY = [(np.float32, 12, 32), (np.float32, 12, 4), (np.float32, 8, 32), (np.float32, 16, 16),
             (np.float32, 21, 64), (np.float32, 24, 32), (np.float32, 32, 32), (np.float64, 21, 32),
             (np.float64, 40, 16), (np.float64, 8, 32), (np.float64, 53, 8), (np.float64, 60, 32),
             (np.float64, 12, 1024)]


@pytest.mark.parametrize("dtype,bits,B", ARBITRARY)
def test_decompress_arbitrary_streams(fz, port, dtype, bits, B):
    """Streams that no compress writes: random payload bytes and block
    exponents over the range where the reference's decode is defined (its
    shift counts stay below the width: E + bias >= -24 for float32, >= -53
    for float64, up to 40000), which includes E + bias < 0, a range no
    compress writes.  Decompress, block runs and gather equal the oracle."""
    from oracle.oracle import compressed_bytes, storage_bits

    print(f"arbitrary {np.dtype(dtype).name} b{bits} B{B}", flush=True)
    rng = np.random.default_rng(bits * 7 + B)
    n = 50_003
    U = storage_bits(bits)
    nb = -(-n // B)
    s = rng.integers(0, 256, compressed_bytes(bits, B, n), dtype=np.uint8)
    et = {8: np.int8, 16: np.int16, 32: np.int32, 64: np.int64}[U]
    info = np.iinfo(et)
    lo = max(int(info.min), -151 if dtype == np.float32 else -1076)
    hi = min(info.max, 40_000)
    E = rng.integers(lo, hi, nb, endpoint=True).astype(et)
    E[: min(nb, 4)] = np.array([lo, hi, -1, 0], dtype=et)[: min(nb, 4)]
    s[: nb * U // 8] = E.view(np.uint8)
    want = port.decompress(s, n, dtype, bits, B)
    sd = torch.from_numpy(s).cuda()
    y = fz.decompress(sd, n, tdt(dtype), bits, B).cpu().numpy()
    assert np.array_equal(ubits(y), ubits(want))
    yb = torch.empty(n, dtype=tdt(dtype), device="cuda")
    cut = (nb // 3) * B
    fz.decompress_blocks(sd, yb[:cut], n, 0, bits, B)
    fz.decompress_blocks(sd, yb[cut:], n, nb // 3, bits, B)
    assert np.array_equal(ubits(yb.cpu().numpy()), ubits(want))
    idx = rng.integers(0, n, 4096).astype(np.int64)
    g = fz.gather(sd, n, tdt(dtype), bits, B, torch.from_numpy(idx).cuda()).cpu().numpy()
    assert np.array_equal(ubits(g), ubits(want[idx]))


def generic_tile_values(dtype, bits, B, compress):
    """Values per generic tile, by the host's plan rule: TB whole blocks, a
    multiple of 16 / gcd(CB, 16) (the tile's payload then a whole number of
    16-byte chunks), as many as fit 4096 values for float32 below 32 bits
    and float64 compress up to 16 bits, 2048 otherwise."""
    from math import gcd

    from oracle.oracle import storage_bits

    U = storage_bits(bits)
    CB = -(-B * bits // U) * (U // 8)
    step = 16 // gcd(CB, 16)
    wide = bits < 32 if dtype == np.float32 else (compress and bits <= 16)
    TB = (4096 if wide else 2048) // B // step * step
    return max(TB, step) * B


@pytest.mark.parametrize("dtype,bits,B", [(np.float32, 12, 3), (np.float32, 8, 5), (np.float32, 21, 6),
                                          (np.float32, 16, 7), (np.float32, 13, 9), (np.float32, 32, 12),
                                          (np.float32, 9, 100), (np.float64, 12, 5), (np.float64, 40, 100)])
def test_generic_tile_boundaries(fz, port, dtype, bits, B):
    """Arrays that end one value before, at, and one value after a whole
    number of generic tiles (compress and decompress tiles differ in size
    for float64), and that leave a last tile of one value: stream, decode
    and gather equal the oracle."""
    rng = np.random.default_rng(bits * 101 + B)
    sizes = set()
    for compress in (True, False):
        V = generic_tile_values(dtype, bits, B, compress)
        sizes.update({V - 1, V, V + 1, 2 * V - 1, 2 * V, 2 * V + 1, 3 * V + 1})
    for n in sorted(sizes):
        x = mixed_values(rng, n, dtype)
        print(f"{np.dtype(dtype).name} bits={bits} B={B} n={n}", flush=True)
        s, y = gpu_roundtrip(fz, x, bits, B)
        want = port.compress(x, bits, B)
        assert np.array_equal(s, want), n
        assert np.array_equal(ubits(y), ubits(port.decompress(want, n, dtype, bits, B))), n
        idx = np.array([0, n - 1, n // 2, min(n - 1, B)], dtype=np.int64)
        g = fz.gather(torch.from_numpy(s).cuda(), n, tdt(dtype), bits, B, torch.from_numpy(idx).cuda())
        assert np.array_equal(ubits(g.cpu().numpy()), ubits(y[idx])), n
