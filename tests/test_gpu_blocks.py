"""Runs of exponent blocks (frsz2_compress_blocks / frsz2_decompress_blocks
and their host-memory variants), the per-block methods of the reference
view (compress_cpu_block / decompress_cpu_block, include/frsz2.hpp:1108-1240)
and the pipelined host-memory path built on them.

Bar: bit-exact.  Runs over a partition of the blocks together give the bytes
of the whole-array compress (itself pinned to the reference by the golden
vectors); a partial block inside the array is checked against the oracle's
compress_cpu_block, which the compiled reference pins on the CPU."""
import threading

import numpy as np
import pytest

from oracle.oracle import compressed_bytes, storage_bits

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


def values(rng, n, dtype):
    x = (rng.standard_normal(n) * np.exp2(rng.integers(-30, 30, n))).astype(dtype)
    x[rng.random(n) < 0.05] = 0
    return x


def partition(rng, nb, parts):
    cuts = np.sort(rng.choice(np.arange(1, nb), size=min(parts - 1, nb - 1), replace=False)) if nb > 1 else []
    edges = [0, *[int(c) for c in cuts], nb]
    return list(zip(edges[:-1], edges[1:]))


CASES = [
    (np.float32, 16, 32, 100_003),
    (np.float32, 12, 16, 65_537),
    (np.float32, 32, 64, 70_001),
    (np.float32, 8, 5, 33_333),
    (np.float64, 32, 32, 50_001),
    (np.float64, 21, 8, 40_003),
    (np.float64, 40, 7, 20_011),
    (np.float64, 64, 1, 9_999),
    (np.float64, 9, 1024, 123_457),
]


@pytest.mark.parametrize("dtype,bits,B,n", CASES)
def test_runs_equal_whole_stream(fz, dtype, bits, B, n):
    print(f"runs {np.dtype(dtype).name} b{bits} B{B} n{n}", flush=True)
    rng = np.random.default_rng(bits * 1000 + B)
    x = values(rng, n, dtype)
    xd = torch.from_numpy(x).cuda()
    whole = fz.compress(xd, bits, B)
    yd = fz.decompress(whole, n, tdt(dtype), bits, B)
    nb = -(-n // B)
    s = torch.full((whole.numel(),), 0xAB, dtype=torch.uint8, device="cuda")
    y = torch.full((n,), -7.0, dtype=tdt(dtype), device="cuda")
    runs = partition(rng, nb, 7)
    for b0, b1 in rng.permutation(np.array(runs, dtype=np.int64)):
        b0, b1 = int(b0), int(b1)
        v0, v1 = b0 * B, min(n, b1 * B)
        fz.compress_blocks(xd[v0:v1], s, n, b0, bits, B)
    assert torch.equal(s, whole)
    for b0, b1 in runs:
        v0, v1 = b0 * B, min(n, b1 * B)
        fz.decompress_blocks(whole, y[v0:v1], n, b0, bits, B)
    assert torch.equal(y.view(torch.uint8), yd.view(torch.uint8))


@pytest.mark.parametrize("dtype,bits,B", [(np.float32, 16, 32), (np.float32, 12, 16), (np.float64, 21, 9),
                                          (np.float64, 32, 32), (np.float32, 7, 64)])
def test_partial_block_inside_array(fz, port, dtype, bits, B):
    """compress_cpu_block(k, data, exp_block_elements < B) on an existing
    stream: exponent and codes as the reference writes them; the words of the
    block past the codes, which the reference leaves as they were, are zero."""
    rng = np.random.default_rng(21 + bits)
    n = 40 * B + 3
    x = values(rng, n, dtype)
    base = port.compress(x, bits, B)
    for k, ln in ((0, 1), (5, B // 2 + 1), (17, B - 1), (39, 1)):
        newv = values(rng, ln, dtype)
        want = base.copy()
        port.compress_block(want, n, bits, B, k, newv)
        U = storage_bits(bits)
        ub = U // 8
        cb = -(-(B * bits) // U) * ub
        R = compressed_bytes(bits, B, n) - (-(-n // B)) * cb
        touched = -(-(ln * bits) // U)
        want[R + k * cb + touched * ub: R + (k + 1) * cb] = 0
        got = torch.from_numpy(base.copy()).cuda()
        fz.compress_blocks(torch.from_numpy(newv).cuda(), got, n, k, bits, B)
        assert np.array_equal(got.cpu().numpy(), want), (k, ln)
        out = torch.empty(min(B, n - k * B), dtype=tdt(dtype), device="cuda")
        fz.decompress_blocks(got, out, n, k, bits, B)
        ref = port.decompress_block(want, n, dtype, bits, B, k)
        assert np.array_equal(ubits(out.cpu().numpy()), ubits(ref))


def test_run_that_ends_the_stream_writes_the_pad(fz):
    x = torch.randn(1000, device="cuda")
    whole = fz.compress(x, 16, 32)
    s = torch.full_like(whole, 0xCD)
    fz.compress_blocks(x[: 31 * 32], s, 1000, 0, 16, 32)  # blocks 0..30: pad untouched
    assert torch.all(s[31 * 2: 256] == 0xCD)
    fz.compress_blocks(x[31 * 32:], s, 1000, 31, 16, 32)  # last block: pad zeroed
    assert torch.equal(s, whole)


@pytest.mark.parametrize("dtype,bits,B,n", [(np.float32, 16, 32, 10_007), (np.float64, 21, 5, 7_777),
                                            (np.float64, 64, 128, 3_001)])
def test_host_block_runs(fz, port, dtype, bits, B, n):
    from frsz_amd import _capi

    rng = np.random.default_rng(3)
    x = values(rng, n, dtype)
    want = port.compress(x, bits, B)
    code = 0 if dtype == np.float32 else 1
    vs = np.dtype(dtype).itemsize
    s = np.full(want.size, 0x5A, np.uint8)
    runs = partition(rng, -(-n // B), 5)
    for b0, b1 in runs[::-1]:
        v0, v1 = b0 * B, min(n, b1 * B)
        _capi.check(_capi.lib.frsz2_compress_blocks_host(code, bits, B, n, b0, x.ctypes.data + v0 * vs, v1 - v0,
                                                         s.ctypes.data, s.size))
    assert np.array_equal(s, want)
    y = np.full(n, -1, dtype)
    for b0, b1 in runs:
        v0, v1 = b0 * B, min(n, b1 * B)
        _capi.check(_capi.lib.frsz2_decompress_blocks_host(code, bits, B, n, b0, s.ctypes.data, s.size, v1 - v0,
                                                           y.ctypes.data + v0 * vs))
    assert np.array_equal(ubits(y), ubits(port.decompress(want, n, dtype, bits, B)))


@pytest.mark.parametrize("dtype,bits,B,n", [(np.float32, 16, 32, 30_000_001), (np.float64, 21, 5, 10_000_003),
                                            (np.float32, 12, 1 << 20, 40_000_000)])
def test_pipelined_host_path_many_chunks(fz, dtype, bits, B, n):
    """Several 32 MiB chunks through the three pipeline slots: the host
    path's stream and values equal the device path's."""
    from frsz_amd import _capi

    print(f"pipeline {np.dtype(dtype).name} b{bits} B{B} n{n}", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(n)
    xd = (torch.randn(n, device="cuda", generator=g, dtype=torch.float64) * 1e3).to(tdt(dtype))
    sd = fz.compress(xd, bits, B)
    yd = fz.decompress(sd, n, tdt(dtype), bits, B)
    x = xd.cpu().numpy()
    code = 0 if dtype == np.float32 else 1
    s = np.empty(sd.numel(), np.uint8)
    _capi.check(_capi.lib.frsz2_compress_host(code, bits, B, x.ctypes.data, n, s.ctypes.data, s.size))
    assert np.array_equal(s, sd.cpu().numpy())
    y = np.empty(n, dtype)
    _capi.check(_capi.lib.frsz2_decompress_host(code, bits, B, s.ctypes.data, s.size, n, y.ctypes.data))
    assert np.array_equal(ubits(y), ubits(yd.cpu().numpy()))


def test_host_path_concurrent_callers(fz, port):
    """thread_safe = multiple (src/frsz2.cc:40): concurrent host calls each
    hold their own pipeline."""
    from frsz_amd import _capi

    rng = np.random.default_rng(9)
    xs = [values(rng, 3_000_000 + 17 * i, np.float32) for i in range(4)]
    outs = [np.empty(compressed_bytes(16, 32, x.size), np.uint8) for x in xs]
    rcs = [None] * 4

    def work(i):
        rcs[i] = _capi.lib.frsz2_compress_host(0, 16, 32, xs[i].ctypes.data, xs[i].size, outs[i].ctypes.data,
                                               outs[i].size)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert rcs == [0, 0, 0, 0]
    for x, o in zip(xs, outs):
        assert np.array_equal(o, port.with_threads(8).compress(x, 16, 32))


@pytest.mark.parametrize("dtype,bits,B", [(np.float32, 16, 32), (np.float64, 21, 5), (np.float32, 12, 1 << 22)])
def test_host_path_tiny_and_empty(fz, port, dtype, bits, B):
    """n = 0 (the stream is the zeroed pad only), n = 1 and a few values,
    including a block size far above n (one short chunk)."""
    from frsz_amd import _capi

    code = 0 if dtype == np.float32 else 1
    rng = np.random.default_rng(12)
    for n in (0, 1, 3, 257):
        x = values(rng, n, dtype)
        want = port.compress(x, bits, B)
        s = np.full(want.size, 0x99, np.uint8)
        _capi.check(_capi.lib.frsz2_compress_host(code, bits, B, x.ctypes.data if n else None, n, s.ctypes.data,
                                                  s.size))
        assert np.array_equal(s, want), n
        y = np.full(n, -3, dtype)
        _capi.check(_capi.lib.frsz2_decompress_host(code, bits, B, s.ctypes.data, s.size, n,
                                                    y.ctypes.data if n else None))
        assert np.array_equal(ubits(y), ubits(port.decompress(want, n, dtype, bits, B))), n

