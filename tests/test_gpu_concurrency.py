"""Concurrent callers.  The reference plugin declares
``pressio:thread_safe = multiple`` (src/frsz2.cc:40): separate callers may run
at once, and the codec keeps no state between calls.  Here several threads
call the C ABI at the same time, each on a stream of its own (device entry
points) or through the host-memory pipeline, and every result must equal the
bytes of the same call made alone (which the oracle checks for one case)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = [(np.float32, 12, 32, 300_001), (np.float64, 21, 16, 200_003), (np.float32, 32, 64, 150_000),
         (np.float64, 8, 32, 250_007), (np.float32, 7, 4, 100_003), (np.float64, 53, 1024, 180_000)]


@pytest.fixture(scope="module")
def fz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frsz_amd

    return frsz_amd


def values(rng, n, dtype):
    x = (np.exp2(rng.uniform(-20, 20, n)) * rng.standard_normal(n)).astype(dtype)
    x[rng.random(n) < 0.05] = 0
    return x


def run_threads(target, count):
    errors = []

    def wrap(k):
        try:
            target(k)
        except Exception as e:  # reported on the main thread
            errors.append(f"thread {k}: {type(e).__name__}: {e}")

    threads = [threading.Thread(target=wrap, args=(k,)) for k in range(count)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors


def test_threads_on_own_streams_match_single_calls(fz, port):
    rng = np.random.default_rng(31)
    ref = []
    for dt, bits, B, n in CASES:
        xh = values(rng, n, dt)
        x = torch.from_numpy(xh).cuda()
        s = fz.compress(x, bits, B)
        y = fz.decompress(s, n, x.dtype, bits, B)
        idx = torch.from_numpy(rng.integers(0, n, 4096).astype(np.int64)).cuda()
        g = fz.gather(s, n, x.dtype, bits, B, idx)
        ref.append((x, bits, B, s, y, idx, g))
    torch.cuda.synchronize()
    x0, bits0, B0, s0 = ref[0][:4]
    assert np.array_equal(s0.cpu().numpy(), port.compress(x0.cpu().numpy(), bits0, B0))

    def worker(k):
        x, bits, B, s, y, idx, g = ref[k % len(ref)]
        st = torch.cuda.Stream()
        with torch.cuda.stream(st):
            for it in range(12):
                s2 = fz.compress(x, bits, B)
                y2 = fz.decompress(s2, x.numel(), x.dtype, bits, B)
                g2 = fz.gather(s2, x.numel(), x.dtype, bits, B, idx)
                st.synchronize()
                case = (k, it, str(x.dtype), bits, B, x.numel())
                assert torch.equal(s2, s), f"stream bytes differ {case}"
                assert torch.equal(y2.view(torch.uint8), y.view(torch.uint8)), f"decoded values differ {case}"
                assert torch.equal(g2.view(torch.uint8), g.view(torch.uint8)), f"gathered values differ {case}"

    run_threads(worker, 8)


def test_threads_through_host_pipeline_match_single_calls(fz):
    """frsz2_*_host from several threads: each call leases its own staging
    pipeline; 6e6 float64 values are two 32 MiB chunks per call."""
    from frsz_amd import _capi

    rng = np.random.default_rng(32)
    n = 6_000_000
    cases = [(1, 21, 32), (1, 32, 16), (0, 12, 64), (0, 16, 32)]
    ref = []
    for dtype, bits, B in cases:
        x = values(rng, n, np.float32 if dtype == 0 else np.float64)
        size = fz.compressed_size(torch.float32 if dtype == 0 else torch.float64, bits, B, n)
        s = np.zeros(size, np.uint8)
        _capi.check(_capi.lib.frsz2_compress_host(dtype, bits, B, x.ctypes.data, n, s.ctypes.data, size))
        y = np.empty_like(x)
        _capi.check(_capi.lib.frsz2_decompress_host(dtype, bits, B, s.ctypes.data, size, n, y.ctypes.data))
        ref.append((dtype, bits, B, x, s, y))

    def worker(k):
        dtype, bits, B, x, s, y = ref[k % len(ref)]
        for it in range(2):
            s2 = np.zeros_like(s)
            _capi.check(_capi.lib.frsz2_compress_host(dtype, bits, B, x.ctypes.data, n, s2.ctypes.data, s.size))
            y2 = np.empty_like(y)
            _capi.check(_capi.lib.frsz2_decompress_host(dtype, bits, B, s2.ctypes.data, s.size, n, y2.ctypes.data))
            case = (k, it, dtype, bits, B)
            assert np.array_equal(s2, s), f"stream bytes differ {case}"
            assert np.array_equal(y2.view(np.uint8), y.view(np.uint8)), f"decoded values differ {case}"

    run_threads(worker, 6)
