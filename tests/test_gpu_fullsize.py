"""Parity at the workloads' full sizes against the oracle (multithreaded,
byte-identical to the serial restatement): the 10^8-value lognormal edge-case
sweep over bits {8,12,16,21,32} x B {16,32,64} with a 10^7-index random
gather, and the 1 GiB float64 field at bits 32, B 32."""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

torch = pytest.importorskip("torch")

N_EDGE = 100_000_037  # not a multiple of any block size


@pytest.fixture(scope="module")
def edge():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from frsz_amd.workloads import lognormal_edge

    x = lognormal_edge(N_EDGE, seed=3)
    return x, torch.from_numpy(x).cuda()


@pytest.fixture(scope="module")
def port_mt(port):
    return port.with_threads(0)


@pytest.mark.parametrize("bits", [8, 12, 16, 21, 32])
@pytest.mark.parametrize("B", [16, 32, 64])
def test_edge_sweep_full_size(edge, port_mt, bits, B):
    import frsz_amd

    x, xd = edge
    print(f"edge sweep n={N_EDGE} bits={bits} B={B}", flush=True)
    s = frsz_amd.compress(xd, bits, B)
    want = port_mt.compress(x, bits, B)
    got = s.cpu().numpy()
    assert got.size == want.size
    assert np.array_equal(got, want)
    y = frsz_amd.decompress(s, N_EDGE, torch.float32, bits, B).cpu().numpy()
    assert np.array_equal(y.view(np.uint32), port_mt.decompress(want, N_EDGE, np.float32, bits, B).view(np.uint32))
    if (bits, B) == (16, 32):
        rng = np.random.default_rng(7)
        idx = rng.integers(0, N_EDGE, 10_000_000, dtype=np.uint64)
        g = frsz_amd.gather(s, N_EDGE, torch.float32, bits, B, torch.from_numpy(idx.view(np.int64)).cuda())
        assert np.array_equal(g.cpu().numpy().view(np.uint32), y[idx.astype(np.int64)].view(np.uint32))


def test_field_1gib_full_size(port_mt):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frsz_amd
    from frsz_amd.workloads import gaussian_field_torch

    x = gaussian_field_torch(512, seed=2).reshape(-1).contiguous()
    s = frsz_amd.compress(x, 32, 32)
    y = frsz_amd.decompress(s, x.numel(), torch.float64, 32, 32)
    xh = x.cpu().numpy()
    want = port_mt.compress(xh, 32, 32)
    assert np.array_equal(s.cpu().numpy(), want)
    assert np.array_equal(y.cpu().numpy().view(np.uint64),
                          port_mt.decompress(want, xh.size, np.float64, 32, 32).view(np.uint64))
    # truncation bound at full size: |x - y| < 2^(E - 30) per block
    err = (x - y).abs().view(-1, 32).amax(dim=1)
    E = torch.from_numpy(want[: x.numel() // 32 * 4].view(np.int32).astype(np.int64)).cuda()
    assert bool((err < torch.exp2((E - 30).double())).all())
