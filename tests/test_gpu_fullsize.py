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


@pytest.mark.parametrize("bits,B", [(12, 32), (32, 64), (8, 1), (12, 5000)])
def test_more_than_2p32_values(port, bits, B):
    """Maximum sizes: n = 2^32 + 4133 float32 values (17 GB in, value and byte
    offsets past 32 bits).  Blocks are independent, so block k of the stream
    equals the one-block stream of its values: checked by the oracle for the
    first blocks, those around value 2^32 and the partial last block, with the
    exponent pad, the decoded values there and gathers past 2^32.  B = 1 and
    B = 5000 take the one-thread-per-block / per-word / per-value kernels,
    whose grids then stride over more than 2^32 items."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frsz_amd
    from oracle.oracle import compressed_bytes, storage_bits

    n = (1 << 32) + 4133
    print(f"2^32+ values bits={bits} B={B}", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(bits)
    x = torch.randn(n, device="cuda", generator=g) * 1e3
    s = frsz_amd.compress(x, bits, B)
    assert s.numel() == compressed_bytes(bits, B, n)
    y = frsz_amd.decompress(s, n, torch.float32, bits, B)
    U8 = storage_bits(bits) // 8
    nb = -(-n // B)
    cb = -(-(B * bits) // (U8 * 8)) * U8
    R = s.numel() - nb * cb
    assert not torch.any(s[nb * U8:R]), "exponent pad"
    mid = (1 << 32) // B
    for k in [0, 1, mid - 2, mid - 1, mid, mid + 1, nb - 2, nb - 1]:
        v0, v1 = k * B, min(n, (k + 1) * B)
        vals = x[v0:v1].cpu().numpy()
        one = port.compress(vals, bits, B)  # the one-block stream of these values
        got_e = s[k * U8:(k + 1) * U8].cpu().numpy()
        got_p = s[R + k * cb:R + (k + 1) * cb].cpu().numpy()
        assert np.array_equal(got_e, one[:U8]), ("exponent", k)
        assert np.array_equal(got_p, one[256:256 + cb]), ("payload", k)
        want_y = port.decompress(one, v1 - v0, np.float32, bits, B)
        assert np.array_equal(y[v0:v1].cpu().numpy().view(np.uint32), want_y.view(np.uint32)), ("values", k)
    idx = torch.tensor([0, (1 << 32) - 1, 1 << 32, (1 << 32) + 1, n - 1, n, (1 << 33)], dtype=torch.int64,
                       device="cuda")
    gv = frsz_amd.gather(s, n, torch.float32, bits, B, idx).cpu().numpy().view(np.uint32)
    yv = y[idx[:5]].cpu().numpy().view(np.uint32)
    assert np.array_equal(gv[:5], yv)
    assert np.array_equal(gv[5:], np.zeros(2, np.uint32)), "indices >= n gather 0"


def test_gather_more_than_2p32_indices():
    """A gather of 2^32 + 7 indices in one call (34 GB of indices): the
    one-thread-per-index kernel strides over more work-items than one
    dispatch holds; every output (and, printed on failure, samples around
    index 2^32 and at the end) equals the decoded value it indexes."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frsz_amd

    n, bits, B = 1 << 20, 16, 32
    m = (1 << 32) + 7
    print(f"gather of {m} indices", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    x = torch.randn(n, device="cuda", generator=g)
    s = frsz_amd.compress(x, bits, B)
    y = frsz_amd.decompress(s, n, torch.float32, bits, B)
    idx = torch.randint(0, n, (m,), device="cuda", dtype=torch.int64, generator=g)
    out = frsz_amd.gather(s, n, torch.float32, bits, B, idx)
    pos = torch.tensor([0, 1, 12345, (1 << 32) - 2, (1 << 32) - 1, 1 << 32, m - 2, m - 1], device="cuda")
    got = out[pos].cpu().numpy().view(np.uint32)
    want = y[idx[pos]].cpu().numpy().view(np.uint32)
    assert np.array_equal(got, want)
    # every output: bit patterns equal to the decoded values they index
    assert torch.equal(out.view(torch.int32), y[idx].view(torch.int32))
