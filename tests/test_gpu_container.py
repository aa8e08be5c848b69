"""Self-describing container: header + the unchanged stream.  Bar: the bytes
after the 64-byte header are exactly the plain compress output (itself
pinned to the reference), the header decodes to what was written, and the
device, host and file paths agree byte for byte."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frsz_amd
    from frsz_amd import container

    return frsz_amd, container


@pytest.mark.parametrize("dtype,bits,B,n", [(torch.float32, 16, 32, 1 << 20), (torch.float64, 21, 5, 77_777),
                                            (torch.float64, 64, 1, 1000), (torch.float32, 8, 64, 0)])
def test_container_wraps_unchanged_stream(mods, tmp_path, dtype, bits, B, n):
    fz, container = mods
    g = torch.Generator(device="cuda")
    g.manual_seed(n + bits)
    x = (torch.randn(n, device="cuda", generator=g, dtype=torch.float64) * 100).to(dtype)
    c = container.compress(x, bits, B)
    s = fz.compress(x, bits, B)
    assert c.numel() == 64 + s.numel()
    assert torch.equal(c[64:], s)
    f = container.parse(c)
    assert (f["dtype"], f["bits"], f["block"], f["n"], f["stream_bytes"]) == (
        fz.dtype_code(dtype), bits, B, n, s.numel())
    y = container.decompress(c)
    assert y.dtype == dtype and y.numel() == n
    assert torch.equal(y.view(torch.uint8), fz.decompress(s, n, dtype, bits, B).view(torch.uint8))
    # file image through the host path: same bytes, same values
    path = str(tmp_path / "a.frsz2")
    container.save(path, x, bits, B)
    assert np.array_equal(np.fromfile(path, dtype=np.uint8), c.cpu().numpy())
    yh = container.load(path)
    assert np.array_equal(yh.view(np.uint8), y.cpu().numpy().view(np.uint8))


def test_container_at_unaligned_offset(mods):
    fz, container = mods
    x = torch.randn(100_000, device="cuda")
    c = container.compress(x, 12, 16)
    buf = torch.zeros(c.numel() + 8, dtype=torch.uint8, device="cuda")
    buf[8:] = c
    y = container.decompress(buf[8:])  # stream 8-byte aligned only: generic kernels
    assert torch.equal(y, container.decompress(c))
