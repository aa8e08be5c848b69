"""Self-describing container: header + the unchanged stream.  Bar: the bytes
after the 64-byte header are exactly the oracle's stream, the header holds
the fields written at the offsets the format fixes, the decode is the
oracle's, and the device, host and file paths agree byte for byte."""
import numpy as np
import pytest

from conftest import oracle_codec

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
def test_container_wraps_unchanged_stream(mods, port, tmp_path, dtype, bits, B, n):
    import struct

    fz, container = mods
    g = torch.Generator(device="cuda")
    g.manual_seed(n + bits)
    x = (torch.randn(n, device="cuda", generator=g, dtype=torch.float64) * 100).to(dtype)
    c = container.compress(x, bits, B)
    s, want_y = oracle_codec(port, x, bits, B)
    assert c.numel() == 64 + s.numel()
    assert torch.equal(c[64:], s)
    # the header fields at their fixed offsets (little-endian), independent of the parser
    magic, ver, dt, b, hdr, blk, nn, sb = struct.unpack("<8sIIIIQQQ", bytes(c[:48].cpu().numpy()))
    assert (magic, ver, dt, b, hdr, blk, nn, sb) == (b"FRSZ2CT\0", 1, fz.dtype_code(dtype), bits, 64, B, n, s.numel())
    assert not c[48:64].any()
    f = container.parse(c)
    assert (f["dtype"], f["bits"], f["block"], f["n"], f["stream_bytes"]) == (
        fz.dtype_code(dtype), bits, B, n, s.numel())
    y = container.decompress(c)
    assert y.dtype == dtype and y.numel() == n
    assert torch.equal(y.view(torch.uint8), want_y.view(torch.uint8))
    # file image through the host path: same bytes, same values
    path = str(tmp_path / "a.frsz2")
    container.save(path, x, bits, B)
    assert np.array_equal(np.fromfile(path, dtype=np.uint8), c.cpu().numpy())
    yh = container.load(path)
    assert np.array_equal(yh.view(np.uint8), y.cpu().numpy().view(np.uint8))


def test_container_at_unaligned_offset(mods, port):
    fz, container = mods
    x = torch.randn(100_000, device="cuda")
    c = container.compress(x, 12, 16)
    buf = torch.zeros(c.numel() + 8, dtype=torch.uint8, device="cuda")
    buf[8:] = c
    y = container.decompress(buf[8:])  # stream 8-byte aligned only
    assert torch.equal(y.view(torch.uint8), oracle_codec(port, x, 12, 16)[1].view(torch.uint8))
