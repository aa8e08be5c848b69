"""CPU-side checks of the C-ABI library: it loads, exports every function
include/*.h declares, and rejects bad parameters with the reference's codes
and messages before touching the GPU (no kernel is launched here)."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

capi = pytest.importorskip("frsz_amd._capi")


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(frsz2_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(capi.lib, n)]
    assert not missing, missing


def test_version():
    assert "frsz2" in capi.version()


@pytest.mark.parametrize("dtype,bits,block,rc,msg", [
    (2, 16, 32, 1, "unsupported dtype"),
    (0, 3, 32, 2, "Unsupported number of bits: 3"),
    (0, 33, 32, 2, "Unsupported number of bits: 33"),
    (1, 65, 32, 2, "Unsupported number of bits: 65"),
    (1, 16, 0, 2, "Unsupported exponent block size: 0"),
    (0, 16, (1 << 24) + 1, 2, f"Unsupported exponent block size: {(1 << 24) + 1}"),
])
def test_check_rejects(dtype, bits, block, rc, msg):
    assert capi.lib.frsz2_check(dtype, bits, block) == rc
    assert capi.last_error() == msg
    # the device entry points validate before any launch
    assert capi.lib.frsz2_compress(dtype, bits, block, None, 10, None, 0, None) == rc
    assert capi.last_error() == msg


def test_messages_match_reference_dispatch():
    # the reference's runtime dispatch messages (include/frsz2.hpp:1300, :1326)
    for e in json.load(open(os.path.join(GOLDEN, "dispatch_errors.json"))):
        bits, block = e["bits"], e["block"]
        if bits == 5 or block in (3, 2048):
            continue  # supported here: any bits 4..W and any block size 1..2^24
        assert capi.lib.frsz2_check(0, bits, block) == e["rc"]
        assert capi.last_error() == e["message"]


@pytest.mark.parametrize("dtype,bits,B,n", [(0, 16, 32, 1 << 20), (1, 32, 32, 512 ** 3), (0, 21, 16, 100_000_037),
                                            (1, 9, 5, 2049), (0, 4, 1, 0)])
def test_compressed_size_matches_layout(dtype, bits, B, n):
    from oracle.oracle import compressed_bytes

    out = ctypes.c_uint64()
    assert capi.lib.frsz2_compressed_size(dtype, bits, B, n, ctypes.byref(out)) == 0
    assert out.value == compressed_bytes(bits, B, n)


def test_capacity_and_pointer_errors():
    fake = ctypes.c_void_p(1 << 20)
    # capacity below the stream size
    assert capi.lib.frsz2_compress(0, 16, 32, fake, 1000, fake, 100, None) == capi.ECAPACITY
    assert capi.lib.frsz2_decompress(0, 16, 32, fake, 100, 1000, fake, None) == capi.ECAPACITY
    # misaligned value pointer
    bad = ctypes.c_void_p((1 << 20) + 2)
    assert capi.lib.frsz2_compress(0, 16, 32, bad, 1000, fake, 1 << 20, None) == capi.EARG
    assert "misaligned" in capi.last_error()
    assert capi.lib.frsz2_compress(0, 16, 32, fake, 1000, None, 1 << 20, None) == capi.EARG


def test_python_plugin_surface_defaults():
    from frsz_amd import Frsz2Compressor

    c = Frsz2Compressor()
    assert c.get_options() == {"frsz2:bits": 32, "frsz2:max_exp_block_size": 32}
    c.set_options({"frsz2:bits": 16})
    assert c.clone().get_options()["frsz2:bits"] == 16
    assert c.get_configuration()["pressio:thread_safe"] == "multiple"


def test_block_run_errors():
    fake = ctypes.c_void_p(1 << 20)
    L = capi.lib
    # 1000 values, B 32 -> 32 blocks; a run from block 31 holds at most 8 values
    assert L.frsz2_compress_blocks(0, 16, 32, 1000, 31, fake, 9, fake, 1 << 20, None) == capi.EARG
    assert "exceed" in capi.last_error()
    assert L.frsz2_compress_blocks(0, 16, 32, 1000, 33, fake, 0, fake, 1 << 20, None) == capi.EARG
    assert L.frsz2_decompress_blocks(0, 16, 32, 1000, 0, fake, 1 << 20, 1001, fake, None) == capi.EARG
    assert L.frsz2_compress_blocks(0, 16, 32, 1000, 0, fake, 10, fake, 100, None) == capi.ECAPACITY
    assert L.frsz2_compress_blocks(0, 3, 32, 1000, 0, fake, 10, fake, 1 << 20, None) == 2
    assert L.frsz2_compress_blocks_host(0, 16, 32, 1000, 31, fake, 9, fake, 1 << 20) == capi.EARG
    assert L.frsz2_decompress_blocks_host(2, 16, 32, 1000, 0, fake, 1 << 20, 10, fake) == capi.EDTYPE
    # a run with no values is a no-op unless it ends the stream (then it writes the pad)
    assert L.frsz2_compress_blocks_host(0, 16, 32, 1000, 3, None, 0, fake, 1 << 20) == 0


def _header(dtype=0, bits=16, block=32, n=1000, stream=None, magic=b"FRSZ2CT\0", version=1, hdr=64):
    import struct

    from oracle.oracle import compressed_bytes

    if stream is None:
        stream = compressed_bytes(bits, block, n)
    return bytearray(struct.pack("<8sIIIIQQQ16x", magic, version, dtype, bits, hdr, block, n, stream))


def _parse(buf):
    info = capi.ContainerInfo()
    h = np.frombuffer(bytes(buf), dtype=np.uint8)
    rc = capi.lib.frsz2_container_parse(h.ctypes.data, h.size, ctypes.byref(info))
    return rc, info


def test_container_header_parse():
    rc, info = _parse(_header(1, 21, 5, 2049))
    assert rc == 0
    assert (info.dtype, info.bits, info.block, info.n, info.header_bytes) == (1, 21, 5, 2049, 64)
    from oracle.oracle import compressed_bytes

    assert info.stream_bytes == compressed_bytes(21, 5, 2049)
    size = ctypes.c_uint64()
    assert capi.lib.frsz2_container_size(1, 21, 5, 2049, ctypes.byref(size)) == 0
    assert size.value == 64 + info.stream_bytes


@pytest.mark.parametrize("buf,rc", [
    (_header(magic=b"FRSZ2XX\0"), capi.EFORMAT),
    (_header(version=2), capi.EFORMAT),
    (_header(hdr=128), capi.EFORMAT),
    (_header(stream=12345), capi.EFORMAT),
    (_header()[:40], capi.EFORMAT),
    (_header(bits=33), capi.EPARAM),
    (_header(dtype=3), capi.EDTYPE),
    (_header(block=0, stream=256), capi.EPARAM),
])
def test_container_header_rejects(buf, rc):
    got, _ = _parse(buf)
    assert got == rc
    assert capi.last_error()
