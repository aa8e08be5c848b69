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


DISPATCH_ERRORS = json.load(open(os.path.join(GOLDEN, "dispatch_errors.json")))
# the bits list the reference dispatch ran with when the fixture was written
# (tests/golden/make_golden.py); its block list is the plugin's
FIXTURE_BITS = (4, 8, 12, 16, 21, 32)


def test_messages_match_reference_dispatch():
    # the C ABI accepts a superset (any bits 4..W, any block 1..2^24); where it
    # rejects, code and message are the reference dispatch's
    # (include/frsz2.hpp:1300, :1326)
    for e in DISPATCH_ERRORS:
        bits, block = e["bits"], e["block"]
        if bits == 5 or block in (3, 2048):
            continue
        assert capi.lib.frsz2_check(0, bits, block) == e["rc"]
        assert capi.last_error() == e["message"]


@pytest.mark.parametrize("case", DISPATCH_ERRORS, ids=lambda e: f"b{e['bits']}-B{e['block']}")
def test_plugin_check_matches_reference_dispatch(case):
    """The Python plugin surface rejects exactly what the reference plugin
    does (src/frsz2.cc:83-134), with the dispatch's code and message.  The
    fixture's bits 5 is outside the bits list it was written with but inside
    the plugin's 4..32, so the plugin accepts it."""
    import torch

    from frsz_amd.codec import Frsz2Error, plugin_check

    bits, block = case["bits"], case["block"]
    if bits not in FIXTURE_BITS and 4 <= bits <= 32 and block in (2, 4, 8, 16, 32, 64, 128, 256, 512, 1024):
        plugin_check(torch.float32, bits, block)
        return
    with pytest.raises(Frsz2Error) as ei:
        plugin_check(torch.float32, bits, block)
    assert (ei.value.code, ei.value.message) == (case["rc"], case["message"])


def test_plugin_check_lists():
    import torch

    from frsz_amd.codec import Frsz2Error, PLUGIN_BLOCK_SIZES, plugin_check

    for B in range(0, 2050):
        ok = B in PLUGIN_BLOCK_SIZES
        try:
            plugin_check(torch.float64, 16, B)
            assert ok, B
        except Frsz2Error as e:
            assert not ok and e.code == 2 and e.message == f"Unsupported exponent block size: {B}"
    for dt, W in ((torch.float32, 32), (torch.float64, 64)):
        for b in range(-1, 70):
            try:
                plugin_check(dt, b, 32)
                assert 4 <= b <= W, b
            except Frsz2Error as e:
                assert not 4 <= b <= W and e.code == 2 and e.message == f"Unsupported number of bits: {b}"
    with pytest.raises(Frsz2Error) as ei:
        plugin_check(torch.float16, 16, 32)
    assert (ei.value.code, ei.value.message) == (1, "unsupported dtype")
    # bits go through int like the plugin's dispatch call: the low 32 bits
    plugin_check(torch.float32, (1 << 32) + 16, 32)


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


def test_batch_rejects_undersized_outputs():
    """A batch job's out_bytes must hold its output: the stream for compress,
    n values for decompress, m values for gather (validated before any
    device call)."""
    from frsz_amd.batch import BatchConfig, Job

    fake = 1 << 20
    h = ctypes.c_void_p()

    def create(job):
        arr = (Job * 1)(job)
        return capi.lib.frsz2_batch_create(arr, 1, None, ctypes.byref(h))

    from oracle.oracle import compressed_bytes

    S = compressed_bytes(16, 32, 1000)
    assert create(Job(0, 0, 16, 0, 32, 1000, fake, fake, 0, S - 1)) == capi.ECAPACITY
    assert create(Job(1, 0, 16, 0, 32, 1000, fake, fake, S, 3999)) == capi.ECAPACITY
    assert "below 1000 values" in capi.last_error()
    assert create(Job(1, 1, 16, 0, 32, 1000, fake, fake, S, 7999)) == capi.ECAPACITY
    assert create(Job(2, 0, 16, 0, 32, 1000, fake, fake, S, 4 * 99, fake, 100)) == capi.ECAPACITY
    assert "below 100 values" in capi.last_error()
    assert create(Job(2, 0, 16, 0, 32, 1000, fake, fake, S, 400, fake + 4, 100)) == capi.EARG
    assert BatchConfig  # the structure the plan functions take


def test_batch_rejects_bad_tile_shapes():
    """Per-class tile shapes are checked before any device call: values per
    thread 4, 8, 16 (2 only for float64 classes with U >= 16), workgroup size
    64 or 256, 0 = built-in."""
    from frsz_amd.batch import BATCH_CLASSES, BatchConfig, job_class

    import torch

    h = ctypes.c_void_p()
    f32_u16 = job_class(0, torch.float32, 16)
    f64_u32 = job_class(1, torch.float64, 21)
    assert (f32_u16, f64_u32) == (2, 14 + 6 + 2 * 2 + 1)
    for cls, vpt, wgs, msg in [(f32_u16, 3, 0, "values per thread 3"), (f32_u16, 2, 0, "values per thread 2"),
                               (f64_u32, 32, 0, "values per thread 32"), (BATCH_CLASSES - 1, 0, 128, "workgroup size 128")]:
        cfg = BatchConfig.from_dict({"tile_shapes": {cls: [vpt, wgs]}})
        assert capi.lib.frsz2_batch_create(None, 0, ctypes.byref(cfg), ctypes.byref(h)) == capi.EARG
        assert msg in capi.last_error() and f"class {cls}:" in capi.last_error()
    cfg = BatchConfig.from_dict({"tile_shapes": {f64_u32: [2, 64], f32_u16: [16, 256]}})
    assert cfg.as_dict()["tile_shapes"] == {f32_u16: [16, 256], f64_u32: [2, 64]}


def test_batch_job_helpers_check_tensors():
    import torch

    from frsz_amd import batch
    from frsz_amd.codec import Frsz2Error

    cpu = torch.zeros(8)
    with pytest.raises(Frsz2Error):
        batch.gather_job(torch.zeros(64, dtype=torch.uint8), 8, cpu, torch.zeros(4, dtype=torch.int64), 16, 32)
    with pytest.raises(Frsz2Error):
        batch.compress_job(cpu, torch.zeros(64, dtype=torch.uint8), 16, 32)


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
