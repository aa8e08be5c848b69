"""The "frsz2" LibPressio plugin (frsz_amd/plugin/frsz2_plugin.cc, built
against the LibPressio test double) end to end on the GPU: compress_impl /
decompress_impl of host data through the plugin, against the golden vectors
the reference codec wrote, for every golden case whose (bits, B) the
reference plugin supports (src/frsz2.cc:83-134)."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_CASES

import plugin_double

pytestmark = pytest.mark.gpu

PLUGIN_BLOCKS = (2, 4, 8, 16, 32, 64, 128, 256, 512, 1024)
CASES = [c for c in GOLDEN_CASES if c.B in PLUGIN_BLOCKS and 4 <= c.bits <= (32 if c.dtype == np.float32 else 64)]


@pytest.fixture(scope="module")
def plugin():
    return plugin_double.Plugin(plugin_double.load())


@pytest.mark.parametrize("case", CASES, ids=repr)
def test_plugin_compress_decompress_golden(plugin, case):
    plugin.set("frsz2:bits", case.bits)
    plugin.set("frsz2:max_exp_block_size", case.B)
    rc, stream = plugin.compress(case.x)
    assert rc == 0, plugin.error()
    # the output buffer is the true stream size (the reference's output_size is short)
    assert stream.size == case.stream.size
    assert np.array_equal(stream, case.stream)
    y = np.full(case.n, np.nan, dtype=case.dtype)
    assert plugin.decompress(stream, y) == 0, plugin.error()
    assert np.array_equal(y.view(np.uint8), case.y.view(np.uint8))


def test_plugin_small_reference_case(plugin):
    """The reference's CPU-sized case through the plugin: float32 2^20
    values of the smooth signal, bits 16, B 32, against the hashes of the
    reference codec's stream and decode."""
    from frsz_amd.workloads import smooth_signal

    ref = json.load(open(os.path.join(GOLDEN, "plain_roundtrip_f32.json")))
    x = smooth_signal(ref["n"], seed=1)
    assert hashlib.sha256(x.tobytes()).hexdigest() == ref["sha256_input"]
    plugin.set("frsz2:bits", ref["bits"])
    plugin.set("frsz2:max_exp_block_size", ref["block"])
    rc, stream = plugin.compress(x)
    assert rc == 0, plugin.error()
    assert stream.size == ref["stream_bytes"]
    assert hashlib.sha256(stream.tobytes()).hexdigest() == ref["sha256_stream"]
    y = np.empty_like(x)
    assert plugin.decompress(stream, y) == 0
    assert hashlib.sha256(y.tobytes()).hexdigest() == ref["sha256_decoded"]


def test_plugin_short_stream_is_rejected(plugin):
    plugin.set("frsz2:bits", 16)
    plugin.set("frsz2:max_exp_block_size", 32)
    rc, stream = plugin.compress(np.linspace(-1, 1, 5000, dtype=np.float32))
    assert rc == 0
    y = np.empty(5000, np.float32)
    assert plugin.decompress(stream[:-1], y) == 3
    assert plugin.error()[0] == 3
