"""The streaming-ceiling program (tests/native/stream_ceiling.hip) runs on
the device and its copies and read:write mixes store the expected bytes: the
ceilings the codec's kernels are measured against come from kernels that
move the bytes they claim."""
import json
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "native", "stream_ceiling")

pytestmark = pytest.mark.gpu


def test_stream_ceiling_patterns():
    if not os.path.exists(EXE):
        pytest.fail("tests/native/stream_ceiling is not built (python tests/native/build.py)")
    res = subprocess.run([EXE, "--mib", "64"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout[-2000:]
    rows = [json.loads(ln) for ln in res.stdout.splitlines() if ln.startswith("{")]
    patterns = {(r["read"], r["write"]) for r in rows}
    assert patterns == {(1, 0), (0, 1), (1, 1), (2, 1), (1, 2), (1, 4), (4, 1)}
    assert len(rows) == 168 and all(r["gbs"] > 0 for r in rows)


def test_random_access_ceiling():
    """--random: the gather's memory pattern without the decode; every
    variant stores the words the host expects."""
    if not os.path.exists(EXE):
        pytest.fail("tests/native/stream_ceiling is not built (python tests/native/build.py)")
    res = subprocess.run([EXE, "--random", "100000", "--values", "1000003"], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, timeout=120)
    assert res.returncode == 0, res.stdout[-2000:]
    rows = [json.loads(ln) for ln in res.stdout.splitlines() if ln.startswith("{")]
    assert {(r["loads"], r["nt_code"]) for r in rows} == {("code+exponent", 0), ("code+exponent", 1), ("code", 0),
                                                         ("code", 1), ("exponent", 0)}
    assert all(r["ok"] == 1 and r["us"] > 0 for r in rows)
