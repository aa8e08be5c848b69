"""The drop-in C++ header (include/frsz2.hpp) on the GPU: the program
tests/native/test_header_api drives it like the reference's GPU test drives
its header (test/test_frsz_gpu.cu:382-410); the streams it writes must equal
the golden vectors the reference codec wrote for the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN_CASES, ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "native", "test_header_api")


def test_header_api_program(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/native/test_header_api not built (run __graft_entry__.build())")
    res = subprocess.run([EXE, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
    golden = {(np.dtype(c.dtype).name, c.bits, c.B): c for c in GOLDEN_CASES if c.name == "repeat9"}
    assert len(golden) == 11
    for (tname, bits, B), case in golden.items():
        fn = tmp_path / f"{'float' if tname == 'float32' else 'double'}_b{bits}_B{B}.bin"
        got = np.fromfile(fn, dtype=np.uint8)
        assert np.array_equal(got, case.stream), fn.name
