"""The host copy pool of the pipelined host-memory path
(frsz_amd/csrc/copy_pool.hpp) under ThreadSanitizer: several pools used from
several threads at once, every copy size class, byte-exact results and no
data race reported.  CPU only."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_copy_pool_under_thread_sanitizer(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "copy_pool_stress")
    src = os.path.join(ROOT, "tests", "native", "copy_pool_stress.cpp")
    inc = os.path.join(ROOT, "frsz_amd", "csrc")
    subprocess.run([gxx, "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread", f"-I{inc}", "-o", exe, src],
                   check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    res = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300,
                         env={**os.environ, "TSAN_OPTIONS": "halt_on_error=1 exitcode=66"})
    assert "ThreadSanitizer" not in res.stdout, res.stdout[-4000:]
    assert res.returncode == 0, res.stdout[-4000:]
    assert res.stdout.strip().endswith("ok")
