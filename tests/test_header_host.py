"""The drop-in header (include/frsz2.hpp) in plain host C++, no GPU: the
program tests/native/header_host.cpp, compiled with g++, checks the view's
host-side element decode and the reference's public static helpers
(compressed_to_uint_fp / uint_fp_to_compressed / write_shift_value /
retrieve_shift_value, include/frsz2.hpp:550-644) and the static decoders
decompress_gpu_value / decompress_gpu_element_impl (:646-775, callable on the
host here) on every golden vector the reference codec wrote."""
import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN_CASES, ROOT

LIB = os.path.join(ROOT, "tests", "native", "libheader_host.so")


@pytest.fixture(scope="module")
def hh():
    pytest.importorskip("frsz_amd._capi")  # loads torch's HIP runtime first, then the codec library
    if not os.path.exists(LIB):
        pytest.fail("tests/native/libheader_host.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    lib.hh_check.restype = ctypes.c_int
    lib.hh_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                             ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    lib.hh_size_f32_b16_B32.restype = ctypes.c_uint64
    lib.hh_size_f32_b16_B32.argtypes = [ctypes.c_uint64]
    return lib


WHAT = {1: "decompress_cpu_element", 2: "retrieve_shift_value overloads disagree",
        3: "retrieve_shift_value code decodes to another value", 4: "write_shift_value repack differs",
        5: "decompress_gpu_value / decompress_gpu_element_impl differ from the decode"}


@pytest.mark.parametrize("case", GOLDEN_CASES, ids=repr)
def test_header_host_helpers_on_golden(hh, case):
    s = np.ascontiguousarray(case.stream)
    y = np.ascontiguousarray(case.y)
    where = ctypes.c_uint64(0)
    rc = hh.hh_check(0 if case.dtype == np.float32 else 1, case.bits, case.B, s.ctypes.data, case.n, y.ctypes.data,
                     ctypes.byref(where))
    assert rc != -1, f"(bits {case.bits}, B {case.B}) not instantiated in header_host.cpp"
    assert rc == 0, f"{WHAT.get(rc, rc)} at index {where.value}"


def test_header_host_size(hh):
    # compute_compressed_memory_size_byte at the small reference case (2^20 float32, bits 16, B 32)
    assert hh.hh_size_f32_b16_B32(1 << 20) == 2_162_944
