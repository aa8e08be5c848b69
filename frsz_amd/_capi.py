"""ctypes binding of the C ABI in include/frsz2_capi.h (libfrsz2_hip.so).

The library is the product: every codec call below runs a HIP kernel.  If
the library has not been built, importing this module raises immediately;
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfrsz2_hip.so")

FLOAT32 = 0
FLOAT64 = 1

OK, EDTYPE, EPARAM, ECAPACITY, EARG, EHIP, EFORMAT = range(7)


class Frsz2Error(RuntimeError):
    """A non-zero status of the C ABI; ``code`` follows the reference plugin
    (1 unsupported dtype, 2 unsupported bits / block size) plus 3..6."""

    def __init__(self, code: int, message: str):
        super().__init__(f"frsz2 error {code}: {message}")
        self.code = code
        self.message = message


def _load() -> ctypes.CDLL:
    # torch ships its own libamdhip64 (SONAME libamdhip64.so.7).  Loading it
    # first makes the dynamic linker bind this library's libamdhip64.so.7 to
    # that copy; loading ours first would leave two HIP runtimes in the
    # process, and the second one finds no device.
    import torch  # noqa: F401

    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: the frsz2 HIP library has not been built "
            "(run __graft_entry__.build()); there is no CPU fallback")
    return ctypes.CDLL(LIB_PATH)


lib = _load()

_u64 = ctypes.c_uint64
_int = ctypes.c_int
_vp = ctypes.c_void_p
_u64p = ctypes.POINTER(ctypes.c_uint64)

lib.frsz2_last_error.restype = ctypes.c_char_p
lib.frsz2_last_error.argtypes = []
lib.frsz2_version.restype = ctypes.c_char_p
lib.frsz2_version.argtypes = []
lib.frsz2_check.restype = _int
lib.frsz2_check.argtypes = [_int, _int, _u64]
lib.frsz2_compressed_size.restype = _int
lib.frsz2_compressed_size.argtypes = [_int, _int, _u64, _u64, _u64p]
lib.frsz2_compress.restype = _int
lib.frsz2_compress.argtypes = [_int, _int, _u64, _vp, _u64, _vp, _u64, _vp]
lib.frsz2_decompress.restype = _int
lib.frsz2_decompress.argtypes = [_int, _int, _u64, _vp, _u64, _u64, _vp, _vp]
lib.frsz2_gather.restype = _int
lib.frsz2_gather.argtypes = [_int, _int, _u64, _vp, _u64, _u64, _vp, _u64, _vp, _vp]
lib.frsz2_tune_set.restype = _int
lib.frsz2_tune_set.argtypes = [_int, _int, _int, _int, _int]
lib.frsz2_tune_get.restype = _int
lib.frsz2_tune_get.argtypes = [_int, _int, _int, ctypes.POINTER(_int), ctypes.POINTER(_int)]
lib.frsz2_tune_auto.restype = _int
lib.frsz2_tune_auto.argtypes = [_int, _int, _int, _u64, _u64, _vp, _u64, _vp, _u64, _int, ctypes.POINTER(_int),
                                ctypes.POINTER(_int), _vp]
lib.frsz2_compress_host.restype = _int
lib.frsz2_compress_host.argtypes = [_int, _int, _u64, _vp, _u64, _vp, _u64]
lib.frsz2_decompress_host.restype = _int
lib.frsz2_decompress_host.argtypes = [_int, _int, _u64, _vp, _u64, _u64, _vp]
lib.frsz2_gather_host.restype = _int
lib.frsz2_gather_host.argtypes = [_int, _int, _u64, _vp, _u64, _u64, _vp, _u64, _vp]
lib.frsz2_compress_blocks.restype = _int
lib.frsz2_compress_blocks.argtypes = [_int, _int, _u64, _u64, _u64, _vp, _u64, _vp, _u64, _vp]
lib.frsz2_decompress_blocks.restype = _int
lib.frsz2_decompress_blocks.argtypes = [_int, _int, _u64, _u64, _u64, _vp, _u64, _u64, _vp, _vp]
lib.frsz2_compress_blocks_host.restype = _int
lib.frsz2_compress_blocks_host.argtypes = [_int, _int, _u64, _u64, _u64, _vp, _u64, _vp, _u64]
lib.frsz2_decompress_blocks_host.restype = _int
lib.frsz2_decompress_blocks_host.argtypes = [_int, _int, _u64, _u64, _u64, _vp, _u64, _u64, _vp]


class ContainerInfo(ctypes.Structure):
    _fields_ = [("dtype", _int), ("bits", _int), ("block", _u64), ("n", _u64), ("stream_bytes", _u64),
                ("header_bytes", _u64)]


CONTAINER_HEADER_BYTES = 64
lib.frsz2_container_size.restype = _int
lib.frsz2_container_size.argtypes = [_int, _int, _u64, _u64, _u64p]
lib.frsz2_container_parse.restype = _int
lib.frsz2_container_parse.argtypes = [_vp, _u64, ctypes.POINTER(ContainerInfo)]
lib.frsz2_compress_container.restype = _int
lib.frsz2_compress_container.argtypes = [_int, _int, _u64, _vp, _u64, _vp, _u64, _vp]
lib.frsz2_decompress_container.restype = _int
lib.frsz2_decompress_container.argtypes = [ctypes.POINTER(ContainerInfo), _vp, _u64, _vp, _u64, _vp]
lib.frsz2_compress_container_host.restype = _int
lib.frsz2_compress_container_host.argtypes = [_int, _int, _u64, _vp, _u64, _vp, _u64]
lib.frsz2_decompress_container_host.restype = _int
lib.frsz2_decompress_container_host.argtypes = [_vp, _u64, _vp, _u64]


def check(rc: int) -> None:
    if rc != OK:
        raise Frsz2Error(rc, lib.frsz2_last_error().decode())


def last_error() -> str:
    return lib.frsz2_last_error().decode()


def version() -> str:
    return lib.frsz2_version().decode()


TUNE_NONTEMPORAL = 1
TUNE_STAGED_STORES = 2
TUNE_LINEAR_ORDER = 4
TUNE_SMALL_WORKGROUPS = 8


def tune_set(op: int, dtype: int, bits: int, values_per_thread: int, nontemporal: bool = True,
             staged_stores: bool = False, linear_order: bool = False, small_workgroups: bool = False) -> None:
    flags = ((TUNE_NONTEMPORAL if nontemporal else 0) | (TUNE_STAGED_STORES if staged_stores else 0)
             | (TUNE_LINEAR_ORDER if linear_order else 0) | (TUNE_SMALL_WORKGROUPS if small_workgroups else 0))
    check(lib.frsz2_tune_set(op, dtype, bits, values_per_thread, flags))


def tune_get(op: int, dtype: int, bits: int):
    """(values_per_thread, nontemporal, staged_stores, linear_order, small_workgroups) of (op, dtype, bits)."""
    v, fl = ctypes.c_int(), ctypes.c_int()
    check(lib.frsz2_tune_get(op, dtype, bits, ctypes.byref(v), ctypes.byref(fl)))
    return (v.value, bool(fl.value & TUNE_NONTEMPORAL), bool(fl.value & TUNE_STAGED_STORES),
            bool(fl.value & TUNE_LINEAR_ORDER), bool(fl.value & TUNE_SMALL_WORKGROUPS))
