"""MI355X-native FRSZ2 fixed-rate floating-point compressor.

The codec runs as hand-written gfx950 HIP kernels in ``lib/libfrsz2_hip.so``
behind the C ABI of ``include/frsz2_capi.h``; this package is the Python
view of that ABI over torch device tensors.  Importing it fails loudly when
the library has not been built.
"""
from ._capi import FLOAT32, FLOAT64, Frsz2Error, version
from .codec import (
    Frsz2Compressor,
    compress,
    compress_blocks,
    compress_into,
    compressed_size,
    decompress,
    decompress_blocks,
    decompress_into,
    dtype_code,
    gather,
)

__all__ = [
    "FLOAT32",
    "FLOAT64",
    "Frsz2Error",
    "Frsz2Compressor",
    "compress",
    "compress_blocks",
    "compress_into",
    "compressed_size",
    "decompress",
    "decompress_blocks",
    "decompress_into",
    "dtype_code",
    "gather",
    "version",
]
