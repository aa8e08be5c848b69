"""Python view of the FRSZ2 codec over device tensors.

This mirrors the reference's operator surface for the path:

* ``Frsz2Compressor`` follows the ``"frsz2"`` LibPressio plugin
  (src/frsz2.cc:25-200): options ``frsz2:bits`` and
  ``frsz2:max_exp_block_size`` (defaults 32 / 32, src/frsz2.cc:201-202),
  ``compress`` replaces the output with a new byte buffer of the stream size
  (src/frsz2.cc:152-169), ``decompress`` takes the element count and dtype
  from the output (src/frsz2.cc:171-186), and errors carry the plugin's codes
  (1 unsupported dtype, 2 unsupported bits / block size).
* ``compressed_size`` is ``compute_compressed_memory_size_byte``
  (include/frsz2.hpp:542-548); ``gather`` is ``decompress_cpu_element``
  (include/frsz2.hpp:1199-1214) over an index tensor.

Tensors are torch tensors on the current HIP device; every call is ordered
on ``torch.cuda.current_stream()``.  PyTorch is used for device memory and
streams only; all codec work is done by the HIP kernels of libfrsz2_hip.so.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi
from ._capi import FLOAT32, FLOAT64, Frsz2Error, check

__all__ = [
    "Frsz2Compressor",
    "Frsz2Error",
    "PLUGIN_BLOCK_SIZES",
    "plugin_check",
    "compressed_size",
    "compress",
    "compress_into",
    "compress_blocks",
    "decompress",
    "decompress_blocks",
    "decompress_into",
    "gather",
    "dtype_code",
]


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return FLOAT32
    if dtype == torch.float64:
        return FLOAT64
    raise Frsz2Error(_capi.EDTYPE, "unsupported dtype")


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def compressed_size(dtype: torch.dtype, bits: int, block: int, n: int) -> int:
    out = ctypes.c_uint64(0)
    check(_capi.lib.frsz2_compressed_size(dtype_code(dtype), bits, block, n, ctypes.byref(out)))
    return int(out.value)


def _flat(x: torch.Tensor) -> torch.Tensor:
    if not x.is_cuda:
        raise Frsz2Error(_capi.EARG, "tensor must be on the GPU")
    return x.contiguous().view(-1)


def compress_into(x: torch.Tensor, out: torch.Tensor, bits: int, block: int) -> torch.Tensor:
    """Compress ``x`` (float32/float64, any shape, flattened) into the uint8
    tensor ``out`` (at least ``compressed_size`` bytes).  Returns ``out``."""
    x = _flat(x)
    if out.dtype != torch.uint8 or not out.is_cuda or not out.is_contiguous():
        raise Frsz2Error(_capi.EARG, "output must be a contiguous uint8 GPU tensor")
    check(_capi.lib.frsz2_compress(dtype_code(x.dtype), bits, block, x.data_ptr(), x.numel(), out.data_ptr(),
                                   out.numel(), _stream()))
    return out


def compress(x: torch.Tensor, bits: int, block: int) -> torch.Tensor:
    x = _flat(x)
    out = torch.empty(compressed_size(x.dtype, bits, block, x.numel()), dtype=torch.uint8, device=x.device)
    return compress_into(x, out, bits, block)


def decompress_into(stream: torch.Tensor, out: torch.Tensor, bits: int, block: int) -> torch.Tensor:
    """Decompress into ``out``; like the plugin (src/frsz2.cc:174-179) the
    value count and dtype come from ``out``."""
    if not out.is_cuda or not out.is_contiguous():
        raise Frsz2Error(_capi.EARG, "output must be a contiguous GPU tensor")
    s = _flat(stream)
    check(_capi.lib.frsz2_decompress(dtype_code(out.dtype), bits, block, s.data_ptr(), s.numel(), out.numel(),
                                     out.data_ptr(), _stream()))
    return out


def decompress(stream: torch.Tensor, n: int, dtype: torch.dtype, bits: int, block: int) -> torch.Tensor:
    out = torch.empty(n, dtype=dtype, device=stream.device)
    return decompress_into(stream, out, bits, block)


def gather(stream: torch.Tensor, n: int, dtype: torch.dtype, bits: int, block: int,
           idx: torch.Tensor) -> torch.Tensor:
    """Random-access decode of ``idx`` (int64 / uint64 GPU tensor)."""
    s = _flat(stream)
    idx = _flat(idx)
    if idx.dtype not in (torch.int64, torch.uint64):
        raise Frsz2Error(_capi.EARG, "indices must be 64-bit integers")
    out = torch.empty(idx.numel(), dtype=dtype, device=s.device)
    check(_capi.lib.frsz2_gather(dtype_code(dtype), bits, block, s.data_ptr(), s.numel(), n, idx.data_ptr(),
                                 idx.numel(), out.data_ptr(), _stream()))
    return out


def compress_blocks(values: torch.Tensor, stream: torch.Tensor, n: int, first_block: int, bits: int,
                    block: int) -> torch.Tensor:
    """Compress ``values`` -- values ``[first_block*block, +values.numel())``
    of an ``n``-value array -- into their exponent slots and payload blocks
    of that array's ``stream`` (compress_cpu_block, include/frsz2.hpp:1108-1179,
    over a run of blocks).  Runs over disjoint blocks together give the bytes
    of ``compress``.  Returns ``stream``."""
    # the kernel writes into the caller's stream: a copy made to flatten a
    # non-contiguous tensor would take the bytes and drop them
    if stream.dtype != torch.uint8 or not stream.is_cuda or not stream.is_contiguous():
        raise Frsz2Error(_capi.EARG, "stream must be a contiguous uint8 GPU tensor")
    v = _flat(values)
    s = stream.view(-1)
    check(_capi.lib.frsz2_compress_blocks(dtype_code(v.dtype), bits, block, n, first_block, v.data_ptr(), v.numel(),
                                          s.data_ptr(), s.numel(), _stream()))
    return stream


def decompress_blocks(stream: torch.Tensor, out: torch.Tensor, n: int, first_block: int, bits: int,
                      block: int) -> torch.Tensor:
    """Decode values ``[first_block*block, +out.numel())`` of the ``n``-value
    array whose stream is ``stream`` into ``out`` (decompress_cpu_block,
    include/frsz2.hpp:1217-1240, over a run of blocks).  Returns ``out``."""
    if not out.is_cuda or not out.is_contiguous():
        raise Frsz2Error(_capi.EARG, "output must be a contiguous GPU tensor")
    s = _flat(stream)
    check(_capi.lib.frsz2_decompress_blocks(dtype_code(out.dtype), bits, block, n, first_block, s.data_ptr(),
                                            s.numel(), out.numel(), out.data_ptr(), _stream()))
    return out


# The plugin's supported block sizes (src/frsz2.cc:86, :98, :112, :124); its
# bit lists are 4..32 for float and 4..64 for double (:88-89, :100-103).
PLUGIN_BLOCK_SIZES = (2, 4, 8, 16, 32, 64, 128, 256, 512, 1024)


def _as_int(v: int) -> int:
    # the plugin keeps both options as uint64 and hands them to the dispatch
    # as int (src/frsz2.cc:91-92, :201-202): the low 32 bits, signed
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >= 1 << 31 else v


def plugin_check(dtype: torch.dtype, bits: int, block: int) -> None:
    """Raise the plugin's error for (dtype, bits, block): code 1 "unsupported
    dtype" (src/frsz2.cc:164, :181), then code 2 with the dispatch's message,
    bits checked before the block size (include/frsz2.hpp:1300, :1326)."""
    code = dtype_code(dtype)
    b, B = _as_int(bits), _as_int(block)
    if not 4 <= b <= (32 if code == FLOAT32 else 64):
        raise Frsz2Error(_capi.EPARAM, f"Unsupported number of bits: {b}")
    if B not in PLUGIN_BLOCK_SIZES:
        raise Frsz2Error(_capi.EPARAM, f"Unsupported exponent block size: {B}")


class Frsz2Compressor:
    """The ``"frsz2"`` compressor with the plugin's option names and defaults
    (src/frsz2.cc:28-64, :201-202) and its supported lists: compress and
    decompress reject what the plugin rejects, with its codes and messages
    (``plugin_check``).  The module-level functions accept any bits 4..W and
    any block size 1..2^24."""

    prefix = "frsz2"
    version = "0.0.2"

    def __init__(self, bits: int = 32, max_exp_block_size: int = 32):
        self.bits = int(bits)
        self.max_exp_block_size = int(max_exp_block_size)

    def get_options(self) -> dict:
        return {"frsz2:bits": self.bits, "frsz2:max_exp_block_size": self.max_exp_block_size}

    def set_options(self, options: dict) -> int:
        if "frsz2:bits" in options:
            self.bits = int(options["frsz2:bits"])
        if "frsz2:max_exp_block_size" in options:
            self.max_exp_block_size = int(options["frsz2:max_exp_block_size"])
        return 0

    def get_configuration(self) -> dict:
        return {"pressio:thread_safe": "multiple", "pressio:stability": "experimental"}

    def output_size(self, dtype: torch.dtype, n: int) -> int:
        """True stream size.  The reference plugin's output_size
        (src/frsz2.cc:136-150) omits the exponent pad and is 1..256 bytes
        short; this returns the size the codec writes."""
        return compressed_size(dtype, self.bits, self.max_exp_block_size, n)

    def compress(self, x: torch.Tensor) -> torch.Tensor:
        plugin_check(x.dtype, self.bits, self.max_exp_block_size)
        return compress(x, _as_int(self.bits), _as_int(self.max_exp_block_size))

    def decompress(self, stream: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        plugin_check(out.dtype, self.bits, self.max_exp_block_size)
        return decompress_into(stream, out, _as_int(self.bits), _as_int(self.max_exp_block_size))

    def gather(self, stream: torch.Tensor, n: int, dtype: torch.dtype, idx: torch.Tensor) -> torch.Tensor:
        plugin_check(dtype, self.bits, self.max_exp_block_size)
        return gather(stream, n, dtype, _as_int(self.bits), _as_int(self.max_exp_block_size), idx)

    def clone(self) -> "Frsz2Compressor":
        return Frsz2Compressor(self.bits, self.max_exp_block_size)
