"""Self-describing container: a 64-byte header (magic, version, dtype, bits,
B, n, stream size) in front of the unchanged FRSZ2 stream, so that a stream
can be stored and read back without carrying n, dtype, bits and B out of
band (the reference stream has no header; src/frsz2.cc:152-186 passes them
separately).  ``container[64:]`` is byte for byte what ``compress`` writes.

    c = container.compress(x_gpu, bits=16, block=32)      # uint8 GPU tensor
    y = container.decompress(c)                           # dtype and n from the header
    container.save("field.frsz2", x_host_or_gpu, 16, 32)  # host path, file image
    x2 = container.load("field.frsz2")                    # numpy array
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi
from ._capi import CONTAINER_HEADER_BYTES, ContainerInfo, Frsz2Error, check
from .codec import _flat, _stream, dtype_code

__all__ = ["ContainerInfo", "container_size", "parse", "compress", "decompress", "save", "load"]


def container_size(dtype: torch.dtype, bits: int, block: int, n: int) -> int:
    out = ctypes.c_uint64(0)
    check(_capi.lib.frsz2_container_size(dtype_code(dtype), bits, block, n, ctypes.byref(out)))
    return int(out.value)


def parse(header) -> dict:
    """Header fields of a container (a uint8 tensor on any device, a numpy
    array or bytes; only the first 64 bytes are read)."""
    if isinstance(header, torch.Tensor):
        h = np.ascontiguousarray(header.view(-1)[:CONTAINER_HEADER_BYTES].cpu().numpy())
    else:
        h = np.frombuffer(bytes(header[:CONTAINER_HEADER_BYTES]), dtype=np.uint8)
    info = ContainerInfo()
    check(_capi.lib.frsz2_container_parse(h.ctypes.data, h.size, ctypes.byref(info)))
    return {k: int(getattr(info, k)) for k, _ in ContainerInfo._fields_}


def compress(x: torch.Tensor, bits: int, block: int) -> torch.Tensor:
    x = _flat(x)
    out = torch.empty(container_size(x.dtype, bits, block, x.numel()), dtype=torch.uint8, device=x.device)
    check(_capi.lib.frsz2_compress_container(dtype_code(x.dtype), bits, block, x.data_ptr(), x.numel(),
                                             out.data_ptr(), out.numel(), _stream()))
    return out


def decompress(c: torch.Tensor) -> torch.Tensor:
    """Values of a container on the GPU; reads its header back to the host
    (one small synchronous copy) for n and the dtype."""
    c = _flat(c)
    f = parse(c)
    info = ContainerInfo(**f)
    dt = torch.float32 if f["dtype"] == _capi.FLOAT32 else torch.float64
    out = torch.empty(f["n"], dtype=dt, device=c.device)
    check(_capi.lib.frsz2_decompress_container(ctypes.byref(info), c.data_ptr(), c.numel(), out.data_ptr(),
                                               out.numel(), _stream()))
    return out


def save(path: str, x, bits: int, block: int) -> int:
    """Compress ``x`` (numpy array or tensor) through the host path and write
    the container to ``path``; returns the bytes written."""
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    x = np.ascontiguousarray(x).reshape(-1)
    if x.dtype not in (np.float32, np.float64):
        raise Frsz2Error(_capi.EDTYPE, "unsupported dtype")
    tdt = torch.float32 if x.dtype == np.float32 else torch.float64
    buf = np.empty(container_size(tdt, bits, block, x.size), dtype=np.uint8)
    check(_capi.lib.frsz2_compress_container_host(dtype_code(tdt), bits, block, x.ctypes.data, x.size,
                                                  buf.ctypes.data, buf.size))
    buf.tofile(path)
    return buf.size


def load(path: str) -> np.ndarray:
    buf = np.fromfile(path, dtype=np.uint8)
    f = parse(buf)
    y = np.empty(f["n"], dtype=np.float32 if f["dtype"] == _capi.FLOAT32 else np.float64)
    check(_capi.lib.frsz2_decompress_container_host(buf.ctypes.data, buf.size, y.ctypes.data, y.size))
    return y
