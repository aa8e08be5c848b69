"""Kernel-variant tuning of the plain kernels (C ABI frsz2_tune_*).

The library ships a table of measured variants per (direction, dtype,
storage width, packed); ``autotune`` re-measures it on the caller's own data
and keeps the fastest, e.g. for a GPU run at other clocks or power limits:

    chosen = frsz_amd.tune.autotune(x_gpu, bits=16, block=32)

Every variant writes the same bytes, so tuning never changes results.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi
from .codec import _flat, _stream, compressed_size, dtype_code

__all__ = ["autotune", "get_variant", "set_variant"]

_FLAG_NAMES = (("nontemporal", _capi.TUNE_NONTEMPORAL), ("staged_stores", _capi.TUNE_STAGED_STORES),
               ("linear_order", _capi.TUNE_LINEAR_ORDER), ("small_workgroups", _capi.TUNE_SMALL_WORKGROUPS))


def _describe(vpt: int, flags: int) -> dict:
    out = {"values_per_thread": vpt}
    out.update({name: bool(flags & bit) for name, bit in _FLAG_NAMES})
    return out


def get_variant(op: int, dtype: torch.dtype, bits: int) -> dict:
    v, nt, st, lin, sm = _capi.tune_get(op, dtype_code(dtype), bits)
    return {"values_per_thread": v, "nontemporal": nt, "staged_stores": st, "linear_order": lin,
            "small_workgroups": sm}


def set_variant(op: int, dtype: torch.dtype, bits: int, values_per_thread: int, nontemporal: bool = True,
                staged_stores: bool = False, linear_order: bool = False, small_workgroups: bool = False) -> None:
    _capi.tune_set(op, dtype_code(dtype), bits, values_per_thread, nontemporal, staged_stores, linear_order,
                   small_workgroups)


def autotune(x: torch.Tensor, bits: int, block: int, reps: int = 5) -> dict:
    """Tune compress on ``x`` and decompress on its stream; returns the
    chosen variants {"compress": {...}, "decompress": {...}}."""
    x = _flat(x)
    code = dtype_code(x.dtype)
    n = x.numel()
    size = compressed_size(x.dtype, bits, block, n)
    stream = torch.empty(size, dtype=torch.uint8, device=x.device)
    y = torch.empty_like(x)
    out = {}
    for op, name, src, sb, dst, db in ((0, "compress", x, n * x.element_size(), stream, size),
                                       (1, "decompress", stream, size, y, n * x.element_size())):
        vpt, flags = ctypes.c_int(), ctypes.c_int()
        _capi.check(_capi.lib.frsz2_tune_auto(op, code, bits, block, n, src.data_ptr(), sb, dst.data_ptr(), db, reps,
                                              ctypes.byref(vpt), ctypes.byref(flags), _stream()))
        out[name] = _describe(vpt.value, flags.value)
    return out
