"""Python view of the FRSZ2 serving runtime (C ABI frsz2_server_*).

    srv = Server(hbm_budget=64 << 30)
    t = srv.submit_compress(x, bits=16, block=32)            # stream stays resident
    d = srv.submit_decompress(t, n, torch.float32, 16, 32)   # waits for t's stream
    g = srv.submit_gather(t, n, torch.float32, 16, 32, idx)
    res = srv.wait(d); y = srv.output(d, torch.float32, n)   # copy out of the pool
    srv.release(d); srv.release(g); srv.release(t)

    results = srv.replay(trace)        # open-loop trace of REQUEST_DTYPE records

Request outputs live in the server's HBM pool; ``output`` copies one into a
new torch tensor.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi
from ._capi import check
from .batch import OP_COMPRESS, OP_DECOMPRESS, OP_GATHER
from .codec import dtype_code

REQ_DISCARD = 1


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("slots", ctypes.c_int), ("max_batch", ctypes.c_int),
                ("schedule", ctypes.c_int), ("workgroups", ctypes.c_int), ("xcd_order", ctypes.c_int),
                ("stall_ms", ctypes.c_int), ("large_bytes", ctypes.c_uint64)]


class Request(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int), ("dtype", ctypes.c_int), ("bits", ctypes.c_int), ("flags", ctypes.c_int),
                ("block", ctypes.c_uint64), ("n", ctypes.c_uint64), ("values", ctypes.c_void_p),
                ("source", ctypes.c_uint64), ("idx", ctypes.c_void_p), ("m", ctypes.c_uint64),
                ("arrival", ctypes.c_double)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("reserved", ctypes.c_int), ("t_arrival", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_launch", ctypes.c_double), ("t_done", ctypes.c_double),
                ("out", ctypes.c_void_p), ("out_bytes", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64), ("launched_requests", ctypes.c_uint64),
                ("completed", ctypes.c_uint64), ("pool_used", ctypes.c_uint64), ("pool_peak", ctypes.c_uint64),
                ("max_queue", ctypes.c_uint64)]


# numpy mirrors of the request / result records (same layout as the C structs)
REQUEST_DTYPE = np.dtype([("op", "<i4"), ("dtype", "<i4"), ("bits", "<i4"), ("flags", "<i4"), ("block", "<u8"),
                          ("n", "<u8"), ("values", "<u8"), ("source", "<u8"), ("idx", "<u8"), ("m", "<u8"),
                          ("arrival", "<f8")])
RESULT_DTYPE = np.dtype([("status", "<i4"), ("reserved", "<i4"), ("t_arrival", "<f8"), ("t_admit", "<f8"),
                         ("t_launch", "<f8"), ("t_done", "<f8"), ("out", "<u8"), ("out_bytes", "<u8")])
assert REQUEST_DTYPE.itemsize == ctypes.sizeof(Request) and RESULT_DTYPE.itemsize == ctypes.sizeof(Result)

_L = _capi.lib
_L.frsz2_server_create.restype = ctypes.c_int
_L.frsz2_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(ctypes.c_void_p)]
_L.frsz2_server_submit.restype = ctypes.c_int
_L.frsz2_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
_L.frsz2_server_wait.restype = ctypes.c_int
_L.frsz2_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(Result)]
_L.frsz2_server_release.restype = ctypes.c_int
_L.frsz2_server_release.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
_L.frsz2_server_replay.restype = ctypes.c_int
_L.frsz2_server_replay.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                   ctypes.c_void_p]
_L.frsz2_server_get_stats.restype = ctypes.c_int
_L.frsz2_server_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerStats)]
_L.frsz2_server_destroy.restype = ctypes.c_int
_L.frsz2_server_destroy.argtypes = [ctypes.c_void_p]

# the HIP runtime torch loaded (same SONAME), for copies out of the pool
_hip = ctypes.CDLL("libamdhip64.so.7")
_hip.hipMemcpy.restype = ctypes.c_int
_hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
_HIP_D2D = 3


def result_dict(r: Result) -> dict:
    return {k: getattr(r, k) for k, _ in Result._fields_}


class Server:
    def __init__(self, hbm_budget: int = 64 << 30, slots: int = 4, max_batch: int = 1024, schedule: int = 0,
                 workgroups: int = 0, xcd_order: int = 0, large_bytes: int = 0, stall_ms: int = 0):
        """large_bytes: requests moving at least this many HBM bytes are
        batched only with each other, on at most slots - 1 slots at once
        (0 = 32 MiB, 2**64 - 1 = none).  stall_ms: once nothing has run for
        this long and nothing can be admitted, every queued request that has
        waited this long fails with code 3 (0 = 10000, negative = never)."""
        cfg = ServerConfig(hbm_budget, slots, max_batch, schedule, workgroups, xcd_order, stall_ms, large_bytes)
        h = ctypes.c_void_p()
        check(_L.frsz2_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def _submit(self, req: Request, keep) -> int:
        t = ctypes.c_uint64()
        check(_L.frsz2_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)))
        self._keep[t.value] = keep
        return t.value

    def submit_compress(self, x: torch.Tensor, bits: int, block: int, discard: bool = False) -> int:
        x = x.view(-1)
        req = Request(OP_COMPRESS, dtype_code(x.dtype), bits, REQ_DISCARD if discard else 0, block, x.numel(),
                      x.data_ptr(), 0, None, 0, 0.0)
        return self._submit(req, x)

    def submit_decompress(self, source: int, n: int, dtype: torch.dtype, bits: int, block: int,
                          discard: bool = False) -> int:
        req = Request(OP_DECOMPRESS, dtype_code(dtype), bits, REQ_DISCARD if discard else 0, block, n, None,
                      source, None, 0, 0.0)
        return self._submit(req, None)

    def submit_gather(self, source: int, n: int, dtype: torch.dtype, bits: int, block: int, idx: torch.Tensor,
                      discard: bool = False) -> int:
        idx = idx.view(-1)
        req = Request(OP_GATHER, dtype_code(dtype), bits, REQ_DISCARD if discard else 0, block, n, None, source,
                      idx.data_ptr(), idx.numel(), 0.0)
        return self._submit(req, idx)

    def wait(self, ticket: int) -> dict:
        r = Result()
        check(_L.frsz2_server_wait(self._h, ticket, ctypes.byref(r)))
        self._keep.pop(ticket, None)
        return result_dict(r)

    def output(self, result: dict, dtype: torch.dtype, count: int) -> torch.Tensor:
        """Copy a held output (from ``wait``) into a new tensor of ``count``
        elements of ``dtype`` (uint8 for a compressed stream)."""
        out = torch.empty(count, dtype=dtype, device="cuda")
        nbytes = count * out.element_size()
        if nbytes > result["out_bytes"]:
            raise ValueError("output is smaller than requested")
        torch.cuda.synchronize()
        if _hip.hipMemcpy(out.data_ptr(), int(result["out"]), nbytes, _HIP_D2D) != 0:
            raise RuntimeError("hipMemcpy failed")
        return out

    def release(self, ticket: int) -> None:
        check(_L.frsz2_server_release(self._h, ticket))

    def replay(self, trace: np.ndarray):
        """Open-loop replay of a REQUEST_DTYPE array; returns (results, tickets)."""
        trace = np.ascontiguousarray(trace, dtype=REQUEST_DTYPE)
        results = np.zeros(trace.size, dtype=RESULT_DTYPE)
        tickets = np.zeros(trace.size, dtype=np.uint64)
        check(_L.frsz2_server_replay(self._h, trace.ctypes.data, trace.size, results.ctypes.data,
                                     tickets.ctypes.data))
        return results, tickets

    def stats(self) -> dict:
        s = ServerStats()
        check(_L.frsz2_server_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self) -> None:
        if self._h:
            _L.frsz2_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
