"""Batched FRSZ2 over torch device tensors (C ABI frsz2_batch_*).

    jobs = [batch.compress_job(x, stream_buf, bits=16, block=32), ...]
    plan = batch.BatchPlan(jobs)          # descriptors uploaded, launch captured in a hipGraph
    plan.autotune()                        # optional: time the tile schedules and each class's
                                           # tile shape on these jobs
    plan.launch()                          # one graph launch on the current stream

Each job means exactly one frsz_amd.compress_into / decompress_into call
(same validation, same output bytes).  The plan keeps references to the
job tensors so they outlive it.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi
from ._capi import Frsz2Error, check
from .codec import compressed_size, dtype_code

OP_COMPRESS = 0
OP_DECOMPRESS = 1
OP_GATHER = 2


class Job(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int),
        ("dtype", ctypes.c_int),
        ("bits", ctypes.c_int),
        ("reserved", ctypes.c_int),
        ("block", ctypes.c_uint64),
        ("n", ctypes.c_uint64),
        ("in_", ctypes.c_void_p),
        ("out", ctypes.c_void_p),
        ("in_bytes", ctypes.c_uint64),
        ("out_bytes", ctypes.c_uint64),
        ("idx", ctypes.c_void_p),
        ("m", ctypes.c_uint64),
    ]


SCHEDULE_TILES = 0       # one workgroup per tile
SCHEDULE_PERSISTENT = 1  # per class, a grid of what the device holds striding over its tiles


BATCH_CLASSES = 28  # kernel classes of tiled compress / decompress jobs (FRSZ2_BATCH_CLASSES)


def job_class(op: int, dtype: torch.dtype, bits: int) -> int:
    """Kernel class of a compress or decompress job (op * 14 + slot; slot
    2u + packed for float32, 6 + 2u + packed for float64, with u the index
    of the storage width U in (8, 16, 32, 64) and packed = bits < U)."""
    U = 8 if bits <= 8 else 16 if bits <= 16 else 32 if bits <= 32 else 64
    u = (8, 16, 32, 64).index(U)
    packed = 1 if bits < U else 0
    slot = 2 * u + packed if dtype == torch.float32 else 6 + 2 * u + packed
    return op * 14 + slot


class BatchConfig(ctypes.Structure):
    _fields_ = [
        ("schedule", ctypes.c_int),
        ("workgroups", ctypes.c_int),
        ("xcd_order", ctypes.c_int),
        ("use_graph", ctypes.c_int),
        ("plain_launches", ctypes.c_int),
        ("values_per_thread", ctypes.c_ubyte * BATCH_CLASSES),
        ("workgroup_size", ctypes.c_ushort * BATCH_CLASSES),
    ]

    @classmethod
    def from_dict(cls, d: dict) -> "BatchConfig":
        """``d["tile_shapes"]`` maps a kernel class (job_class) to
        ``[values_per_thread, workgroup_size]`` (0 = the built-in value);
        classes not named run their built-in shape."""
        c = cls(d.get("schedule", SCHEDULE_TILES), d.get("workgroups", 0), d.get("xcd_order", 0),
                d.get("use_graph", 1), d.get("plain_launches", 0))
        for k, (vpt, wgs) in d.get("tile_shapes", {}).items():
            c.values_per_thread[int(k)] = vpt
            c.workgroup_size[int(k)] = wgs
        return c

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_[:5]}
        d["tile_shapes"] = {c: [self.values_per_thread[c], self.workgroup_size[c]] for c in range(BATCH_CLASSES)
                            if self.values_per_thread[c] or self.workgroup_size[c]}
        return d


_L = _capi.lib
_L.frsz2_batch_create.restype = ctypes.c_int
_L.frsz2_batch_create.argtypes = [ctypes.POINTER(Job), ctypes.c_size_t, ctypes.POINTER(BatchConfig),
                                  ctypes.POINTER(ctypes.c_void_p)]
_L.frsz2_batch_launch.restype = ctypes.c_int
_L.frsz2_batch_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
_L.frsz2_batch_autotune.restype = ctypes.c_int
_L.frsz2_batch_autotune.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(BatchConfig)]
_L.frsz2_batch_get_config.restype = ctypes.c_int
_L.frsz2_batch_get_config.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchConfig)]
_L.frsz2_batch_set_config.restype = ctypes.c_int
_L.frsz2_batch_set_config.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchConfig)]
_L.frsz2_batch_destroy.restype = ctypes.c_int
_L.frsz2_batch_destroy.argtypes = [ctypes.c_void_p]
_L.frsz2_batch_run_noalloc.restype = ctypes.c_int
_L.frsz2_batch_run_noalloc.argtypes = [ctypes.POINTER(Job), ctypes.c_size_t, ctypes.c_void_p]

NOALLOC_MAX_JOBS = 10922


def run_noalloc(jobs, stream: int | None = None) -> None:
    """Run ``jobs`` (from compress_job / decompress_job / gather_job) once, as
    a plan with the grouped shape would, without allocating device memory
    (frsz2_batch_run_noalloc).  Calls on one device that overlap in time
    share its descriptor scratch and race; calls on one stream do not."""
    arr = (Job * len(jobs))(*[j for j, _ in jobs])
    s = torch.cuda.current_stream().cuda_stream if stream is None else stream
    check(_L.frsz2_batch_run_noalloc(arr, len(jobs), s))


def _require(t: torch.Tensor, what: str, dtypes=None) -> None:
    # a job names raw device addresses: a tensor that is not on the GPU, not
    # contiguous or of another dtype would be read or written out of bounds
    if not t.is_cuda or not t.is_contiguous():
        raise Frsz2Error(_capi.EARG, f"{what} must be a contiguous GPU tensor")
    if dtypes is not None and t.dtype not in dtypes:
        raise Frsz2Error(_capi.EARG, f"{what} must be of dtype {' or '.join(str(d) for d in dtypes)}")


_VALUES = (torch.float32, torch.float64)


def compress_job(x: torch.Tensor, out: torch.Tensor, bits: int, block: int):
    _require(x, "values", _VALUES)
    _require(out, "stream", (torch.uint8,))
    x = x.view(-1)
    return (Job(OP_COMPRESS, dtype_code(x.dtype), bits, 0, block, x.numel(), x.data_ptr(), out.data_ptr(), 0,
                out.numel()), (x, out))


def decompress_job(stream: torch.Tensor, out: torch.Tensor, bits: int, block: int):
    _require(stream, "stream", (torch.uint8,))
    _require(out, "output", _VALUES)
    out_flat = out.view(-1)
    return (Job(OP_DECOMPRESS, dtype_code(out.dtype), bits, 0, block, out_flat.numel(), stream.data_ptr(),
                out_flat.data_ptr(), stream.numel(), out_flat.numel() * out.element_size()), (stream, out))


def gather_job(stream: torch.Tensor, n: int, out: torch.Tensor, idx: torch.Tensor, bits: int, block: int):
    """Gather of the ``idx.numel()`` values ``idx`` names into ``out``, which
    must hold at least that many values (the library checks the byte
    capacity as well)."""
    _require(stream, "stream", (torch.uint8,))
    _require(out, "output", _VALUES)
    _require(idx, "indices", (torch.int64, torch.uint64))
    if out.numel() < idx.numel():
        raise Frsz2Error(_capi.EARG, f"output holds {out.numel()} values, below the {idx.numel()} indices")
    idx = idx.view(-1)
    return (Job(OP_GATHER, dtype_code(out.dtype), bits, 0, block, n, stream.data_ptr(), out.data_ptr(),
                stream.numel(), out.numel() * out.element_size(), idx.data_ptr(), idx.numel()), (stream, out, idx))


class BatchPlan:
    def __init__(self, jobs, config: dict | None = None):
        self._keep = [k for _, k in jobs]
        arr = (Job * len(jobs))(*[j for j, _ in jobs])
        handle = ctypes.c_void_p()
        cfg = None if config is None else BatchConfig.from_dict(config)
        check(_L.frsz2_batch_create(arr, len(jobs), ctypes.byref(cfg) if cfg is not None else None,
                                    ctypes.byref(handle)))
        self._h = handle

    def launch(self, stream: int | None = None) -> None:
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        check(_L.frsz2_batch_launch(self._h, s))

    def autotune(self, reps: int = 3) -> dict:
        chosen = BatchConfig()
        check(_L.frsz2_batch_autotune(self._h, torch.cuda.current_stream().cuda_stream, reps, ctypes.byref(chosen)))
        return chosen.as_dict()

    @property
    def config(self) -> dict:
        c = BatchConfig()
        check(_L.frsz2_batch_get_config(self._h, ctypes.byref(c)))
        return c.as_dict()

    @config.setter
    def config(self, cfg: dict) -> None:
        cur = self.config
        cur.update(cfg)
        check(_L.frsz2_batch_set_config(self._h, ctypes.byref(BatchConfig.from_dict(cur))))

    def close(self) -> None:
        if self._h:
            _L.frsz2_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["BatchPlan", "BatchConfig", "BATCH_CLASSES", "job_class", "run_noalloc", "NOALLOC_MAX_JOBS", "compress_job", "decompress_job",
           "gather_job", "compressed_size", "OP_COMPRESS", "OP_DECOMPRESS", "OP_GATHER", "SCHEDULE_TILES",
           "SCHEDULE_PERSISTENT"]
