"""ctypes access to the FRSZ2 oracle libraries.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
CPU leg of ``bench.py`` (``cpu_leg``, which yields ``cpu_baseline``) import
this module, and only as the checker (or the timed CPU baseline).  The
product package never imports it.

Two libraries are exposed with the same numpy-level API:

* ``Port`` -- ``oracle/libfrsz2_oracle.so``, the plain-C restatement of the
  reference CPU codec (oracle/frsz2_oracle.c), with OpenMP variants.
* ``Reference`` -- ``oracle/_ref/libfrsz2_ref.so``, the reference's own
  ``include/frsz2.hpp`` CPU codec compiled by ``make -C oracle ref`` where the
  reference checkout is present.  It is a git-ignored build product.

dtype codes: 0 = float32, 1 = float64.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_SO = os.path.join(HERE, "libfrsz2_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfrsz2_ref.so")

_u64 = ctypes.c_uint64
_int = ctypes.c_int
_vp = ctypes.c_void_p


def dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float32:
        return 0
    if dt == np.float64:
        return 1
    raise ValueError(f"unsupported dtype {dt}")


def np_dtype(code: int):
    return np.float32 if code == 0 else np.float64


def storage_bits(bits: int) -> int:
    return 8 if bits <= 8 else 16 if bits <= 16 else 32 if bits <= 32 else 64


def compressed_bytes(bits: int, B: int, n: int) -> int:
    """compute_compressed_memory_size_byte, include/frsz2.hpp:534-548."""
    U = storage_bits(bits)
    nb = -(-n // B)
    x = nb * (U // 8)
    r = x + 256 - x % 256
    cb = -(-(B * bits) // U) * (U // 8)
    return r + nb * cb


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class _Codec:
    def __init__(self, so_path: str, prefix: str):
        if not os.path.exists(so_path):
            raise FileNotFoundError(so_path)
        self.lib = ctypes.CDLL(so_path)
        self.prefix = prefix

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def compress(self, x: np.ndarray, bits: int, B: int, fill: int = 0) -> np.ndarray:
        x = np.ascontiguousarray(x)
        code = dtype_code(x.dtype)
        out = np.full(compressed_bytes(bits, B, x.size), fill, dtype=np.uint8)
        rc = self._compress(code, bits, B, x, out)
        if rc != 0:
            raise RuntimeError(f"{self.prefix}compress rc={rc} (bits={bits}, B={B})")
        return out

    def decompress(self, stream: np.ndarray, n: int, dtype, bits: int, B: int) -> np.ndarray:
        code = dtype_code(dtype)
        out = np.empty(n, dtype=np_dtype(code))
        stream = np.ascontiguousarray(stream, dtype=np.uint8)
        rc = self._decompress(code, bits, B, stream, n, out)
        if rc != 0:
            raise RuntimeError(f"{self.prefix}decompress rc={rc} (bits={bits}, B={B})")
        return out

    def compress_into(self, x: np.ndarray, out: np.ndarray, bits: int, B: int) -> None:
        """compress into a caller-owned contiguous uint8 buffer of at least
        compressed_bytes(bits, B, x.size) bytes (nothing allocated: for
        timing the codec alone)."""
        assert x.flags.c_contiguous and out.flags.c_contiguous and out.dtype == np.uint8
        assert out.size >= compressed_bytes(bits, B, x.size)
        rc = self._compress(dtype_code(x.dtype), bits, B, x, out)
        if rc != 0:
            raise RuntimeError(f"{self.prefix}compress rc={rc} (bits={bits}, B={B})")

    def decompress_into(self, stream: np.ndarray, out: np.ndarray, bits: int, B: int) -> None:
        """decompress out.size values into a caller-owned contiguous array."""
        assert stream.flags.c_contiguous and out.flags.c_contiguous and stream.dtype == np.uint8
        assert stream.size >= compressed_bytes(bits, B, out.size)
        rc = self._decompress(dtype_code(out.dtype), bits, B, stream, out.size, out)
        if rc != 0:
            raise RuntimeError(f"{self.prefix}decompress rc={rc} (bits={bits}, B={B})")

    def gather(self, stream: np.ndarray, n: int, dtype, bits: int, B: int, idx: np.ndarray) -> np.ndarray:
        code = dtype_code(dtype)
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        out = np.empty(idx.size, dtype=np_dtype(code))
        stream = np.ascontiguousarray(stream, dtype=np.uint8)
        fn = self._fn("gather")
        rc = fn(_int(code), _int(bits), self._blk(B), _vp(_ptr(stream)), _u64(n), _vp(_ptr(idx)),
                _u64(idx.size), _vp(_ptr(out)))
        if rc != 0:
            raise RuntimeError(f"{self.prefix}gather rc={rc}")
        return out

    def compress_block(self, stream: np.ndarray, n: int, bits: int, B: int, k: int, vals: np.ndarray) -> None:
        """compress_cpu_block(k, vals, vals.size) into `stream` (in place) of
        a view over n values (include/frsz2.hpp:1108-1179)."""
        vals = np.ascontiguousarray(vals)
        rc = self._fn("compress_block")(_int(dtype_code(vals.dtype)), _int(bits), self._blk(B), _u64(n), _u64(k),
                                        _vp(_ptr(vals)), _u64(vals.size), _vp(_ptr(stream)))
        if rc != 0:
            raise RuntimeError(f"{self.prefix}compress_block rc={rc}")

    def decompress_block(self, stream: np.ndarray, n: int, dtype, bits: int, B: int, k: int) -> np.ndarray:
        """decompress_cpu_block(k, out) of a view over n values
        (include/frsz2.hpp:1217-1240): min(B, n - k*B) values."""
        code = dtype_code(dtype)
        out = np.empty(min(B, n - k * B), dtype=np_dtype(code))
        stream = np.ascontiguousarray(stream, dtype=np.uint8)
        rc = self._fn("decompress_block")(_int(code), _int(bits), self._blk(B), _u64(n), _u64(k),
                                          _vp(_ptr(stream)), _vp(_ptr(out)))
        if rc != 0:
            raise RuntimeError(f"{self.prefix}decompress_block rc={rc}")
        return out


class Port(_Codec):
    """The C restatement (oracle/frsz2_oracle.c)."""

    def __init__(self, so_path: str = PORT_SO):
        super().__init__(so_path, "fo_")
        L = self.lib
        L.fo_compressed_bytes.restype = _u64
        L.fo_compressed_bytes.argtypes = [_int, _u64, _u64]
        for name in ("compress_mt", "decompress_mt", "compress_block", "decompress_block"):
            getattr(L, "fo_" + name).restype = _int
        L.fo_gather.restype = _int
        L.fo_max_threads.restype = _int
        L.fo_exponent_f32.restype = ctypes.c_int32
        L.fo_exponent_f32.argtypes = [ctypes.c_float]
        L.fo_significand_f32.restype = ctypes.c_uint32
        L.fo_significand_f32.argtypes = [ctypes.c_float]
        L.fo_sign_f32.restype = _int
        L.fo_sign_f32.argtypes = [ctypes.c_float]
        L.fo_exponent_f64.restype = ctypes.c_int32
        L.fo_exponent_f64.argtypes = [ctypes.c_double]
        L.fo_significand_f64.restype = ctypes.c_uint64
        L.fo_significand_f64.argtypes = [ctypes.c_double]
        L.fo_sign_f64.restype = _int
        L.fo_sign_f64.argtypes = [ctypes.c_double]
        L.fo_floating_to_fixed_f32.restype = ctypes.c_uint32
        L.fo_floating_to_fixed_f32.argtypes = [ctypes.c_float, ctypes.c_int16]
        L.fo_floating_to_fixed_f64.restype = ctypes.c_uint64
        L.fo_floating_to_fixed_f64.argtypes = [ctypes.c_double, ctypes.c_int16]
        L.fo_fixed_to_floating_f32.restype = ctypes.c_float
        L.fo_fixed_to_floating_f32.argtypes = [ctypes.c_uint32, ctypes.c_int64]
        L.fo_fixed_to_floating_f64.restype = ctypes.c_double
        L.fo_fixed_to_floating_f64.argtypes = [ctypes.c_uint64, ctypes.c_int64]
        self.threads = 1

    @staticmethod
    def _blk(B):
        return _u64(B)

    def _compress(self, code, bits, B, x, out):
        return self.lib.fo_compress_mt(_int(code), _int(bits), _u64(B), _vp(_ptr(x)), _u64(x.size),
                                       _vp(_ptr(out)), _int(self.threads))

    def _decompress(self, code, bits, B, stream, n, out):
        return self.lib.fo_decompress_mt(_int(code), _int(bits), _u64(B), _vp(_ptr(stream)), _u64(n),
                                         _vp(_ptr(out)), _int(self.threads))

    def with_threads(self, threads: int) -> "Port":
        p = Port.__new__(Port)
        p.__dict__.update(self.__dict__)
        p.threads = threads
        return p

    def max_threads(self) -> int:
        return int(self.lib.fo_max_threads())


class Reference(_Codec):
    """The reference header's CPU codec compiled by oracle/Makefile (target ref)."""

    def __init__(self, so_path: str = REF_SO):
        super().__init__(so_path, "ref_")
        L = self.lib
        for name in ("compress", "decompress", "gather", "compressed_size", "compress_block", "decompress_block"):
            getattr(L, "ref_" + name).restype = _int
        for tag, ct in (("f32a", ctypes.c_float), ("f64a", ctypes.c_double)):
            getattr(L, "ref_exponent_" + tag).restype = ctypes.c_int32
            getattr(L, "ref_exponent_" + tag).argtypes = [ct]
            getattr(L, "ref_significand_" + tag).restype = ctypes.c_uint64
            getattr(L, "ref_significand_" + tag).argtypes = [ct]
            getattr(L, "ref_sign_" + tag).restype = _int
            getattr(L, "ref_sign_" + tag).argtypes = [ct]
            getattr(L, "ref_floating_to_fixed_" + tag).restype = ctypes.c_uint64
            getattr(L, "ref_floating_to_fixed_" + tag).argtypes = [ct, ctypes.c_int16]
            getattr(L, "ref_fixed_to_floating_" + tag).restype = ct
            getattr(L, "ref_fixed_to_floating_" + tag).argtypes = [ctypes.c_uint64, ctypes.c_int16]
            getattr(L, "ref_dispatch_compress_" + tag).restype = _int

    @staticmethod
    def _blk(B):
        return _int(B)

    def _compress(self, code, bits, B, x, out):
        return self.lib.ref_compress(_int(code), _int(bits), _int(B), _vp(_ptr(x)), _u64(x.size),
                                     _vp(_ptr(out)))

    def _decompress(self, code, bits, B, stream, n, out):
        return self.lib.ref_decompress(_int(code), _int(bits), _int(B), _vp(_ptr(stream)), _u64(n),
                                       _vp(_ptr(out)))

    def compressed_size(self, dtype, bits, B, n) -> int:
        out = _u64(0)
        rc = self.lib.ref_compressed_size(_int(dtype_code(dtype)), _int(bits), _int(B), _u64(n),
                                          ctypes.byref(out))
        if rc != 0:
            raise RuntimeError(f"ref_compressed_size rc={rc}")
        return int(out.value)

    def dispatch_compress(self, x: np.ndarray, bits: int, B: int):
        """The reference's runtime dispatch (include/frsz2.hpp:1296-1347);
        returns (rc, message)."""
        x = np.ascontiguousarray(x)
        tag = "f32a" if x.dtype == np.float32 else "f64a"
        out = np.zeros(max(compressed_bytes(max(bits, 4), max(B, 1), x.size), 1), dtype=np.uint8)
        msg = ctypes.create_string_buffer(256)
        rc = getattr(self.lib, "ref_dispatch_compress_" + tag)(
            _int(bits), _int(B), _vp(_ptr(x)), _u64(x.size), _vp(_ptr(out)), msg, _int(256))
        return rc, msg.value.decode()


def reference_available() -> bool:
    return os.path.exists(REF_SO)
