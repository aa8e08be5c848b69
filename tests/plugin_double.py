"""ctypes view of tests/native/libfrsz2_plugin_double.so: the "frsz2"
LibPressio plugin (frsz_amd/plugin/frsz2_plugin.cc) compiled against the
LibPressio test double, looked up in the compressor registry by name."""
import ctypes
import os

import numpy as np

from conftest import ROOT

LIB_PATH = os.path.join(ROOT, "tests", "native", "libfrsz2_plugin_double.so")

_vp, _u64, _i32, _int = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int


def load():
    import frsz_amd._capi  # noqa: F401  (torch's HIP runtime first, then the codec library)

    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "pd_create": (_vp, [ctypes.c_char_p]),
        "pd_clone": (_vp, [_vp]),
        "pd_free": (None, [_vp]),
        "pd_set_u64": (_int, [_vp, ctypes.c_char_p, _u64]),
        "pd_set_i32": (_int, [_vp, ctypes.c_char_p, _i32]),
        "pd_get_u64": (_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_u64)]),
        "pd_get_config_i64": (_int, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
        "pd_get_config_str": (_int, [_vp, ctypes.c_char_p, ctypes.c_char_p, _int]),
        "pd_get_doc_str": (_int, [_vp, ctypes.c_char_p, ctypes.c_char_p, _int]),
        "pd_version": (_int, [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_int),
                              ctypes.c_char_p, _int]),
        "pd_compress": (_int, [_vp, _int, _vp, _u64]),
        "pd_stream_bytes": (_u64, [_vp]),
        "pd_copy_stream": (_int, [_vp, _vp, _u64]),
        "pd_decompress": (_int, [_vp, _vp, _u64, _int, _vp, _u64]),
        "pd_error_code": (_int, [_vp]),
        "pd_error_msg": (ctypes.c_char_p, [_vp]),
        "register_frsz": (None, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


class Plugin:
    """One plugin instance from the registry."""

    def __init__(self, lib, handle=None):
        self.lib = lib
        self.h = handle if handle is not None else lib.pd_create(b"frsz2")
        assert self.h, "no frsz2 plugin registered"

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.pd_free(self.h)
            self.h = None

    def set(self, key, value):
        return self.lib.pd_set_u64(self.h, key.encode(), value)

    def get(self, key):
        v = _u64()
        rc = self.lib.pd_get_u64(self.h, key.encode(), ctypes.byref(v))
        return v.value if rc == 0 else None

    def config_int(self, key):
        v = ctypes.c_int64()
        return v.value if self.lib.pd_get_config_i64(self.h, key.encode(), ctypes.byref(v)) == 0 else None

    def config_str(self, key):
        buf = ctypes.create_string_buffer(512)
        return buf.value.decode() if self.lib.pd_get_config_str(self.h, key.encode(), buf, 512) == 0 else None

    def doc(self, key):
        buf = ctypes.create_string_buffer(1024)
        return buf.value.decode() if self.lib.pd_get_doc_str(self.h, key.encode(), buf, 1024) == 0 else None

    def version(self):
        a, b, c = _int(), _int(), _int()
        buf = ctypes.create_string_buffer(64)
        self.lib.pd_version(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), buf, 64)
        return (a.value, b.value, c.value), buf.value.decode()

    def clone(self):
        return Plugin(self.lib, self.lib.pd_clone(self.h))

    def error(self):
        return self.lib.pd_error_code(self.h), self.lib.pd_error_msg(self.h).decode()

    def compress(self, x, dtype_code=None):
        """rc, stream bytes (None on error) for host array x."""
        x = np.ascontiguousarray(x)
        if dtype_code is None:
            dtype_code = {np.dtype(np.float32): 0, np.dtype(np.float64): 1}.get(x.dtype, 2)
        rc = self.lib.pd_compress(self.h, dtype_code, x.ctypes.data, x.size)
        if rc:
            return rc, None
        out = np.empty(self.lib.pd_stream_bytes(self.h), np.uint8)
        assert self.lib.pd_copy_stream(self.h, out.ctypes.data, out.size) == 0
        return rc, out

    def decompress(self, stream, out):
        """rc; decodes into the host array out (its dtype and size drive the decode)."""
        stream = np.ascontiguousarray(stream)
        code = {np.dtype(np.float32): 0, np.dtype(np.float64): 1}.get(out.dtype, 2)
        return self.lib.pd_decompress(self.h, stream.ctypes.data, stream.size, code, out.ctypes.data, out.size)
