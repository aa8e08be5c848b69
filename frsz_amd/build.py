"""In-tree build of the codec library, frsz_amd/lib/libfrsz2_hip.so (hipcc
for gfx950).

    python frsz_amd/build.py    # run as a script: importing the package would
                                # load the library that is about to be rebuilt

Every HIP source is compiled with the packed-fp32 instructions disabled.  The
test helpers (tests/native/build.py) and the oracle (oracle/Makefile) have
build recipes of their own.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "frsz_amd")
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")
LIB_PATH = os.path.join(LIBDIR, "libfrsz2_hip.so")
ARCH = "gfx950"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["frsz2_plain.hip", "frsz2_host.hip", "frsz2_pipeline.hip", "frsz2_container.hip", "frsz2_batch.hip", "frsz2_serve.hip"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd, cwd=ROOT):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
    lines = [ln for ln in res.stdout.splitlines() if "packed-fp32-ops" not in ln]
    if res.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")
    return "\n".join(lines)


def hipcc() -> str:
    exe = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def build_library() -> str:
    """Compile every HIP source for gfx950 and link libfrsz2_hip.so."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(ROOT, "_build", "obj")
    os.makedirs(objdir, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in srcs]

    def compile_one(pair):
        src, obj = pair
        return _run([hipcc(), *HIPCC_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)])

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB_PATH + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build_library()
    print("built", LIB_PATH)
