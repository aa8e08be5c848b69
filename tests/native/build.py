"""Build of the native test helpers (test infrastructure):

* libdevmath_host.so: the device value arithmetic of include/frsz2_device.hpp
  compiled for the host, which tests/test_devmath.py checks against the oracle;
* libheader_host.so: the drop-in header compiled by g++ alone, whose
  host-side element decode and static helpers tests/test_header_host.py
  checks on the golden vectors;
* libfrsz2_plugin_double.so: the "frsz2" LibPressio plugin
  (frsz_amd/plugin/frsz2_plugin.cc) compiled against the LibPressio test
  double in tests/pressio_double, plus tests/native/plugin_driver.cc, for
  tests/test_plugin.py and tests/test_gpu_plugin.py;
* test_header_api: the C++ header-API program (include/frsz2.hpp) that
  tests/test_gpu_header_api.py runs, linked against frsz_amd/lib/libfrsz2_hip.so
  (build that first);
* stream_ceiling: the HBM streaming ceilings (read-only, write-only, 1:1,
  2:1, 1:2, 1:4, 4:1 read:write) the codec's kernels are measured against.

    python tests/native/build.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "frsz_amd", "lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def _run(cmd):
    res = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
        sys.stderr.write("\n".join(ln for ln in res.stdout.splitlines() if "packed-fp32-ops" not in ln) + "\n")
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")


def build_test_helpers() -> None:
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
          f"-I{INCLUDE}", "-o", os.path.join(HERE, "libdevmath_host.so"), os.path.join(HERE, "devmath_host.cpp")])
    # the drop-in header in plain host C++ (g++, no HIP headers)
    _run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-Wall", f"-I{INCLUDE}", "-o",
          os.path.join(HERE, "libheader_host.so"), os.path.join(HERE, "header_host.cpp"), f"-L{LIBDIR}",
          "-lfrsz2_hip", "-Wl,-rpath,$ORIGIN/../../frsz_amd/lib"])
    # the "frsz2" LibPressio plugin source against the LibPressio test double,
    # with the C entry points the plugin tests drive it through
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{INCLUDE}",
          f"-I{os.path.join(ROOT, 'tests', 'pressio_double')}", "-o", os.path.join(HERE, "libfrsz2_plugin_double.so"),
          os.path.join(ROOT, "frsz_amd", "plugin", "frsz2_plugin.cc"), os.path.join(HERE, "plugin_driver.cc"),
          f"-L{LIBDIR}", "-lfrsz2_hip", "-Wl,-rpath,$ORIGIN/../../frsz_amd/lib"])
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    _run([hipcc, "--offload-arch=gfx950", "-std=c++17", "-O2", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", f"-I{INCLUDE}", "-o", os.path.join(HERE, "test_header_api"),
          os.path.join(HERE, "test_header_api.cpp"), f"-L{LIBDIR}", "-lfrsz2_hip", f"-Wl,-rpath,{LIBDIR}",
          "-Wl,-rpath,$ORIGIN/../../frsz_amd/lib"])
    _run([hipcc, "--offload-arch=gfx950", "-std=c++17", "-O3", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-o", os.path.join(HERE, "stream_ceiling"), os.path.join(HERE, "stream_ceiling.hip")])


if __name__ == "__main__":
    build_test_helpers()
    print("built test helpers in", HERE)
