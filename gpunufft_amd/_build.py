"""Build recipe for the native libraries (run by __graft_entry__.build()).

  lib/libgpunufft_hip.so   HIP kernels + C ABI (include/gpunufft_hip.h), gfx950
  lib/libgpuNUFFT.so       C++ gpuNUFFT:: API over the C ABI (g++)
  gpuNUFFT*.so             pybind11 module `gpuNUFFT` (NUFFTOp)

Everything is built in-tree so the shared objects travel with the source.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_SOURCES = ["precompute.hip", "gridding.hip", "prepost.hip", "fft.hip", "capi.hip", "cg.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROCM, "include"),
    "-Wall", "-Wno-unused-result",
]
HOST_SOURCES = ["gpuNUFFT_operator.cpp", "gpuNUFFT_operator_factory.cpp"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def _newer(target, sources):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(s) > t for s in sources)


def build_hip(force=True, jobs=8):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    hdrs = [os.path.join(CSRC, "hip", "gnufft_internal.hpp"),
            os.path.join(ROOT, "include", "gpunufft_hip.h")]
    objs, cmds = [], []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, "hip", src)
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(o)
        if force or _newer(o, [s] + hdrs):
            cmds.append([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIB, "libgpunufft_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + objs +
         ["-L", os.path.join(ROCM, "lib"), "-lrocfft", "-lamdhip64",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    return out


def build_host():
    """C++ API layer (namespace gpuNUFFT), linked against the C ABI library."""
    srcs = [os.path.join(CSRC, "host", s) for s in HOST_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    out = os.path.join(LIB, "libgpuNUFFT.so")
    _run(["g++"] + CXX_FLAGS + ["-shared", "-o", out] + srcs +
         ["-L", LIB, "-lgpunufft_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_python():
    src = os.path.join(CSRC, "python", "gpuNUFFT_python.cpp")
    if not os.path.exists(src):
        return None
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "gpuNUFFT" + suffix)
    inc = sysconfig.get_paths()["include"]
    _run(["g++"] + CXX_FLAGS + ["-shared", "-I", pybind11.get_include(), "-I", inc, "-o", out, src,
                                "-L", LIB, "-lgpuNUFFT", "-lgpunufft_hip",
                                "-Wl,-rpath,$ORIGIN/lib"])
    return out


def build_cpp_check():
    """tests/cpp/cpp_api_check: GPU self-check of the C++ API (run by a test)."""
    src = os.path.join(ROOT, "tests", "cpp", "cpp_api_check.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "cpp_api_check")
    _run(["g++"] + CXX_FLAGS + ["-o", out, src, "-L", LIB, "-lgpuNUFFT", "-lgpunufft_hip",
                                "-Wl,-rpath,$ORIGIN/../../gpunufft_amd/lib"])
    return out


def build_all(force=True):
    outs = [build_hip(force=force)]
    outs.append(build_host())
    outs.append(build_python())
    outs.append(build_cpp_check())
    return [o for o in outs if o]


if __name__ == "__main__":
    for o in build_all(force="--incremental" not in sys.argv):
        print(o)
