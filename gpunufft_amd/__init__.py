"""MI355X-native sector-based NUFFT (gpuNUFFT-compatible).

  gpunufft_amd.capi       ctypes binding of the C ABI (include/gpunufft_hip.h)
  gpunufft_amd.gpuNUFFT   pybind11 module with the reference's NUFFTOp
  gpunufft_amd.workloads  seeded synthetic trajectories / coil maps
"""

import os

# torch ships its own copy of the HIP runtime under another file name. When
# torch is present it must be loaded first: the native libraries then bind
# to that copy (same soname) and the process holds one runtime, so device
# pointers from torch tensors are valid in the library and vice versa.
try:  # pragma: no cover - depends on the environment
    import torch  # noqa: F401
except ImportError:
    pass

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")


def native_libraries():
    """Paths of the in-tree shared objects this package loads."""
    return [os.path.join(LIB_DIR, "libgpunufft_hip.so"), os.path.join(LIB_DIR, "libgpuNUFFT.so")]
