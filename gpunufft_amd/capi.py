"""ctypes binding of the C ABI in include/gpunufft_hip.h (libgpunufft_hip.so).

Device buffers are torch tensors on a ROCm device (torch provides device
memory and the stream); host buffers are numpy arrays. The library is the
only compute path: if it is not built, loading fails loudly.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libgpunufft_hip.so")

OK = 0
(ERR_INVALID_ARGUMENT, ERR_OUT_OF_MEMORY, ERR_DEVICE, ERR_FFT, ERR_UNSUPPORTED,
 ERR_STATE) = 1, 2, 3, 4, 5, 6
LOOKUP_NEAREST, LOOKUP_LINEAR = 0, 1
CONVOLUTION, FFT, DEAPODIZATION, DENSITY_ESTIMATION = 0, 1, 2, 3
FFT_FUSED, FFT_ROCFFT = 0, 1
HOST, DEVICE = 0, 1

EXPORTS = [
    "gnufft_last_error", "gnufft_abi_version", "gnufft_malloc", "gnufft_free", "gnufft_memcpy",
    "gnufft_zero", "gnufft_stream_synchronize", "gnufft_mem_info", "gnufft_plan_create",
    "gnufft_plan_destroy", "gnufft_plan_get_info", "gnufft_plan_set_trajectory",
    "gnufft_plan_load_precomputed", "gnufft_plan_set_sens", "gnufft_plan_get_tables",
    "gnufft_plan_get_assigned_sectors", "gnufft_adjoint", "gnufft_forward",
    "gnufft_spread_sorted", "gnufft_interp_sorted", "gnufft_plan_set_timing",
    "gnufft_plan_get_timing", "gnufft_batch_create", "gnufft_batch_destroy",
    "gnufft_batch_adjoint", "gnufft_batch_forward", "gnufft_cg_create", "gnufft_cg_destroy",
    "gnufft_cg_solve", "gnufft_plan_set_fft", "gnufft_plan_get_fft",
    "gnufft_plan_set_workgroups_per_cu", "gnufft_check_precomputed",
    "gnufft_adjoint_noalloc",
    "gnufft_noalloc_capacity",
]

T_SPREAD, T_INTERP, T_FFT, T_ZERO, T_PREPOST = range(5)
T_NAMES = ["spread", "interp", "fft", "zero", "prepost"]


class GnufftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gpunufft error %d: %s" % (code, msg))
        self.code = code


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("img", ctypes.c_uint32 * 3),
        ("grid", ctypes.c_uint32 * 3),
        ("sectors", ctypes.c_uint32 * 3),
        ("sector_count", ctypes.c_uint32),
        ("sector_width", ctypes.c_uint32),
        ("kernel_width", ctypes.c_uint32),
        ("kernel_count", ctypes.c_uint32),
        ("sector_pad_width", ctypes.c_uint32),
        ("sector_offset", ctypes.c_uint32),
        ("n_samples", ctypes.c_uint32),
        ("n_sens_coils", ctypes.c_uint32),
        ("n_chunks", ctypes.c_uint32),
        ("lookup", ctypes.c_int32),
        ("has_dens", ctypes.c_int32),
        ("osf", ctypes.c_float),
        ("radius_sqr", ctypes.c_float),
        ("dist_multiplier", ctypes.c_float),
        ("grid_count", ctypes.c_uint64),
        ("img_count", ctypes.c_uint64),
    ]


class Timing(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * 5), ("calls", ctypes.c_uint32 * 5)]


_lib = None


def load(path=LIB_PATH):
    """Load the C ABI library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libgpunufft_hip.so is not built (%s); run __graft_entry__.build()"
                           % path)
    # torch ships its own HIP runtime; loading it first makes the library
    # bind to that one, so the process holds a single runtime and device
    # pointers from torch tensors are valid here.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    vp, u32, i32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float
    u32p, f32p = ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_float)
    L.gnufft_last_error.restype = ctypes.c_char_p
    L.gnufft_abi_version.restype = i32
    L.gnufft_malloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    L.gnufft_free.argtypes = [vp]
    L.gnufft_memcpy.argtypes = [vp, vp, ctypes.c_size_t, i32, i32, vp]
    L.gnufft_zero.argtypes = [vp, ctypes.c_size_t, vp]
    L.gnufft_stream_synchronize.argtypes = [vp]
    L.gnufft_mem_info.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    L.gnufft_plan_create.argtypes = [ctypes.POINTER(vp), u32p, u32, u32, f32, i32, vp]
    L.gnufft_plan_destroy.argtypes = [vp]
    L.gnufft_plan_get_info.argtypes = [vp, ctypes.POINTER(PlanInfo)]
    L.gnufft_plan_set_trajectory.argtypes = [vp, vp, u32, vp, i32, vp]
    L.gnufft_plan_load_precomputed.argtypes = [vp, vp, vp, vp, vp, vp, vp, u32, i32, vp]
    L.gnufft_check_precomputed.argtypes = [u32p, u32, f32, f32p, u32p, u32p, u32p, u32]
    L.gnufft_plan_set_sens.argtypes = [vp, vp, u32, i32, vp]
    L.gnufft_plan_get_tables.argtypes = [vp, u32p, u32p, u32p, f32p, f32p, f32p, f32p]
    L.gnufft_plan_get_assigned_sectors.argtypes = [vp, u32p]
    L.gnufft_adjoint.argtypes = [vp, vp, vp, u32, i32, vp]
    L.gnufft_adjoint_noalloc.argtypes = [vp, vp, vp, u32, i32, vp]
    L.gnufft_noalloc_capacity.restype = ctypes.c_size_t
    L.gnufft_forward.argtypes = [vp, vp, vp, u32, i32, vp]
    L.gnufft_spread_sorted.argtypes = [vp, vp, vp, u32, vp]
    L.gnufft_interp_sorted.argtypes = [vp, vp, vp, u32, vp]
    L.gnufft_plan_set_fft.argtypes = [vp, i32]
    L.gnufft_plan_get_fft.argtypes = [vp, ctypes.POINTER(i32)]
    L.gnufft_plan_set_workgroups_per_cu.argtypes = [vp, i32]
    L.gnufft_plan_set_timing.argtypes = [vp, i32]
    L.gnufft_plan_get_timing.argtypes = [vp, ctypes.POINTER(Timing), i32]
    vpp = ctypes.POINTER(vp)
    L.gnufft_batch_create.argtypes = [ctypes.POINTER(vp), vpp, u32, u32p, vp]
    L.gnufft_batch_destroy.argtypes = [vp]
    L.gnufft_batch_adjoint.argtypes = [vp, vpp, vpp, vp]
    L.gnufft_batch_forward.argtypes = [vp, vpp, vpp, vp]
    L.gnufft_cg_create.argtypes = [ctypes.POINTER(vp), vp, u32, f32]
    L.gnufft_cg_destroy.argtypes = [vp]
    L.gnufft_cg_solve.argtypes = [vp, vp, vp, u32, i32, vp]
    for name in EXPORTS:
        getattr(L, name).restype = getattr(L, name).restype or i32
    L.gnufft_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise GnufftError(rc, load().gnufft_last_error().decode())


def _stream_ptr(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def _ptr_loc(a):
    """(pointer, location) of a torch device tensor or a numpy host array."""
    if a is None:
        return None, HOST
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return ctypes.c_void_p(a.ctypes.data), HOST
    assert a.is_contiguous(), "device tensors must be contiguous"
    return ctypes.c_void_p(a.data_ptr()), (DEVICE if a.is_cuda else HOST)


def check_precomputed(img_dims, sector_width, osf, traj_sorted, data_indices,
                      sector_data_count, sector_centers):
    """Host-only table check (gnufft_check_precomputed); returns the status code."""
    dims = (list(img_dims) + [0, 0, 0])[:3]
    tr = np.ascontiguousarray(traj_sorted, np.float32)
    di = np.ascontiguousarray(data_indices, np.uint32)
    sdc = np.ascontiguousarray(sector_data_count, np.uint32)
    cen = np.ascontiguousarray(sector_centers, np.uint32)
    return load().gnufft_check_precomputed((ctypes.c_uint32 * 3)(*dims), sector_width, float(osf),
                                           _f32p(tr), _u32p(di), _u32p(sdc), _u32p(cen),
                                           tr.shape[-1])


def _u32p(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _f32p(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Plan:
    """One operator on the current device (owns device tables)."""

    def __init__(self, img_dims, kernel_width=3, sector_width=8, osf=2.0,
                 lookup=LOOKUP_NEAREST, stream=None):
        L = load()
        dims = (list(img_dims) + [0, 0, 0])[:3]
        self._h = ctypes.c_void_p()
        check(L.gnufft_plan_create(ctypes.byref(self._h), (ctypes.c_uint32 * 3)(*dims),
                                   kernel_width, sector_width, float(osf), lookup,
                                   _stream_ptr(stream)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().gnufft_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self):
        i = PlanInfo()
        check(load().gnufft_plan_get_info(self._h, ctypes.byref(i)))
        return i

    def set_trajectory(self, traj, dens=None, stream=None):
        tp, loc = _ptr_loc(traj)
        dp, dloc = _ptr_loc(dens)
        assert dens is None or dloc == loc
        M = traj.shape[-1]
        check(load().gnufft_plan_set_trajectory(self._h, tp, M, dp, loc, _stream_ptr(stream)))

    def load_precomputed(self, traj_sorted, data_indices, sector_data_count, sector_centers,
                         dens_sorted=None, deapo=None, stream=None):
        arrs = [traj_sorted, data_indices, sector_data_count, sector_centers, dens_sorted, deapo]
        ptrs = [_ptr_loc(a)[0] for a in arrs]
        loc = _ptr_loc(traj_sorted)[1]
        check(load().gnufft_plan_load_precomputed(self._h, *ptrs, traj_sorted.shape[-1], loc,
                                                  _stream_ptr(stream)))

    def set_sens(self, sens, n_coils, stream=None):
        sp, loc = _ptr_loc(sens)
        check(load().gnufft_plan_set_sens(self._h, sp, n_coils if sens is not None else 0, loc,
                                          _stream_ptr(stream)))

    def tables(self):
        i = self.info()
        M, S, d = i.n_samples, i.sector_count, i.dim
        out = dict(
            data_indices=np.zeros(M, np.uint32),
            sector_data_count=np.zeros(S + 1, np.uint32),
            sector_centers=np.zeros(S * d, np.uint32),
            traj_sorted=np.zeros((d, M), np.float32),
            deapo=np.zeros(int(i.img_count), np.float32),
            lut=np.zeros(i.kernel_count, np.float32),
        )
        dens = np.zeros(M, np.float32) if i.has_dens else None
        check(load().gnufft_plan_get_tables(
            self._h, _u32p(out["data_indices"]), _u32p(out["sector_data_count"]),
            _u32p(out["sector_centers"]), _f32p(out["traj_sorted"]), _f32p(dens),
            _f32p(out["deapo"]), _f32p(out["lut"])))
        out["dens_sorted"] = dens
        return out

    def assigned_sectors(self):
        a = np.zeros(self.info().n_samples, np.uint32)
        check(load().gnufft_plan_get_assigned_sectors(self._h, _u32p(a)))
        return a

    def adjoint(self, kdata, out, n_coils, stop=DEAPODIZATION, stream=None):
        check(load().gnufft_adjoint(self._h, _ptr_loc(kdata)[0], _ptr_loc(out)[0], n_coils,
                                    stop, _stream_ptr(stream)))
        return out

    def adjoint_noalloc(self, kdata, out, n_coils, stop=DEAPODIZATION, stream=None):
        """gnufft_adjoint_noalloc: the adjoint with its work memory in the
        library's fixed per-device area (no device allocation)."""
        check(load().gnufft_adjoint_noalloc(self._h, _ptr_loc(kdata)[0], _ptr_loc(out)[0],
                                            n_coils, stop, _stream_ptr(stream)))
        return out

    def forward(self, img, out, n_coils, stop=DEAPODIZATION, stream=None):
        check(load().gnufft_forward(self._h, _ptr_loc(img)[0], _ptr_loc(out)[0], n_coils,
                                    stop, _stream_ptr(stream)))
        return out

    def spread_sorted(self, data_sorted, grid, n_coils, stream=None):
        check(load().gnufft_spread_sorted(self._h, _ptr_loc(data_sorted)[0], _ptr_loc(grid)[0],
                                          n_coils, _stream_ptr(stream)))
        return grid

    def set_workgroups_per_cu(self, n):
        """Workgroups per CU the kernel-width-3 gridding kernels size their
        coil block for (0: default)."""
        check(load().gnufft_plan_set_workgroups_per_cu(self._h, int(n)))

    def set_fft(self, kind):
        """FFT_FUSED (pruned per-axis passes with the pre / post stages
        fused) or FFT_ROCFFT (rocFFT plus separate pre / post kernels)."""
        check(load().gnufft_plan_set_fft(self._h, int(kind)))

    def fft(self):
        k = ctypes.c_int()
        check(load().gnufft_plan_get_fft(self._h, ctypes.byref(k)))
        return k.value

    def set_timing(self, enable=True):
        check(load().gnufft_plan_set_timing(self._h, 1 if enable else 0))

    def timing(self, reset=True):
        """{stage: (total_ms, calls)} since the last reset (synchronises)."""
        t = Timing()
        check(load().gnufft_plan_get_timing(self._h, ctypes.byref(t), 1 if reset else 0))
        return {T_NAMES[i]: (t.ms[i], t.calls[i]) for i in range(5)}

    def interp_sorted(self, grid, data_sorted, n_coils, stream=None):
        check(load().gnufft_interp_sorted(self._h, _ptr_loc(grid)[0], _ptr_loc(data_sorted)[0],
                                          n_coils, _stream_ptr(stream)))
        return data_sorted


class Batch:
    """Independent plans run together, one launch per stage
    (gnufft_batch_*). The plans must outlive the batch."""

    def __init__(self, plans, n_coils, stream=None):
        L = load()
        self._plans = list(plans)
        n = len(self._plans)
        arr = (ctypes.c_void_p * n)(*[p.handle.value for p in self._plans])
        coils = (ctypes.c_uint32 * n)(*[int(c) for c in n_coils])
        self._h = ctypes.c_void_p()
        check(L.gnufft_batch_create(ctypes.byref(self._h), arr, n, coils, _stream_ptr(stream)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().gnufft_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ptrs(tensors):
        return (ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])

    def adjoint(self, kdata, out, stream=None):
        """kdata[i] (C_i, M_i) and out[i] (1 or C_i, *img_i): device tensors."""
        check(load().gnufft_batch_adjoint(self._h, self._ptrs(kdata), self._ptrs(out),
                                          _stream_ptr(stream)))
        return out

    def forward(self, img, out, stream=None):
        check(load().gnufft_batch_forward(self._h, self._ptrs(img), self._ptrs(out),
                                          _stream_ptr(stream)))
        return out


class CGSense:
    """Device-resident CG-SENSE solver on a plan with sensitivities
    (gnufft_cg_*): x = argmin of |A x - y|^2 + alpha |x|^2 by plain CG."""

    def __init__(self, plan, n_coils, alpha=0.0):
        self._plan = plan
        self._h = ctypes.c_void_p()
        check(load().gnufft_cg_create(ctypes.byref(self._h), plan.handle, int(n_coils),
                                      float(alpha)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().gnufft_cg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, kdata, x_out, iterations=10, use_graph=True, stream=None):
        """kdata (C, M) and x_out (1, *img) device tensors; asynchronous."""
        check(load().gnufft_cg_solve(self._h, _ptr_loc(kdata)[0], _ptr_loc(x_out)[0],
                                     int(iterations), 1 if use_graph else 0, _stream_ptr(stream)))
        return x_out
