"""ctypes binding of lib/libfftfpga.so (the drop-in FFTFPGA C ABI).

The signatures mirror include/fftfpga/fftfpga.h (the reference's
api/include/fftfpga/fftfpga.h) and include/fftfpga/fftfpga_hip.h.  Loading
fails loudly when the native library has not been built: there is no Python
or CPU fallback for any transform.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libfftfpga.so")


class XForm3D(ctypes.Structure):
    """fftamd_xform3d: one transform of a multi-transform plan (fftfpga_hip.h)."""

    _fields_ = [
        ("n", ctypes.c_uint * 3),
        ("reserved", ctypes.c_uint),
        ("in_offset", ctypes.c_ulonglong),
        ("out_offset", ctypes.c_ulonglong),
    ]


class FpgaT(ctypes.Structure):
    """fpga_t: per-call timings in ms plus the validity flag (fftfpga.h:32-39)."""

    _fields_ = [
        ("pcie_read_t", ctypes.c_double),
        ("pcie_write_t", ctypes.c_double),
        ("exec_t", ctypes.c_double),
        ("svm_copyin_t", ctypes.c_double),
        ("svm_copyout_t", ctypes.c_double),
        ("valid", ctypes.c_bool),
    ]

    def __repr__(self) -> str:  # pragma: no cover - debugging aid
        return (f"fpga_t(valid={self.valid}, write={self.pcie_write_t:.4f} ms, "
                f"exec={self.exec_t:.4f} ms, read={self.pcie_read_t:.4f} ms)")


# Every symbol declared in include/fftfpga/*.h, with (restype, argtypes).
_vp = ctypes.c_void_p
_u = ctypes.c_uint
_b = ctypes.c_bool
_i = ctypes.c_int
SIGNATURES = {
    # fftfpga.h
    "fpga_initialize": (_i, [ctypes.c_char_p, ctypes.c_char_p, _b]),
    "fftfpga_init_fpga": (_i, [ctypes.c_char_p, ctypes.c_char_p, _b]),
    "fpga_final": (None, []),
    "fftfpga_complex_malloc": (_vp, [ctypes.c_size_t]),
    "fftfpgaf_complex_malloc": (_vp, [ctypes.c_size_t]),
    "fftfpga_c2c_1d": (FpgaT, [_u, _vp, _vp, _b, _u]),
    "fftfpgaf_c2c_1d": (FpgaT, [_u, _vp, _vp, _b, _u]),
    "fftfpgaf_c2c_1d_svm": (FpgaT, [_u, _vp, _vp, _b, _u]),
    "fftfpgaf_c2c_2d_bram": (FpgaT, [_u, _vp, _vp, _b, _b, _u]),
    "fftfpgaf_c2c_2d_bram_svm": (FpgaT, [_u, _vp, _vp, _b, _u]),
    "fftfpgaf_c2c_2d_ddr": (FpgaT, [_u, _vp, _vp, _b]),
    "fftfpgaf_c2c_3d_bram": (FpgaT, [_u, _vp, _vp, _b, _b]),
    "fftfpgaf_c2c_3d_ddr": (FpgaT, [_u, _vp, _vp, _b]),
    "fftfpgaf_c2c_3d_ddr_batch": (FpgaT, [_u, _vp, _vp, _b, _b, _u]),
    "fftfpgaf_c2c_3d_ddr_svm": (FpgaT, [_u, _vp, _vp, _b, _b]),
    "fftfpgaf_c2c_3d_ddr_svm_batch": (FpgaT, [_u, _vp, _vp, _b, _u]),
    # fftfpga_hip.h
    "fftamd_c2c_dev": (_i, [_i, ctypes.POINTER(_u), _u, _vp, _vp, _i, _u, _vp]),
    "fftamd_c2c_dev_noalloc": (_i, [_i, ctypes.POINTER(_u), _u, _vp, _vp, _i, _u, _vp]),
    "fftamd_z2z_dev": (_i, [_i, ctypes.POINTER(_u), _u, _vp, _vp, _i, _u, _vp]),
    "fftamd_c2c_weighted_dev": (_i, [_i, ctypes.POINTER(_u), _u, _vp, _vp, _i, _vp, _vp, _vp]),
    "fftfpgaf_c2c_3d_ddr_dev": (_i, [_u, _vp, _vp, _b, _u, _vp]),
    "fftamd_multi_create": (_i, [ctypes.POINTER(_vp), _u, _vp, _vp, _vp, _i, _u]),
    "fftamd_multi_exec": (_i, [_vp, _vp]),
    "fftamd_multi_check": (_i, [_vp, _vp]),
    "fftamd_multi_tune": (_i, [_vp, _vp, ctypes.c_char_p, _i, ctypes.POINTER(ctypes.c_double)]),
    "fftamd_multi_schedule": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_u), ctypes.POINTER(_u),
                                   ctypes.POINTER(_u), ctypes.POINTER(_u)]),
    "fftamd_multi_set_schedule": (_i, [_vp, _i, _u, _u]),
    "fftamd_multi_destroy": (None, [_vp]),
    "fftamd_multi_set_spin_limit": (None, [_u]),
    "fftamd_graph_capture": (_i, [ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i, _vp]),
    "fftamd_graph_launch": (_i, [_vp, _vp]),
    "fftamd_graph_destroy": (None, [_vp]),
    "fftamd_last_error": (ctypes.c_char_p, []),
    "fftamd_plan_cache_clear": (None, []),
    "fftamd_plan_cache_size": (ctypes.c_size_t, []),
    "fftamd_profile_begin": (_i, [_i]),
    "fftamd_profile_end": (_i, [ctypes.c_char_p, _i]),
    "fftamd_describe_plan": (_i, [_i, ctypes.POINTER(_u), _u, _u, _i, _i,
                                  ctypes.POINTER(ctypes.c_longlong), _i]),
    "fftamd_describe_flow": (_i, [_u, ctypes.POINTER(_u), ctypes.POINTER(_u), ctypes.POINTER(_u), _u]),
}

_lib = None


def _share_torch_hip_runtime() -> None:
    """Map torch's bundled HIP runtime before libfftfpga.so.

    libfftfpga.so needs libamdhip64.so.7.  If PyTorch (which bundles its own
    copy, also with soname libamdhip64.so.7) is installed, loading that copy
    first makes both resolve to one HIP runtime; two runtimes in one process
    cannot both own the GPU.  Set FFTAMD_SYSTEM_HIP=1 to keep the system
    runtime (C/Python processes that never use torch).
    """
    if os.environ.get("FFTAMD_SYSTEM_HIP") == "1":
        return
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    hip = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(hip):
        ctypes.CDLL(hip, mode=ctypes.RTLD_GLOBAL)


def load(path: str | None = None) -> ctypes.CDLL:
    """Load the native library once and attach the declared signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"native library {p} is missing; build it with `make -C fft3d-fpga_amd` "
            "(there is no fallback implementation)")
    _share_torch_hip_runtime()
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    msg = load().fftamd_last_error()
    return msg.decode() if msg else ""
