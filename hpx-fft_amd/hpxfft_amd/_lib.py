"""ctypes binding of include/hpxfft_hip.h (libhpxfft_hip.so).

The shared library is the product's only compute path.  Loading fails
loudly (ImportError-like RuntimeError) when it has not been built; there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libhpxfft_hip.so")

# status codes (hpxfft_status)
OK, INVALID_ARG, UNSUPPORTED_SIZE, NO_DEVICE, OUT_OF_MEMORY, HIP_ERROR, IO_ERROR = range(7)
STATUS_NAMES = {
    OK: "OK", INVALID_ARG: "INVALID_ARG", UNSUPPORTED_SIZE: "UNSUPPORTED_SIZE", NO_DEVICE: "NO_DEVICE",
    OUT_OF_MEMORY: "OUT_OF_MEMORY", HIP_ERROR: "HIP_ERROR", IO_ERROR: "IO_ERROR",
}
# precisions (hpxfft_precision)
F64, F32, F16_STORAGE, BF16_STORAGE = range(4)
# efforts (hpxfft_effort)
ESTIMATE, MEASURE, PATIENT, EXHAUSTIVE = range(4)

# every entry point declared in include/hpxfft_hip.h
EXPORTS = (
    "hpxfft_plan_flag_from_string", "hpxfft_plan_r2c_2d", "hpxfft_plan_r2c_2d_many", "hpxfft_exec_r2c_2d",
    "hpxfft_exec_r2c_2d_scaled", "hpxfft_exec_r2c_2d_host", "hpxfft_exec_r2c_2d_adjoint", "hpxfft_exec_c2r_2d", "hpxfft_get_measurement",
    "hpxfft_write_plan", "hpxfft_plan_describe", "hpxfft_plan_destroy", "hpxfft_plan_r2c_2d_batched",
    "hpxfft_exec_r2c_2d_batched", "hpxfft_plan_batched_set_config", "hpxfft_plan_set_column_split",
    "hpxfft_exec_r2c_2d_noalloc", "hpxfft_plan_set_row_kernel", "hpxfft_last_error", "hpxfft_version",
)


class HpxfftError(RuntimeError):
    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


_lib = None


def lib() -> ctypes.CDLL:
    """Load libhpxfft_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, ci = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    pplan = ctypes.POINTER(ctypes.c_void_p)
    L.hpxfft_plan_flag_from_string.argtypes = [ctypes.c_char_p, ctypes.POINTER(ci)]
    L.hpxfft_plan_r2c_2d.argtypes = [pplan, sz, sz, sz, ci, ci, vp]
    L.hpxfft_plan_r2c_2d_many.argtypes = [pplan, sz, sz, sz, sz, sz, sz, sz, ci, ci, vp]
    L.hpxfft_exec_r2c_2d.argtypes = [vp, vp, vp]
    L.hpxfft_exec_r2c_2d_scaled.argtypes = [vp, vp, vp, vp, vp]
    L.hpxfft_exec_r2c_2d_host.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.hpxfft_exec_r2c_2d_adjoint.argtypes = [vp, vp, vp]
    L.hpxfft_exec_c2r_2d.argtypes = [vp, vp, vp]
    L.hpxfft_get_measurement.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
    L.hpxfft_write_plan.argtypes = [vp, ctypes.c_char_p]
    L.hpxfft_plan_describe.argtypes = [vp, ctypes.c_char_p, sz, ctypes.POINTER(sz)]
    L.hpxfft_plan_destroy.argtypes = [vp]
    L.hpxfft_plan_r2c_2d_batched.argtypes = [pplan, sz, ctypes.POINTER(sz), ctypes.POINTER(sz), ci, ci, vp]
    L.hpxfft_exec_r2c_2d_batched.argtypes = [vp, ctypes.POINTER(vp)]
    L.hpxfft_plan_batched_set_config.argtypes = [vp, ci, ci]
    L.hpxfft_plan_set_column_split.argtypes = [vp, ci]
    L.hpxfft_exec_r2c_2d_noalloc.argtypes = [sz, sz, sz, sz, ci, vp, vp, vp]
    L.hpxfft_plan_set_row_kernel.argtypes = [vp, ci]
    L.hpxfft_last_error.argtypes = []
    L.hpxfft_last_error.restype = ctypes.c_char_p
    L.hpxfft_version.argtypes = []
    L.hpxfft_version.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name not in ("hpxfft_last_error", "hpxfft_version"):
            getattr(L, name).restype = ci
    _lib = L
    return L


def last_error() -> str:
    return lib().hpxfft_last_error().decode()


def check(status: int) -> None:
    if status != OK:
        raise HpxfftError(status, last_error())
