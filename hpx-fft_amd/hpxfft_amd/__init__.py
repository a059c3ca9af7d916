"""hpxfft_amd -- MI355X (gfx950) 2-D real-to-complex FFT behind HPX-FFT's
shared-memory surface.

All arithmetic runs in hand-written HIP kernels in ``lib/libhpxfft_hip.so``
reached through the C ABI of ``include/hpxfft_hip.h``; this package is the
Python host layer (ctypes) and the torch binding.
"""
from ._lib import HpxfftError, lib, last_error, LIB_PATH  # noqa: F401
from .plan import Plan, BatchedPlan, Loop, Sync, Opt, Naive, plan_flag_to_effort  # noqa: F401

__all__ = ["HpxfftError", "lib", "last_error", "LIB_PATH", "Plan", "BatchedPlan", "Loop", "Sync", "Opt", "Naive",
           "plan_flag_to_effort"]


def __getattr__(name):
    # torch-dependent entry points are imported lazily
    if name in ("rfft2", "rfft2_adjoint", "c2r_2d", "rfft2_scaled", "RFFT2", "rfft2_autograd", "clear_plan_cache",
                "set_planning_effort", "planning_effort"):
        from . import torch_api
        return getattr(torch_api, name)
    raise AttributeError(name)
