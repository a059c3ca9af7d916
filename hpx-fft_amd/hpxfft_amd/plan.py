"""Python mirror of the HPX-FFT shared-memory surface over the C ABI.

``Plan`` wraps one hpxfft_plan (device memory is passed as raw pointers,
normally from torch tensors).  ``Loop`` mirrors hpxfft::shared::loop
(core/include/hpxfft/shared/loop.hpp:15-57): ``initialize(values, flag)``
takes an Nx x 2(Ny/2+1) float64 host array, ``fft_2d_r2c_par()`` /
``fft_2d_r2c_seq()`` / ``fft_2d_r2c()`` return it transformed in place,
``get_measurement(key)`` and ``write_plans_to_file(path)`` behave as the
reference's (unknown key -> 0.0; unopenable file -> RuntimeError).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

PRECISIONS = {"f64": _lib.F64, "f32": _lib.F32, "f16": _lib.F16_STORAGE, "bf16": _lib.BF16_STORAGE}


def plan_flag_to_effort(flag: str) -> int:
    """hpxfft::util::fftw_adapter::string_to_fftw_plan_flag (adapter_fftw.hpp:22-44)."""
    eff = ctypes.c_int(0)
    st = lib().hpxfft_plan_flag_from_string(flag.encode(), ctypes.byref(eff))
    if st != _lib.OK:
        raise ValueError(_lib.last_error())   # std::invalid_argument in the reference
    return eff.value


def _stream_handle(stream) -> int:
    if stream is None:
        return 0
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


class Plan:
    """One 2-D r2c plan (optionally a batch of equal grids)."""

    def __init__(self, nx: int, ny: int, *, batch: int = 1, ld_in: int | None = None, dist_in: int | None = None,
                 ld_out: int | None = None, dist_out: int | None = None, precision: str = "f64",
                 effort: int | str = "estimate", stream=None):
        self.nx, self.ny, self.batch = nx, ny, batch
        self.ncols = ny // 2 + 1
        self.ld_in = 2 * self.ncols if ld_in is None else ld_in
        self.ld_out = self.ncols if ld_out is None else ld_out
        self.dist_in = nx * self.ld_in if dist_in is None else dist_in
        self.dist_out = nx * self.ld_out if dist_out is None else dist_out
        self.precision = precision
        eff = plan_flag_to_effort(effort) if isinstance(effort, str) else int(effort)
        h = ctypes.c_void_p()
        check(lib().hpxfft_plan_r2c_2d_many(ctypes.byref(h), nx, ny, batch, self.ld_in, self.dist_in, self.ld_out,
                                            self.dist_out, PRECISIONS[precision], eff, _stream_handle(stream)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def exec(self, d_in: int, d_out: int) -> None:
        check(lib().hpxfft_exec_r2c_2d(self._h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out)))

    def exec_scaled(self, d_in: int, d_out: int, d_in_row_exp: int | None, d_out_col_exp: int) -> None:
        check(lib().hpxfft_exec_r2c_2d_scaled(self._h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out),
                                              ctypes.c_void_p(d_in_row_exp or 0), ctypes.c_void_p(d_out_col_exp)))

    def exec_adjoint(self, d_spec: int, d_real: int) -> None:
        check(lib().hpxfft_exec_r2c_2d_adjoint(self._h, ctypes.c_void_p(d_spec), ctypes.c_void_p(d_real)))

    def exec_c2r(self, d_spec: int, d_real: int) -> None:
        """Unnormalised inverse (FFTW_BACKWARD): exec_c2r after exec gives Nx*Ny*x."""
        check(lib().hpxfft_exec_c2r_2d(self._h, ctypes.c_void_p(d_spec), ctypes.c_void_p(d_real)))

    def exec_host(self, values: np.ndarray) -> None:
        if values.dtype != np.float64 or not values.flags.c_contiguous:
            raise ValueError("exec_host takes a C-contiguous float64 array (the padded Nx x 2(Ny/2+1) buffer)")
        check(lib().hpxfft_exec_r2c_2d_host(self._h, values.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def set_column_split(self, log2_p: int) -> None:
        """Run the four-step column pass with Nx = P*Q, P = 2^log2_p (tests
        use it to check every split the measure effort may keep)."""
        check(lib().hpxfft_plan_set_column_split(self._h, int(log2_p)))

    def set_row_kernel(self, kind: int) -> None:
        """Rows of 2^13 / 2^14 complex points, FP32 arithmetic: 0 = radix-16
        stages, 1 = radix-32 stages across lane pairs."""
        check(lib().hpxfft_plan_set_row_kernel(self._h, int(kind)))

    def measurement(self, key: str) -> float:
        v = ctypes.c_double(0.0)
        check(lib().hpxfft_get_measurement(self._h, key.encode(), ctypes.byref(v)))
        return v.value

    def describe(self) -> str:
        need = ctypes.c_size_t(0)
        check(lib().hpxfft_plan_describe(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value + 1)
        check(lib().hpxfft_plan_describe(self._h, buf, need.value + 1, None))
        return buf.value.decode()

    def write_plan(self, path: str) -> None:
        st = lib().hpxfft_write_plan(self._h, path.encode())
        if st == _lib.IO_ERROR:
            raise RuntimeError(_lib.last_error())   # std::runtime_error in the reference
        check(st)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().hpxfft_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchedPlan:
    """Many independent in-place FP32 grids of uneven sizes (sides 64..2048),
    transformed by two persistent launches (row phase, column phase).

    ``effort`` "measure" (or higher) times every candidate task shape on the
    plan's sizes and keeps the fastest per phase (the autotuner).
    """

    def __init__(self, nx, ny, *, effort: int | str = "estimate", stream=None):
        if len(nx) != len(ny) or len(nx) == 0:
            raise ValueError("need equal, non-empty lists of Nx and Ny")
        self.nx, self.ny = [int(v) for v in nx], [int(v) for v in ny]
        n = len(self.nx)
        eff = plan_flag_to_effort(effort) if isinstance(effort, str) else int(effort)
        arr = ctypes.c_size_t * n
        h = ctypes.c_void_p()
        check(lib().hpxfft_plan_r2c_2d_batched(ctypes.byref(h), n, arr(*self.nx), arr(*self.ny), _lib.F32, eff,
                                               _stream_handle(stream)))
        self._h = h
        self._ptrs = (ctypes.c_void_p * n)()

    def exec(self, d_grids) -> None:
        if len(d_grids) != len(self.nx):
            raise ValueError(f"need {len(self.nx)} device pointers, one per grid")
        for i, p in enumerate(d_grids):
            self._ptrs[i] = p
        check(lib().hpxfft_exec_r2c_2d_batched(self._h, self._ptrs))

    def set_config(self, row_cfg: int, col_cfg: int) -> None:
        """Run the given task shapes (0..3) instead of the chosen ones."""
        check(lib().hpxfft_plan_batched_set_config(self._h, int(row_cfg), int(col_cfg)))

    measurement = Plan.measurement
    describe = Plan.describe
    close = Plan.close
    __del__ = Plan.__del__


class Loop:
    """hpxfft::shared::loop on the GPU, host buffers in and out."""

    def __init__(self):
        self._plan: Plan | None = None
        self._values: np.ndarray | None = None
        self._meas: dict[str, float] = {}

    def initialize(self, values: np.ndarray, plan_flag: str) -> None:
        """loop.cpp:158-189: Nx = n_row, Ny = n_col - 2 (n_col = 2 (Ny/2+1))."""
        effort = plan_flag_to_effort(plan_flag)
        values = np.ascontiguousarray(values, dtype=np.float64)
        n_row, n_col = values.shape
        if n_col % 2 or n_col < 4:
            raise ValueError("n_col must be even and >= 4 (Ny = n_col - 2)")
        if self._plan is not None:
            self._plan.close()
        self._plan = Plan(n_row, n_col - 2, ld_in=n_col, precision="f64", effort=effort)
        self._values = values
        self._meas = {"plan": self._plan.measurement("plan"), "plan_flops": self._plan.measurement("plan_flops")}

    def fft_2d_r2c_par(self) -> np.ndarray:
        """loop.cpp:56-113; the result is moved out (re-initialize before the next call)."""
        if self._values is None or self._plan is None:
            raise RuntimeError("fft_2d_r2c called without initialize")
        v, self._values = self._values, None
        self._plan.exec_host(v)
        for k in ("total", "first_fftw", "first_trans", "second_fftw", "second_trans", "h2d", "d2h"):
            self._meas[k] = self._plan.measurement(k)
        return v

    fft_2d_r2c_seq = fft_2d_r2c_par
    fft_2d_r2c = fft_2d_r2c_par

    def get_measurement(self, name: str) -> float:
        return self._meas.setdefault(name, 0.0)

    def write_plans_to_file(self, file_path: str) -> None:
        if self._plan is None:
            raise RuntimeError("write_plans_to_file called without initialize")
        self._plan.write_plan(file_path)


# the task-graph variants share the surface (sync.hpp:23-27, opt.hpp:23-27, naive.hpp:23-27)
Sync = Opt = Naive = Loop
