"""CPU FP64 oracle for the 2-D r2c FFT path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this package, and only as the checker or the timed CPU comparator.
The product (``hpx-fft_amd``) never imports it.

ctypes wrapper over ``oracle/liboracle.so`` (built from ``oracle/oracle.c`` by
``oracle/Makefile``).  Functions and the reference code they restate:

* :func:`fill` -- SplitMix64 seeded inputs; ``dist="ramp"`` is the harness
  input of examples/hpxfft/shared_loop_2d.cpp:33-40.
* :func:`r2c_2d` / :class:`Loop` -- hpxfft::shared::loop::initialize +
  fft_2d_r2c_par (core/src/shared/loop.cpp:158-189, :56-113): row r2c
  (FFTW r2c_1d, core/src/util/adapter_fftw.cpp:9,14), transpose (:18-25),
  column c2c FFTW_FORWARD (adapter_fftw.cpp:28-29,34), transpose back
  (:46-53), with the same per-phase measurement keys (:106-110).
* :func:`dft_r2c_2d_brute` -- long double O(N^2) DFT, the pin for the above.
* :func:`adjoint_r2c_2d`, :func:`adjoint_brute` -- the adjoint of the r2c map
  (no reference counterpart; pinned by brute force and by the
  <R x, Y> = <x, R* Y> identity in the tests).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

DISTS = {"u01": 0, "upm1": 1, "normal": 2, "ramp": 3}
PHASE_KEYS = ("total", "first_fftw", "first_trans", "second_fftw", "second_trans")


def build(force: bool = False) -> str:
    """Compile oracle.c with gcc (make)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        sz = ctypes.c_size_t
        L.oracle_fill.argtypes = [dp, sz, sz, sz, ctypes.c_uint64, ctypes.c_int]
        L.oracle_fill.restype = None
        L.oracle_twiddle.argtypes = [ctypes.c_uint64, ctypes.c_uint64, dp, dp]
        L.oracle_twiddle.restype = None
        L.oracle_fft_c2c_1d.argtypes = [dp, sz, ctypes.c_int]
        L.oracle_fft_c2c_1d.restype = ctypes.c_int
        L.oracle_fft_r2c_1d.argtypes = [dp, dp, sz]
        L.oracle_fft_r2c_1d.restype = ctypes.c_int
        L.oracle_loop_init.argtypes = [sz, sz, sz]
        L.oracle_loop_init.restype = ctypes.c_void_p
        L.oracle_loop_free.argtypes = [ctypes.c_void_p]
        L.oracle_loop_free.restype = None
        L.oracle_loop_exec.argtypes = [ctypes.c_void_p, dp, ctypes.c_int, dp]
        L.oracle_loop_exec.restype = ctypes.c_int
        L.oracle_r2c_2d.argtypes = [dp, sz, sz, sz, ctypes.c_int]
        L.oracle_r2c_2d.restype = ctypes.c_int
        L.oracle_max_threads.argtypes = []
        L.oracle_max_threads.restype = ctypes.c_int
        L.oracle_dft_r2c_2d_brute.argtypes = [dp, sz, sz, sz, dp]
        L.oracle_dft_r2c_2d_brute.restype = None
        L.oracle_adjoint_r2c_2d_brute.argtypes = [dp, sz, sz, dp, sz]
        L.oracle_adjoint_r2c_2d_brute.restype = None
        L.oracle_adjoint_r2c_2d.argtypes = [dp, sz, sz, dp, sz, ctypes.c_int]
        L.oracle_adjoint_r2c_2d.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def padded_cols(ny: int) -> int:
    """Row pitch (in doubles) of the in-place padded layout: 2*(Ny/2+1)."""
    return 2 * (ny // 2 + 1)


def fill(nx: int, ny: int, seed: int, dist: str = "upm1", ld: int | None = None) -> np.ndarray:
    """Nx x ld float64 array, real samples in [:, :ny], zero padding after."""
    ld = padded_cols(ny) if ld is None else ld
    a = np.empty((nx, ld), dtype=np.float64)
    lib().oracle_fill(_ptr(a), nx, ny, ld, seed & 0xFFFFFFFFFFFFFFFF, DISTS[dist])
    return a


def twiddle(m: int, n: int) -> complex:
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().oracle_twiddle(m, n, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def fft_c2c_1d(x: np.ndarray, sign: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    v = a.view(np.float64)
    if lib().oracle_fft_c2c_1d(_ptr(v), a.size, sign) != 0:
        raise ValueError("oracle_fft_c2c_1d failed")
    return a


def fft_r2c_1d(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.size
    out = np.empty(n + 2, dtype=np.float64)
    if lib().oracle_fft_r2c_1d(_ptr(x), _ptr(out), n) != 0:
        raise ValueError("oracle_fft_r2c_1d: n must be even and >= 2")
    return out.view(np.complex128)


class Loop:
    """Restatement of hpxfft::shared::loop on the CPU (one plan, many execs)."""

    def __init__(self, nx: int, ny: int, ld: int | None = None):
        self.nx, self.ny = nx, ny
        self.ld = padded_cols(ny) if ld is None else ld
        self._h = lib().oracle_loop_init(nx, ny, self.ld)
        if not self._h:
            raise ValueError(f"oracle: unsupported size {nx}x{ny}")
        self.measurements: dict[str, float] = {}

    def exec(self, values: np.ndarray, nthreads: int = 0) -> np.ndarray:
        """In-place 2-D r2c of an Nx x ld float64 array; returns it."""
        assert values.shape == (self.nx, self.ld)
        t = np.zeros(5, dtype=np.float64)
        if lib().oracle_loop_exec(self._h, _ptr(values), nthreads, _ptr(t)) != 0:
            raise RuntimeError("oracle_loop_exec failed")
        self.measurements = dict(zip(PHASE_KEYS, t.tolist()))
        return values

    def close(self):
        if self._h:
            lib().oracle_loop_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


def r2c_2d(x: np.ndarray, ny: int | None = None, nthreads: int = 0) -> np.ndarray:
    """2-D r2c of a real array.

    ``x`` is either an Nx x Ny real array, or (with ``ny`` given) an Nx x ld
    padded array.  Returns the Nx x (Ny/2+1) complex128 spectrum (a new array).
    """
    x = np.asarray(x, dtype=np.float64)
    nx = x.shape[0]
    if ny is None:
        ny = x.shape[1]
        buf = np.zeros((nx, padded_cols(ny)), dtype=np.float64)
        buf[:, :ny] = x
    else:
        buf = np.ascontiguousarray(x).copy()
    ld = buf.shape[1]
    if lib().oracle_r2c_2d(_ptr(buf), nx, ny, ld, nthreads) != 0:
        raise ValueError(f"oracle: unsupported size {nx}x{ny}")
    return buf[:, : padded_cols(ny)].copy().view(np.complex128)


def dft_r2c_2d_brute(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    nx, ny = x.shape
    out = np.empty((nx, padded_cols(ny)), dtype=np.float64)
    lib().oracle_dft_r2c_2d_brute(_ptr(x), nx, ny, ny, _ptr(out))
    return out.view(np.complex128)


def adjoint_brute(Y: np.ndarray, ny: int) -> np.ndarray:
    Y = np.ascontiguousarray(Y, dtype=np.complex128)
    nx = Y.shape[0]
    assert Y.shape[1] == ny // 2 + 1
    out = np.empty((nx, ny), dtype=np.float64)
    lib().oracle_adjoint_r2c_2d_brute(_ptr(Y.view(np.float64)), nx, ny, _ptr(out), ny)
    return out


def adjoint_r2c_2d(Y: np.ndarray, ny: int, nthreads: int = 0) -> np.ndarray:
    Y = np.ascontiguousarray(Y, dtype=np.complex128)
    nx = Y.shape[0]
    assert Y.shape[1] == ny // 2 + 1
    out = np.empty((nx, ny), dtype=np.float64)
    if lib().oracle_adjoint_r2c_2d(_ptr(Y.view(np.float64)), nx, ny, _ptr(out), ny, nthreads) != 0:
        raise ValueError("oracle_adjoint_r2c_2d failed")
    return out


def c2r_2d(X: np.ndarray, ny: int, nthreads: int = 0) -> np.ndarray:
    """Unnormalised inverse 2-D c2r (FFTW_BACKWARD, the direction the reference
    declares at core/include/hpxfft/util/adapter_fftw.hpp:11): the Hermitian
    extension of the half spectrum, i.e. R*(c X) with c(ky) = 2 for
    0 < ky < Ny/2 and 1 at ky = 0, Ny/2 (R* halves exactly those bins)."""
    X = np.array(X, dtype=np.complex128)
    if ny >= 4:
        X[:, 1:ny // 2] *= 2.0
    return adjoint_r2c_2d(X, ny, nthreads)


def rel_l2(a: np.ndarray, ref: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((a - ref).ravel())
    return float(num / den) if den > 0 else float(num)
