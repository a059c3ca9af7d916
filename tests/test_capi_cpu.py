"""C ABI and host-side surface, without a GPU.

* libhpxfft_hip.so loads and exports every function include/hpxfft_hip.h declares;
* plan-flag strings map as adapter_fftw.hpp:22-44 (invalid -> INVALID_ARG,
  "Invalid FFTW plan flag string");
* argument, size and layout errors come back as status codes (no exception
  crosses the ABI), and with no HIP device plan creation returns NO_DEVICE;
* the C++ programs built on the drop-in class report those errors cleanly.
No compute call is made.
"""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import REPO

import hpxfft_amd
from hpxfft_amd import _lib

HEADER = os.path.join(REPO, "include", "hpxfft_hip.h")
BIN = os.path.join(REPO, "hpx-fft_amd", "bin")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hpxfft_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = hpxfft_amd.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_version_and_last_error():
    L = hpxfft_amd.lib()
    assert b"gfx950" in L.hpxfft_version()
    assert isinstance(hpxfft_amd.last_error(), str)


@pytest.mark.parametrize("flag,effort", [("estimate", 0), ("measure", 1), ("patient", 2), ("exhaustive", 3)])
def test_plan_flags(flag, effort):
    assert hpxfft_amd.plan_flag_to_effort(flag) == effort


@pytest.mark.parametrize("flag", ["", "Estimate", "fastest", "estimate "])
def test_invalid_plan_flag(flag):
    with pytest.raises(ValueError, match="Invalid FFTW plan flag string"):
        hpxfft_amd.plan_flag_to_effort(flag)


def plan_status(nx, ny, batch=1, ld_in=None, dist_in=None, ld_out=None, dist_out=None, prec=0, effort=0):
    L = hpxfft_amd.lib()
    ncy = ny // 2 + 1
    ld_in = 2 * ncy if ld_in is None else ld_in
    ld_out = ncy if ld_out is None else ld_out
    dist_in = nx * ld_in if dist_in is None else dist_in
    dist_out = nx * ld_out if dist_out is None else dist_out
    h = ctypes.c_void_p()
    st = L.hpxfft_plan_r2c_2d_many(ctypes.byref(h), nx, ny, batch, ld_in, dist_in, ld_out, dist_out, prec, effort,
                                   None)
    if h.value:
        L.hpxfft_plan_destroy(h)
    return st


@pytest.mark.parametrize("nx,ny", [(16, 1), (16, 15), (0, 16), (16, 0), (1 << 25, 2), (2, 1 << 26),
                                   (4194319, 2), (2, 2 * 4194319), (2 * 2097169, 2)])
def test_unsupported_sizes(nx, ny):
    """Odd Ny, zero sizes, and lengths no plan covers: more than 2^24 points
    (Nx = 2^25, Ny/2 = 2^25), and lengths above 2^22 (the longest that
    Bluestein's algorithm takes over two levels) with no split into two
    one-workgroup lengths: the prime 4194319, and 2 x the prime 2097169."""
    assert plan_status(nx, ny) == _lib.UNSUPPORTED_SIZE
    assert hpxfft_amd.last_error()


@pytest.mark.parametrize("nx,ny", [(1 << 21, 4), (4, 1 << 22), (16, 2 * 4099), (4099, 16), (16, 2 * 8193),
                                   (2 * 4099, 8), ((1 << 20) + 2, 4), (1 << 24, 2), (4194301, 2),
                                   (2, 2 * 4194301)])
def test_long_and_large_prime_lengths_pass_validation(nx, ny):
    """FP64 lengths beyond in-LDS Bluestein with no split into two
    one-workgroup lengths (4099 and 4194301 are prime; 8193 = 3 * 2731, and
    2731 needs an 8192-point FFT) run Bluestein's algorithm over two levels;
    lengths beyond 2^20 with such a split run as two levels (up to 2^24)
    (without a device the plan then stops at the device check)."""
    st = plan_status(nx, ny)
    if st == _lib.OK:
        pytest.skip("a HIP device is present")
    assert st == _lib.NO_DEVICE, hpxfft_amd.last_error()


def test_fp32_bluestein_reaches_twice_as_far():
    """FP32 holds twice the points per buffer: Ny/2 = 8191 (prime) fits its
    16384-point in-place Bluestein FFT; FP64 (8192 points in place) takes it
    by Bluestein's algorithm over two levels.  Both plans pass validation."""
    for prec in (0, 1):
        st = plan_status(16, 2 * 8191, prec=prec)
        if st == _lib.OK:
            pytest.skip("a HIP device is present")
        assert st == _lib.NO_DEVICE, hpxfft_amd.last_error()


@pytest.mark.parametrize("prec", [2, 3])
def test_general_lengths_16bit(prec):
    """16-bit storage takes lengths that are not powers of two within one
    workgroup (8 x 14, Bluestein lengths, 8191, whose Bluestein FFT runs in
    place only, and 10000 = 2^4 5^4, in place in one buffer), and beyond one
    workgroup (20000 as two levels, the prime 8209 by Bluestein's algorithm
    over two levels, 2^17) through an FP32 workspace; lengths no FP32 plan
    covers are refused."""
    for nx, ny in [(8, 14), (67, 16), (16, 134), (8191, 8), (8, 2 * 8191), (10000, 8), (20000, 8),
                   (8, 2 * 20000), (8209, 4), (4, 1 << 18)]:
        st = plan_status(nx, ny, prec=prec)
        if st == _lib.OK:
            pytest.skip("a HIP device is present")
        assert st == _lib.NO_DEVICE, (nx, ny, hpxfft_amd.last_error())
    assert plan_status(4194319, 2, prec=prec) == _lib.UNSUPPORTED_SIZE


@pytest.mark.parametrize("nx,ny", [(8, 14), (3, 16), (1000, 64), (64, 1000), (4095, 2),
                                   (16, 134), (67, 16), (2039, 2 * 1031),
                                   (6000, 8), (16, 2 * 6000), (1 << 17, 4), (4, 1 << 17), (3 * 2039, 2),
                                   (4093, 8), (8, 2 * 4093)])
def test_general_lengths_pass_validation(nx, ny):
    """Lengths that are not powers of two are accepted: prime factors above 61
    through Bluestein's algorithm (4093: an 8192-point FFT, in place only),
    lengths beyond one workgroup as two levels N = P*Q, as are powers of two
    beyond the power-of-two kernels (Nx > 2^16, Ny > 2^15) (without a device
    the plan then stops at the device check)."""
    st = plan_status(nx, ny)
    if st == _lib.OK:
        pytest.skip("a HIP device is present")
    assert st == _lib.NO_DEVICE, hpxfft_amd.last_error()


def test_bad_arguments():
    L = hpxfft_amd.lib()
    assert plan_status(16, 16, batch=0) == _lib.INVALID_ARG
    assert plan_status(16, 16, prec=7) == _lib.INVALID_ARG
    assert plan_status(16, 16, effort=9) == _lib.INVALID_ARG
    assert plan_status(16, 16, ld_in=8) == _lib.INVALID_ARG          # ld_in < Ny
    assert plan_status(16, 16, ld_out=4) == _lib.INVALID_ARG         # ld_out < Ny/2+1
    assert plan_status(16, 16, ld_in=17) == _lib.INVALID_ARG         # odd pitch
    assert plan_status(16, 16, batch=2, dist_in=8) == _lib.INVALID_ARG
    # grids beyond a launch dimension, and strides past a 64-bit element index
    assert plan_status(16, 16, batch=65536) == _lib.UNSUPPORTED_SIZE
    assert "65535 grids" in hpxfft_amd.last_error()
    assert plan_status(16, 16, ld_in=1 << 61, dist_in=0) == _lib.INVALID_ARG
    assert "2^62" in hpxfft_amd.last_error()
    assert plan_status(16, 16, batch=2, dist_in=1 << 62) == _lib.INVALID_ARG
    assert plan_status(16, 16, batch=3, ld_out=1 << 60, dist_out=1 << 61, ld_in=18, dist_in=288) == _lib.INVALID_ARG
    assert L.hpxfft_plan_r2c_2d_many(None, 16, 16, 1, 18, 288, 9, 144, 0, 0, None) == _lib.INVALID_ARG
    assert L.hpxfft_exec_r2c_2d(None, None, None) == _lib.INVALID_ARG
    assert L.hpxfft_exec_r2c_2d_adjoint(None, None, None) == _lib.INVALID_ARG
    assert L.hpxfft_exec_r2c_2d_host(None, None) == _lib.INVALID_ARG
    assert L.hpxfft_exec_r2c_2d_scaled(None, None, None, None, None) == _lib.INVALID_ARG
    v = ctypes.c_double()
    assert L.hpxfft_get_measurement(None, b"total", ctypes.byref(v)) == _lib.INVALID_ARG
    assert L.hpxfft_write_plan(None, b"/tmp/x") == _lib.INVALID_ARG
    assert L.hpxfft_plan_batched_set_config(None, 0, 0) == _lib.INVALID_ARG
    assert L.hpxfft_plan_destroy(None) == _lib.OK
    eff = ctypes.c_int()
    assert L.hpxfft_plan_flag_from_string(None, ctypes.byref(eff)) == _lib.INVALID_ARG


def test_header_builds_as_c99(tmp_path):
    """include/hpxfft_hip.h is plain C: a C99 program includes it, links the
    library and calls the entry points that need no device."""
    exe = tmp_path / "abi_c99"
    lib_dir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                    "-I", os.path.join(REPO, "include"), os.path.join(REPO, "tests", "c", "abi_c99.c"),
                    "-L", lib_dir, "-lhpxfft_hip", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stderr)
    assert "hpxfft-hip" in r.stdout


def test_no_device_is_reported_cleanly():
    st = plan_status(64, 64)
    if st == _lib.OK:
        pytest.skip("a HIP device is present")
    assert st == _lib.NO_DEVICE
    assert "no HIP device" in hpxfft_amd.last_error()
    with pytest.raises(hpxfft_amd.HpxfftError):
        hpxfft_amd.Plan(64, 64)


def test_python_loop_surface_errors():
    loop = hpxfft_amd.Loop()
    import numpy as np

    with pytest.raises(ValueError, match="Invalid FFTW plan flag string"):
        loop.initialize(np.zeros((4, 6)), "bogus")
    with pytest.raises(RuntimeError):
        loop.fft_2d_r2c_par()
    assert loop.get_measurement("anything") == 0.0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()


def test_cpp_vector_2d():
    r = subprocess.run([os.path.join(BIN, "test_vector_2d")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout


def test_cpp_harness_argument_errors(tmp_path):
    exe = os.path.join(BIN, "shared_loop_2d_gpu")
    r = subprocess.run([exe, "--plan=bogus", "--nx", "4", "--ny", "4"], capture_output=True, text=True, timeout=60,
                       cwd=tmp_path)
    assert r.returncode == 2 and "Invalid FFTW plan flag string" in r.stderr
    r = subprocess.run([exe, "--nx", "1", "--ny", "8388638"], capture_output=True, text=True, timeout=60,
                       cwd=tmp_path)
    assert r.returncode == 2 and "prime factors" in r.stderr       # Ny/2 = 4194319, a prime above 2^22


@pytest.mark.parametrize("args,msg", [
    (["--header=maybe"], "('maybe') for option '--header' is invalid"),
    (["--result", "2"], "('2') for option '--result' is invalid"),
    (["--nx=12x"], "('12x') for option '--nx' is invalid"),
    (["--ny=-4"], "('-4') for option '--ny' is invalid"),
    (["--threads=4"], "unrecognised option '--threads'"),
    (["--nx"], "option '--nx' requires an argument"),
    (["8"], "unexpected argument '8'"),
])
def test_cpp_harness_usage_errors(tmp_path, args, msg):
    """Command-line values are checked as program_options checks them
    (shared_loop_2d.cpp:140-146): a usage error, exit status 2, before any
    device work and before any file is written."""
    r = subprocess.run([os.path.join(BIN, "shared_loop_2d_gpu")] + args, capture_output=True, text=True,
                       timeout=60, cwd=tmp_path)
    assert r.returncode == 2 and msg in r.stderr, r.stderr
    assert not (tmp_path / "runtimes").exists()


@pytest.mark.parametrize("word", ["true", "TRUE", "yes", "on", "1", "false", "No", "off", "0"])
def test_cpp_harness_bool_words(tmp_path, word):
    """--header / --result take program_options' bool words, as the
    reference's size sweep passes --header=true.  The unsupported size makes
    the run stop after argument parsing, at plan validation (status 2)."""
    r = subprocess.run([os.path.join(BIN, "shared_loop_2d_gpu"), "--header=" + word, "--result", word, "--nx=1",
                        "--ny=8388638"], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode == 2 and "prime factors" in r.stderr and "usage error" not in r.stderr, r.stderr


SWEEP = os.path.join(REPO, "hpx-fft_amd", "cpp", "run_size_sweep.sh")


def test_size_sweep_script_drives_the_harness(tmp_path):
    """The sweep (argument order of run_hpxfft_size_shared.sh:10-15) with a
    stand-in executable that records its arguments: sizes 2^3..2^5, two runs
    each, --header=true on the first run only."""
    stub = tmp_path / "stub.sh"
    stub.write_text("#!/bin/bash\necho \"$@\" >> calls.txt\n")
    stub.chmod(0o755)
    r = subprocess.run(["bash", SWEEP, str(stub), "measure", "64", "3", "5", "2"], capture_output=True, text=True,
                       timeout=60, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    calls = (tmp_path / "calls.txt").read_text().splitlines()
    want = [f"--nx={n} --ny={n} --plan=measure --header={'true' if i == 0 else 'false'}"
            for i, n in enumerate([8, 8, 16, 16, 32, 32])]
    assert calls == want


def test_size_sweep_script_stops_at_a_failed_run(tmp_path):
    stub = tmp_path / "stub.sh"
    stub.write_text("#!/bin/bash\necho \"$@\" >> calls.txt\ncase \"$1\" in --nx=16) exit 2;; esac\n")
    stub.chmod(0o755)
    r = subprocess.run(["bash", SWEEP, str(stub), "estimate", "1", "3", "6", "1"], capture_output=True, text=True,
                       timeout=60, cwd=tmp_path)
    assert r.returncode == 2 and "nx=ny=16" in r.stderr
    assert len((tmp_path / "calls.txt").read_text().splitlines()) == 2
    r = subprocess.run(["bash", SWEEP, str(stub)], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode == 2 and "usage" in r.stderr


def test_cpp_known_answer_without_device():
    st = plan_status(4, 4)
    if st == _lib.OK:
        pytest.skip("a HIP device is present (covered by the GPU test)")
    r = subprocess.run([os.path.join(BIN, "test_shared_loop_gpu")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "no HIP device" in r.stderr


def test_batched_plan_argument_errors():
    L = hpxfft_amd.lib()
    arr = ctypes.c_size_t * 3
    h = ctypes.c_void_p()
    # side outside [64, 2048] or not a power of two -> UNSUPPORTED_SIZE (checked before any device call)
    assert L.hpxfft_plan_r2c_2d_batched(ctypes.byref(h), 3, arr(64, 128, 32), arr(64, 64, 64), 1, 0,
                                        None) == _lib.UNSUPPORTED_SIZE
    assert L.hpxfft_plan_r2c_2d_batched(ctypes.byref(h), 3, arr(64, 128, 4096), arr(64, 64, 64), 1, 0,
                                        None) == _lib.UNSUPPORTED_SIZE
    assert L.hpxfft_plan_r2c_2d_batched(ctypes.byref(h), 3, arr(64, 96, 64), arr(64, 64, 64), 1, 0,
                                        None) == _lib.UNSUPPORTED_SIZE
    # FP32 only; count >= 1
    assert L.hpxfft_plan_r2c_2d_batched(ctypes.byref(h), 3, arr(64, 64, 64), arr(64, 64, 64), 0, 0,
                                        None) == _lib.INVALID_ARG
    assert L.hpxfft_plan_r2c_2d_batched(ctypes.byref(h), 0, arr(64, 64, 64), arr(64, 64, 64), 1, 0,
                                        None) == _lib.INVALID_ARG
    assert L.hpxfft_exec_r2c_2d_batched(None, None) == _lib.INVALID_ARG


def test_torch_planning_effort_setting():
    """The torch entry points' planning effort: the reference's flag strings
    only (adapter_fftw.hpp:22-44), estimate by default."""
    import hpxfft_amd
    assert hpxfft_amd.planning_effort() == "estimate"
    with pytest.raises(ValueError, match="Invalid FFTW plan flag string"):
        hpxfft_amd.set_planning_effort("quick")
    hpxfft_amd.set_planning_effort("measure")
    assert hpxfft_amd.planning_effort() == "measure"
    hpxfft_amd.set_planning_effort("estimate")
