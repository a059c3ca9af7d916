"""Precision variants and the C++ drop-in programs on the GPU.

* FP16 / BF16 storage with FP32 arithmetic and power-of-two scales: the
  spectrum reconstructed as X16 * 2^-col_exp is compared with the FP64
  oracle of the original (unrounded) input, at lengths within one workgroup
  and beyond it (an FP32 workspace).  Bounds: FP16 rel-L2 <= 5e-3,
  BF16 <= 4e-2 (11- and 8-bit significands; these variants report their
  error rather than meet the FP64 bar).  Inputs of magnitude 1e6 exercise the
  input row scales (they would overflow FP16 unscaled).
* FP32 at 2^14 x 2^14 against the oracle: rel-L2 <= 1e-5.
* FP64 at 2^14 x 2^14 (the headline workload) against the oracle: <= 1e-12.
* the C++ known-answer test and the reference-CLI harness.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import REPO

pytestmark = pytest.mark.gpu
BIN = os.path.join(REPO, "hpx-fft_amd", "bin")


def to16(x_pad, ny, dt, scale_rows):
    """FP64 padded rows -> 16-bit tensor plus per-row exponents (value = stored * 2^-e)."""
    import torch
    nx = x_pad.shape[0]
    e = np.zeros(nx, dtype=np.int32)
    if scale_rows:
        m = np.max(np.abs(x_pad[:, :ny]), axis=1)
        e = (10 - np.floor(np.log2(np.maximum(m, 1e-300)))).astype(np.int32)   # row max -> [2^10, 2^11)
    scaled = x_pad * np.exp2(e.astype(np.float64))[:, None]
    t = torch.from_numpy(scaled).to(dtype=dt, device="cuda")
    return t, torch.from_numpy(e).cuda(), e


@pytest.mark.parametrize("dt_name,bound", [("f16", 5e-3), ("bf16", 4e-2)])
@pytest.mark.parametrize("nx,ny,amp", [(8, 16, 1.0), (64, 256, 1.0), (256, 1024, 1e6), (1024, 512, 1.0),
                                       (4096, 64, 1e6), (512, 8192, 1.0), (64, 32768, 1.0),
                                       (32768, 64, 1e6), (256, 4096, 1e6), (128, 16384, 1.0),
                                       (32768, 256, 1.0),
                                       # lengths that are not powers of two (general kernels)
                                       (8, 14, 1.0), (100, 134, 1e6), (67, 256, 1.0), (3000, 64, 1e6),
                                       (64, 3000, 1.0), (1000, 1000, 1e6), (6000, 16, 1.0), (4096, 30, 1e6),
                                       # one LDS buffer, in place: 12000 = 2^5 3 5^3, and 8191 (prime)
                                       # through a 16384-point Bluestein FFT; 4093 through 8192 points
                                       (12000, 16, 1e6), (8, 2 * 12000, 1.0), (8191, 8, 1.0),
                                       (16, 2 * 4093, 1e6),
                                       # beyond one workgroup: through an FP32 workspace (two levels,
                                       # Bluestein over two levels for the prime 8209, powers of two
                                       # beyond the power-of-two kernels)
                                       (20000, 16, 1e6), (16, 2 * 20000, 1.0), (8209, 8, 1.0),
                                       (4, 2 * 8209, 1e6), (1 << 17, 4, 1.0), (4, 1 << 17, 1.0)])
def test_16bit_storage(gpu, dt_name, bound, nx, ny, amp):
    import torch
    import hpxfft_amd
    print(f"{dt_name} {nx}x{ny} amp {amp}", flush=True)
    dt = torch.float16 if dt_name == "f16" else torch.bfloat16
    x = oracle.fill(nx, ny, seed=nx + ny, dist="normal") * amp
    ref = oracle.r2c_2d(x, ny)
    t, e_dev, _ = to16(x, ny, dt, scale_rows=amp != 1.0)
    X16, col_exp = hpxfft_amd.rfft2_scaled(t, e_dev if amp != 1.0 else None)
    torch.cuda.synchronize()
    X = X16.float().cpu().numpy().astype(np.float64)
    ce = col_exp.cpu().numpy()
    assert np.all(np.isfinite(X))
    Xc = (X[..., 0] + 1j * X[..., 1]) * np.exp2(-ce.astype(np.float64))[None, :]
    err = oracle.rel_l2(Xc[:, : ny // 2 + 1], ref)
    print(f"  rel-L2 vs FP64 oracle = {err:.3e}", flush=True)
    assert err <= bound, err


@pytest.mark.parametrize("dt_name", ["f16", "bf16"])
def test_16bit_beyond_one_workgroup_plan(gpu, dt_name):
    """16-bit plans with a length beyond one workgroup say how they run: the
    FP32 plan on a workspace, its own two-level / Bluestein lines indented."""
    import hpxfft_amd
    p = hpxfft_amd.Plan(20000, 16, precision=dt_name)
    d = p.describe()
    p.close()
    assert "FP32 workspace" in d and "precision f32" in d and "two levels" in d, d


def test_16bit_row_exponents_are_validated(gpu):
    """The row exponents are read as Nx int32 values on x16's device: any
    other tensor is refused before a kernel runs."""
    import torch
    import hpxfft_amd
    x = torch.zeros(64, 130, dtype=torch.float16, device="cuda")
    good = torch.zeros(64, dtype=torch.int32, device="cuda")
    for bad, err in [(good.cpu(), ValueError), (good.long(), TypeError), (good[:63], ValueError),
                     (torch.zeros(2, 64, dtype=torch.int32, device="cuda"), ValueError)]:
        with pytest.raises(err):
            hpxfft_amd.rfft2_scaled(x, bad)
    with pytest.raises(ValueError):
        hpxfft_amd.rfft2_scaled(x.cpu(), None)
    X16, ce = hpxfft_amd.rfft2_scaled(x, good[::1])
    torch.cuda.synchronize()
    assert X16.shape == (64, 65, 2) and not bool(X16.float().abs().sum())


@pytest.mark.slow
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_headline_grid_vs_oracle(gpu, prec, tol):
    import torch
    import hpxfft_amd
    n = 1 << 14
    print(f"{prec} {n}x{n} vs oracle", flush=True)
    x = oracle.fill(n, n, seed=2, dist="upm1")
    dt = torch.float64 if prec == "f64" else torch.float32
    d = torch.from_numpy(x).to(device="cuda", dtype=dt)
    p = hpxfft_amd.Plan(n, n, precision=prec)
    p.exec(d.data_ptr(), d.data_ptr())
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    del d
    oracle.Loop(n, n).exec(x)
    ref = x.view(np.complex128)
    if prec == "f64":
        err = oracle.rel_l2(got.view(np.complex128), ref)
    else:
        err = oracle.rel_l2(got.astype(np.float64).view(np.complex128), ref)
    print(f"  rel-L2 = {err:.3e}", flush=True)
    p.close()
    assert err <= tol


def test_cpp_known_answer(gpu):
    r = subprocess.run([os.path.join(BIN, "test_shared_loop_gpu")], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout


def test_cpp_harness(gpu, tmp_path):
    exe = os.path.join(BIN, "shared_loop_2d_gpu")
    r = subprocess.run([exe, "--nx=256", "--ny=16384", "--plan=estimate", "--header=1"], capture_output=True,
                       text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FFT 2D runtime" in r.stdout
    rows = (tmp_path / "runtimes" / "runtimes_hpx_shared_loop.txt").read_text().splitlines()
    assert rows[0].split(";")[:3] == ["n_threads", "n_x", "n_y"] and len(rows[1].split(";")) == 15
    assert "GPU r2c 2D plan" in (tmp_path / "plans" / "plan_hpx_shared_loop.txt").read_text()
    r = subprocess.run([exe, "--nx", "4", "--ny", "4", "--result", "1", "--run", "seq"], capture_output=True,
                       text=True, timeout=120, cwd=tmp_path)
    # ramp input x[i][j] = j on 4 x 4: X[0] = (24,0) (-8,8) (-8,0), other rows 0
    assert r.returncode == 0 and "(24 0) (-8 8) (-8 0)" in r.stdout, r.stdout


def test_size_sweep_with_harness(gpu, tmp_path):
    """The size sweep script over 2^3..2^6 grids, two runs each, on the real
    harness: one CSV header (--header=true) then eight rows, sizes in order."""
    sweep = os.path.join(REPO, "hpx-fft_amd", "cpp", "run_size_sweep.sh")
    exe = os.path.join(BIN, "shared_loop_2d_gpu")
    r = subprocess.run(["bash", sweep, exe, "estimate", "1", "3", "6", "2"], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = (tmp_path / "runtimes" / "runtimes_hpx_shared_loop.txt").read_text().splitlines()
    assert rows[0].startswith("n_threads;n_x;n_y;") and len(rows) == 9
    assert [int(x.split(";")[1]) for x in rows[1:]] == [8, 8, 16, 16, 32, 32, 64, 64]
    assert all(float(x.split(";")[7]) > 0 for x in rows[1:])   # fft_2d_total
