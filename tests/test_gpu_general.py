"""Lengths that are not powers of two (the general mixed-radix kernels) on the GPU.

The reference accepts any even Ny and any Nx through FFTW; the general
kernels cover lengths with prime factors <= 61 that fit one workgroup's LDS,
and other lengths by Bluestein's algorithm (a power-of-two FFT of at least
2n-1 points in LDS); longer lengths (and powers of two beyond the
power-of-two kernels) as two levels N = P*Q of such lengths.
Checks against the oracle (itself pinned to brute force and numpy for these
lengths): FP64 <= 1e-12, FP32 <= 1e-5 relative L2, with general rows and/or
general columns combined with every power-of-two kernel family; batches;
the reference harness at its default 8 x 14; and
the backward forms (adjoint, inverse c2r, autograd) on these lengths.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import REPO

pytestmark = pytest.mark.gpu
BIN = os.path.join(REPO, "hpx-fft_amd", "bin")

SIZES = [
    (8, 14),        # the reference harness default
    (3, 10), (5, 6), (7, 30), (1, 6), (4095, 2), (12, 100), (96, 96),
    (1000, 64),     # general columns, power-of-two rows
    (64, 1000),     # general rows, direct power-of-two columns
    (3000, 8), (6, 8190),
    (4096, 30),     # general rows, four-step columns
    (10, 4096),     # general columns, long power-of-two rows
    # prime factors above 61: Bluestein's algorithm
    (8, 134), (67, 16), (67, 134), (127, 262), (1031, 16),
    (4096, 134),    # Bluestein rows, four-step columns
    (67, 4096),     # Bluestein columns, long power-of-two rows
    (3, 2 * 2039),  # Bluestein rows through a 4096-point FFT (the FP64 ping-pong limit)
    (4093, 8), (6, 2 * 4093),   # FP64: an 8192-point Bluestein FFT, in place only
]


def _run(x_pad, nx, ny, prec, batch=1):
    import torch
    import hpxfft_amd
    dt = torch.float64 if prec == "f64" else torch.float32
    d = torch.from_numpy(x_pad).to(device="cuda", dtype=dt).contiguous()
    p = hpxfft_amd.Plan(nx, ny, batch=batch, precision=prec)
    p.exec(d.data_ptr(), d.data_ptr())
    torch.cuda.synchronize()
    p.close()
    return d.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_general_parity(gpu, nx, ny, prec, tol):
    print(f"general {prec} {nx}x{ny}", flush=True)
    x = oracle.fill(nx, ny, seed=5000 + nx + ny, dist="upm1")
    ref = oracle.r2c_2d(x, ny)
    got = _run(x, nx, ny, prec)[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
    err = oracle.rel_l2(got, ref)
    print(f"  rel-L2 = {err:.3e}", flush=True)
    assert err <= tol, err


def test_general_fp32_beyond_fp64_lds(gpu):
    """FP32 holds twice the points per workgroup: Nx = 12000 = 2^5 3 5^3 runs
    one-level in place in FP32, where FP64 takes two levels."""
    nx, ny = 12000, 8
    x = oracle.fill(nx, ny, seed=11, dist="normal")
    got = _run(x, nx, ny, "f32")[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
    assert oracle.rel_l2(got, oracle.r2c_2d(x, ny)) <= 1e-5


# lengths beyond one workgroup's pair of buffers: one buffer in place when the
# radices are at most 7 and the length fits it (6000, 4500 in FP64), else two
# levels N = P*Q (phase A over m1 + P*e with twiddle w_N^{m1 k2}, phase B over
# m1), rows then split in place
SIZES2 = [
    (6000, 16),       # columns 6000: in place (FP64 and FP32)
    (16, 12000),      # rows Ny/2 = 6000
    (3000, 9000),     # direct columns, rows 4500
    (18000, 16),      # columns 18000 = 120 * 150: two levels in both precisions
    (16, 36000),      # rows Ny/2 = 18000
    (300, 34000),     # rows 17000 = 2^3 5^3 17: two levels
    (6, 2 * 9801),    # rows 99 * 99 (P, Q odd; the r2c split fused into phase B pairs k2 with Q - k2)
    (4, 2 * 10000),   # rows 100 * 100 (P, Q even) in FP64
    (5, 2 * 8575),    # rows 49 * 175 in FP64
    (6117, 6),        # 3 * 2039: a Bluestein phase inside two levels
    (1 << 17, 4),     # powers of two beyond the power-of-two kernels
    (4, 1 << 17),
]


@pytest.mark.parametrize("nx,ny", SIZES2)
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_two_level_parity(gpu, nx, ny, prec, tol):
    import hpxfft_amd
    print(f"two-level {prec} {nx}x{ny}", flush=True)
    x = oracle.fill(nx, ny, seed=nx + 5 * ny, dist="upm1")
    ref = oracle.r2c_2d(x, ny)
    got = _run(x, nx, ny, prec)[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
    err = oracle.rel_l2(got, ref)
    print(f"  rel-L2 = {err:.3e}", flush=True)
    assert err <= tol, err
    cap = 4096 if prec == "f64" else 8192     # points of one workgroup's buffer

    def one_level(n):   # direct (prime factors <= 61) or Bluestein within one buffer
        m, f, big = n, 2, 1
        while f * f <= m:
            while m % f == 0:
                big, m = max(big, f), m // f
            f += 1
        big = max(big, m)
        return (n <= cap and big <= 61) or 1 << (2 * n - 2).bit_length() <= cap

    def two(n, pow2_max):   # (every length here splits)
        m, f, big = n, 2, 1
        while f * f <= m:
            while m % f == 0:
                big, m = max(big, f), m // f
            f += 1
        in_place = n <= 2 * cap and max(big, m) <= 7
        return not (n & (n - 1) == 0 and n <= pow2_max) and not one_level(n) and not in_place

    p = hpxfft_amd.Plan(nx, ny, precision=prec)
    assert ("two levels" in p.describe()) == (two(nx, 1 << 16) or two(ny // 2, 1 << 14))
    p.close()


@pytest.mark.parametrize("nx,ny,prec,tol", [(6144, 64, "f32", 1e-5), (3000, 32, "f64", 1e-12),
                                             (2500, 20, "f64", 1e-12)])
def test_measure_times_one_and_two_level_columns(gpu, nx, ny, prec, tol):
    """Narrow general column strips (one 16- or 32-byte segment per row): the
    measure effort times the one-level pass against the two-level form and
    keeps the faster; either way the result is the oracle's."""
    import torch
    import hpxfft_amd
    x = oracle.fill(nx, ny, seed=nx + ny, dist="upm1")
    dt = torch.float64 if prec == "f64" else torch.float32
    d = torch.from_numpy(x).to(device="cuda", dtype=dt)
    p = hpxfft_amd.Plan(nx, ny, precision=prec, effort="measure")
    desc = p.describe()
    assert "general column pass autotuned" in desc, desc
    p.exec(d.data_ptr(), d.data_ptr())
    torch.cuda.synchronize()
    p.close()
    got = d.cpu().numpy().astype(np.float64)[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
    assert oracle.rel_l2(got, oracle.r2c_2d(x, ny)) <= tol


@pytest.mark.parametrize("nx,ny,prec,tol", [(3000, 3000, "f64", 1e-12), (1500, 6000, "f32", 1e-5),
                                             (840, 6300, "f64", 1e-12), (1000, 1000, "f64", 1e-12),
                                             (343, 2 * 2401, "f64", 1e-12),
                                             # Bluestein rows and columns (primes 1031, 127, 67)
                                             (1031, 2 * 127, "f64", 1e-12), (67, 2 * 1031, "f32", 1e-5),
                                             (127, 4096, "f64", 1e-12)])
def test_in_place_and_ping_pong_stages(gpu, nx, ny, prec, tol):
    """Direct general plans whose radices are all at most 7 run ping-pong
    Stockham stages or in-place decimation-in-time stages (inputs scattered
    to their digit-reversed positions); Bluestein plans run their
    power-of-two FFTs ping-pong or in place (decimation in frequency, the
    kernel spectrum in digit-reversed order, decimation in time).  The
    estimate runs in place; measure times both forms of each one-level pass
    and keeps the faster.  Either way the result is the oracle's."""
    import torch
    import hpxfft_amd
    x = oracle.fill(nx, ny, seed=7 * nx + ny, dist="upm1")
    ref = oracle.r2c_2d(x, ny)
    dt = torch.float64 if prec == "f64" else torch.float32
    for effort in ("estimate", "measure"):
        print(f"stages {prec} {nx}x{ny} {effort}", flush=True)
        d = torch.from_numpy(x).to(device="cuda", dtype=dt)
        p = hpxfft_amd.Plan(nx, ny, precision=prec, effort=effort)
        desc = p.describe()
        general = [ny // 2 != 1 << (ny // 2).bit_length() - 1, nx != 1 << nx.bit_length() - 1]
        if effort == "measure":
            assert ("row pass autotuned (ms): ping-pong" in desc) == general[0], desc
            assert ("column pass autotuned (ms): ping-pong" in desc) == general[1], desc
        else:
            assert "mixed-radix Stockham" not in desc and ", Stockham)" not in desc, desc
            assert desc.count("in-place decimation") + desc.count("in place: decimation") == sum(general), desc
        p.exec(d.data_ptr(), d.data_ptr())
        torch.cuda.synchronize()
        p.close()
        got = d.cpu().numpy().astype(np.float64)[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
        err = oracle.rel_l2(got, ref)
        print(f"  rel-L2 = {err:.3e}", flush=True)
        assert err <= tol, err


@pytest.mark.parametrize("nx,ny", [(8191, 4), (4, 2 * 8191)])
def test_fp32_in_place_only_bluestein(gpu, nx, ny):
    """8191 (prime) in FP32: a 16384-point Bluestein FFT, which only the
    in-place stages hold in one workgroup's LDS; both efforts (measure has no
    ping-pong form to time against)."""
    import torch
    import hpxfft_amd
    x = oracle.fill(nx, ny, seed=nx + 3 * ny, dist="upm1")
    ref = oracle.r2c_2d(x, ny)
    for effort in ("estimate", "measure"):
        d = torch.from_numpy(x).to(device="cuda", dtype=torch.float32)
        p = hpxfft_amd.Plan(nx, ny, precision="f32", effort=effort)
        desc = p.describe()
        assert "16384-point FFT" in desc and "in place: decimation" in desc, desc
        assert "autotuned (ms): ping-pong" not in desc, desc
        p.exec(d.data_ptr(), d.data_ptr())
        torch.cuda.synchronize()
        p.close()
        got = d.cpu().numpy().astype(np.float64)[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
        err = oracle.rel_l2(got, ref)
        print(f"f32 {nx}x{ny} {effort}: rel-L2 = {err:.3e}", flush=True)
        assert err <= 1e-5, err


@pytest.mark.parametrize("nx,ny,text", [
    (8, 2 * 2039, "4096-point FFT in LDS, radices 8x8x8x8"),   # 2^12: 4 radix-8 stages, not 6 radix-4
    (6000, 8, "6000-point mixed-radix in-place decimation in time in LDS, radices 4x4x3x5x5x5"),  # 2^4: 4s
    (3000, 8, "radices 8x3x5x5x5"),                             # 2^3: one radix-8 stage
    (1500, 8, "radices 4x3x5x5x5")])                            # 2^2: one radix-4 stage
def test_radix_choice_in_plan_text(gpu, nx, ny, text):
    """The power-of-two part of a general length takes radix-8 stages only
    where they remove a stage (the plan text lists the radices)."""
    import hpxfft_amd
    p = hpxfft_amd.Plan(nx, ny, precision="f64")
    desc = p.describe()
    p.close()
    assert text in desc, desc


def test_two_level_fp32_bluestein_phase(gpu):
    """FP32 Bluestein FFTs of up to 8192 points: 8186 = 2 * 4093 (prime) as a
    2-point phase and a 4093-point Bluestein phase."""
    nx, ny = 2 * 4093, 8
    x = oracle.fill(nx, ny, seed=21, dist="normal")
    got = _run(x, nx, ny, "f32")[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
    assert oracle.rel_l2(got, oracle.r2c_2d(x, ny)) <= 1e-5


@pytest.mark.parametrize("nx,ny", [(6000, 2 * 4500),    # one buffer in place, both dimensions
                                   (18000, 20), (12, 2 * 18000)])   # two-level columns / rows
def test_two_level_batched(gpu, nx, ny):
    batch = 2
    ld = 2 * (ny // 2 + 1)
    xs = [oracle.fill(nx, ny, seed=600 + b, dist="upm1") for b in range(batch)]
    got = _run(np.concatenate(xs), nx, ny, "f64", batch=batch)
    for b, x in enumerate(xs):
        g = got[b * nx:(b + 1) * nx, :ld].copy().view(np.complex128)
        assert oracle.rel_l2(g, oracle.r2c_2d(x, ny)) <= 1e-12, b


@pytest.mark.parametrize("nx,ny", [(6000, 16), (16, 12000), (3000, 9000), (6117, 6), (18000, 16), (16, 36000)])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_two_level_adjoint_and_c2r(gpu, nx, ny, prec, tol):
    import torch
    import hpxfft_amd
    print(f"two-level backward {prec} {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx + ny)
    Y = rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    Yd = torch.from_numpy(Y).to(device="cuda", dtype=cdt)
    Yr = Yd.cpu().numpy().astype(np.complex128)
    adj = hpxfft_amd.rfft2_adjoint(Yd, ny).cpu().numpy().astype(np.float64)
    assert oracle.rel_l2(adj, oracle.adjoint_r2c_2d(Yr, ny)) <= tol
    inv = hpxfft_amd.c2r_2d(Yd, ny).cpu().numpy().astype(np.float64)
    assert oracle.rel_l2(inv, oracle.c2r_2d(Yr, ny)) <= tol
    hpxfft_amd.clear_plan_cache()


def test_bluestein_fp32_beyond_fp64_lds(gpu):
    """FP32 runs Bluestein FFTs of up to 8192 points: Ny/2 = 4093 (prime)."""
    nx, ny = 4, 2 * 4093
    x = oracle.fill(nx, ny, seed=12, dist="normal")
    got = _run(x, nx, ny, "f32")[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
    assert oracle.rel_l2(got, oracle.r2c_2d(x, ny)) <= 1e-5


def test_general_batched(gpu):
    nx, ny, batch = 30, 42, 3
    ld = 2 * (ny // 2 + 1)
    xs = [oracle.fill(nx, ny, seed=300 + b, dist="upm1") for b in range(batch)]
    got = _run(np.concatenate(xs), nx, ny, "f64", batch=batch)
    for b, x in enumerate(xs):
        g = got[b * nx:(b + 1) * nx, :ld].copy().view(np.complex128)
        assert oracle.rel_l2(g, oracle.r2c_2d(x, ny)) <= 1e-12, b


BACK_SIZES = [(8, 14), (5, 6), (1, 6), (12, 100), (1000, 64), (64, 1000), (4096, 30), (10, 4096), (3, 2),
              (67, 134), (127, 262), (4096, 134), (2053, 4), (4, 2 * 2053)]   # 2053: in place only in FP64


@pytest.mark.parametrize("nx,ny", BACK_SIZES)
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_general_adjoint_and_c2r(gpu, nx, ny, prec, tol):
    """The backward forms on general lengths: adjoint (interior halved) and
    the unnormalised inverse c2r, against the oracle's FP64 versions."""
    import torch
    import hpxfft_amd
    print(f"general backward {prec} {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx * 7 + ny)
    Y = rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    Yd = torch.from_numpy(Y).to(device="cuda", dtype=cdt)
    Yr = Yd.cpu().numpy().astype(np.complex128)   # the same (rounded) input
    adj = hpxfft_amd.rfft2_adjoint(Yd, ny).cpu().numpy().astype(np.float64)
    assert oracle.rel_l2(adj, oracle.adjoint_r2c_2d(Yr, ny)) <= tol
    inv = hpxfft_amd.c2r_2d(Yd, ny).cpu().numpy().astype(np.float64)
    assert oracle.rel_l2(inv, oracle.c2r_2d(Yr, ny)) <= tol


def test_general_autograd_and_round_trip(gpu):
    """rfft2_autograd on 8 x 14 against torch.fft FP64 autograd on the CPU, and
    c2r(r2c(x)) = Nx Ny x."""
    import torch
    import hpxfft_amd
    nx, ny = 8, 14
    xc = torch.randn(nx, ny, dtype=torch.float64)
    w = torch.rand(nx, ny // 2 + 1, dtype=torch.float64)
    xr = xc.clone().requires_grad_(True)
    torch.sum(w * torch.abs(torch.fft.rfft2(xr)) ** 2).backward()
    xg = xc.cuda().requires_grad_(True)
    torch.sum(w.cuda() * torch.abs(hpxfft_amd.rfft2_autograd(xg)) ** 2).backward()
    assert oracle.rel_l2(xg.grad.cpu().numpy(), xr.grad.numpy()) <= 1e-12
    back = hpxfft_amd.c2r_2d(hpxfft_amd.rfft2(xc.cuda()), ny).cpu()
    assert torch.allclose(back / (nx * ny), xc, rtol=0, atol=1e-13)


def test_cpp_harness_default_size(gpu, tmp_path):
    """No arguments: the reference's default 8 x 14 ramp, X[0][0] = Nx Ny (Ny-1)/2 = 728."""
    r = subprocess.run([os.path.join(BIN, "shared_loop_2d_gpu"), "--result", "1"], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "(728 0)" in r.stdout, r.stdout
