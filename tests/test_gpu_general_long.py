"""Lengths beyond one workgroup and beyond any split into two one-workgroup
lengths, on the GPU.

* Bluestein's algorithm over two levels: a length N with a prime factor too
  large for in-LDS Bluestein and no usable split runs as a chirp convolution
  of length M (the power of two >= 2N - 1) in three passes M = P*Q.  FP64
  primes above 4096 (FP32: above 8192), and 2 x such a prime.
* Two levels beyond 2^20 points (twiddle tables split in two, multiplied in
  FP64): 1.5e6 = 2^5 3 5^6 and 2^21 / 2^22, as columns and as rows.
* Both together: the prime 1048583 = 2^20 + 7 (M = 2^22).

Checks: against the oracle (FP64 <= 1e-12, FP32 <= 1e-5 relative L2) where
its mixed-radix FFT finishes in seconds; for the prime above 2^20 (O(N^2) in
the oracle) against numpy's pocketfft rfft2 (an independent FFT library) and,
independently of it, against 16 output bins summed directly from the input.
The backward forms (adjoint, inverse c2r) on two-level Bluestein lengths
against the oracle.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _run(x_pad, nx, ny, prec, batch=1):
    import torch
    import hpxfft_amd
    dt = torch.float64 if prec == "f64" else torch.float32
    d = torch.from_numpy(x_pad).to(device="cuda", dtype=dt).contiguous()
    p = hpxfft_amd.Plan(nx, ny, batch=batch, precision=prec)
    desc = p.describe()
    p.exec(d.data_ptr(), d.data_ptr())
    torch.cuda.synchronize()
    p.close()
    out = d.cpu().numpy().astype(np.float64)
    del d
    return out, desc


BLUE = [
    (4099, 6, "f64"),          # columns, prime 4099: M = 16384 = 128 * 128
    (3, 2 * 4099, "f64"),      # rows Ny/2 = 4099
    (2 * 4111, 4, "f64"),      # 2 x prime: no split has both lengths within one workgroup
    (8209, 4, "f64"),
    (8209, 4, "f32"),          # FP32 in-LDS Bluestein ends at 8192 points
    (2, 2 * 8209, "f32"),
    (2, 2 * 8209, "f64"),
]


@pytest.mark.parametrize("nx,ny,prec", BLUE)
def test_two_level_bluestein_parity(gpu, nx, ny, prec):
    tol = 1e-12 if prec == "f64" else 1e-5
    print(f"two-level Bluestein {prec} {nx}x{ny}", flush=True)
    x = oracle.fill(nx, ny, seed=nx + 3 * ny, dist="upm1")
    ref = oracle.r2c_2d(x, ny)
    got, desc = _run(x, nx, ny, prec)
    err = oracle.rel_l2(got[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128), ref)
    print(f"  rel-L2 = {err:.3e}", flush=True)
    assert "Bluestein over two levels" in desc, desc
    assert err <= tol, err


def test_two_level_bluestein_batched(gpu):
    nx, ny, batch = 4099, 6, 2
    ld = 2 * (ny // 2 + 1)
    xs = [oracle.fill(nx, ny, seed=900 + b, dist="normal") for b in range(batch)]
    got, _ = _run(np.concatenate(xs), nx, ny, "f64", batch=batch)
    for b, x in enumerate(xs):
        g = got[b * nx:(b + 1) * nx, :ld].copy().view(np.complex128)
        assert oracle.rel_l2(g, oracle.r2c_2d(x, ny)) <= 1e-12, b


@pytest.mark.parametrize("nx,ny,prec", [(4099, 6, "f64"), (3, 2 * 4099, "f64"), (8209, 4, "f32"),
                                        (2, 2 * 8209, "f32")])
def test_two_level_bluestein_adjoint_and_c2r(gpu, nx, ny, prec):
    import torch
    import hpxfft_amd
    tol = 1e-12 if prec == "f64" else 1e-5
    print(f"two-level Bluestein backward {prec} {nx}x{ny}", flush=True)
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


LONG = [
    (1500000, 4),        # columns 1.5e6 = 2^5 3 5^6: two levels, split twiddle table
    (2, 2 * 1500000),    # rows of 1.5e6 complex points
    (1 << 21, 4),        # power-of-two columns beyond 2^20
    (2, 1 << 22),        # rows of 2^21 complex points
]


@pytest.mark.parametrize("nx,ny", LONG)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_two_level_beyond_2e20(gpu, nx, ny, prec):
    tol = 1e-12 if prec == "f64" else 1e-5
    print(f"two-level long {prec} {nx}x{ny}", flush=True)
    x = oracle.fill(nx, ny, seed=nx + ny, dist="upm1")
    ref = oracle.r2c_2d(x, ny)
    got, desc = _run(x, nx, ny, prec)
    err = oracle.rel_l2(got[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128), ref)
    print(f"  rel-L2 = {err:.3e}", flush=True)
    assert "two levels" in desc and "split in two" in desc, desc
    assert err <= tol, err


def _direct_bins(x, nx, ny, bins):
    """X[kx][ky] summed directly (FP64, exact twiddles by integer reduction)."""
    rows = x[:, :ny]
    out = []
    m = np.arange(nx, dtype=np.int64)
    n = np.arange(ny, dtype=np.int64)
    for kx, ky in bins:
        wy = np.exp(-2j * np.pi * ((n * ky) % ny) / ny)
        wx = np.exp(-2j * np.pi * ((m * kx) % nx) / nx)
        out.append(np.dot(wx, rows @ wy))
    return np.array(out)


@pytest.mark.parametrize("nx,ny,prec", [(1048583, 4, "f64"), (1048583, 4, "f32"), (2, 2 * 1048583, "f64")])
def test_two_level_bluestein_prime_beyond_2e20(gpu, nx, ny, prec):
    """The prime 2^20 + 7: M = 2^22.  numpy's rfft2 (pocketfft) as the
    reference, and 16 bins summed directly from the input."""
    tol = 1e-12 if prec == "f64" else 1e-5
    print(f"two-level Bluestein prime {prec} {nx}x{ny}", flush=True)
    x = oracle.fill(nx, ny, seed=77, dist="upm1")
    got, desc = _run(x, nx, ny, prec)
    assert "Bluestein over two levels" in desc, desc
    X = got[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
    ref = np.fft.rfft2(x[:, :ny])
    err = oracle.rel_l2(X, ref)
    print(f"  rel-L2 vs pocketfft = {err:.3e}", flush=True)
    assert err <= tol, err
    rng = np.random.default_rng(5)
    bins = [(int(rng.integers(nx)), int(rng.integers(ny // 2 + 1))) for _ in range(15)] + [(0, 0)]
    d = _direct_bins(x, nx, ny, bins)
    g = np.array([X[kx, ky] for kx, ky in bins])
    berr = np.linalg.norm(g - d) / np.linalg.norm(d)
    print(f"  16 direct bins: rel err {berr:.3e}", flush=True)
    assert berr <= (1e-9 if prec == "f64" else 1e-4), berr
