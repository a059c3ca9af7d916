"""Seeded random grid shapes on the GPU through the C ABI, against the oracle.

Forty shapes (Nx * Ny <= 1.5e6) mixing powers of two with arbitrary lengths
up to 3000 (prime lengths among them), and ten with one side of 4097 to
20000 points, so the power-of-two kernels, the general one-level kernels,
Bluestein's algorithm and the two-level path all meet.  For each shape the
plan must either be created, and then match the oracle (FP64 <= 1e-12,
FP32 <= 1e-5 relative L2, forward; the adjoint on every fourth shape), or
fail with UNSUPPORTED_SIZE exactly when no plan covers a length (none of
these sides: lengths up to 2^22 always have a plan, by Bluestein's
algorithm over two levels when no split N = P*Q and no one-workgroup
Bluestein FFT covers them).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _shapes():
    rng = np.random.default_rng(2026)
    out = []
    while len(out) < 40:
        kind = rng.integers(0, 3)
        nx = int(rng.integers(1, 3000)) if kind else int(2 ** rng.integers(0, 12))
        ny = 2 * int(rng.integers(1, 3000)) if kind != 1 else int(2 ** rng.integers(1, 13))
        if nx * ny <= 1_500_000:
            out.append((nx, ny))
    return out


def _long_shapes():
    """Ten shapes with one side beyond one workgroup (4097..20000)."""
    rng = np.random.default_rng(77)
    out = []
    for i in range(10):
        long_side = int(rng.integers(4097, 20001))
        short = int(rng.integers(1, 64))
        out.append((long_side, 2 * short) if i % 2 == 0 else (short, 2 * long_side))
    return out


SHAPES = _shapes() + _long_shapes()


def _largest_prime(n):
    m, f, big = n, 2, 1
    while f * f <= m:
        while m % f == 0:
            big, m = max(big, f), m // f
        f += 1
    return max(big, m)


def _bluestein_fft(n):
    return 1 << (2 * n - 2).bit_length()     # the power of two >= 2n - 1


def _one_level(n, cap):
    """Within one workgroup's pair of buffers (cap points)."""
    return (n <= cap and _largest_prime(n) <= 61) or _bluestein_fft(n) <= cap


def _supported(n, cap, pow2_max):
    if n & (n - 1) == 0 and n <= pow2_max:
        return True
    if _one_level(n, cap) or (n <= 2 * cap and _largest_prime(n) <= 7):   # in place, one buffer
        return True
    if n <= 1 << 24 and any(n % p == 0 and n // p >= 2 and _one_level(p, cap) and _one_level(n // p, cap)
                            for p in range(2, min(cap, n // 2) + 1)):
        return True
    # in place, one buffer of 2 cap points; else Bluestein's algorithm over two levels
    return _bluestein_fft(n) <= 2 * cap or n <= 1 << 22


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_random_shape(gpu, nx, ny, prec, tol):
    import torch
    import hpxfft_amd
    from hpxfft_amd import _lib
    cap = 4096 if prec == "f64" else 8192
    expect = _supported(nx, cap, 1 << 16) and _supported(ny // 2, cap, 1 << 14)
    print(f"random {prec} {nx}x{ny} supported={expect}", flush=True)
    if not expect:
        with pytest.raises(RuntimeError) as e:
            hpxfft_amd.Plan(nx, ny, precision=prec)
        assert e.value.status == _lib.UNSUPPORTED_SIZE
        return
    x = oracle.fill(nx, ny, seed=nx * 7 + ny, dist="upm1")
    dt = torch.float64 if prec == "f64" else torch.float32
    d = torch.from_numpy(x).to(device="cuda", dtype=dt).contiguous()
    p = hpxfft_amd.Plan(nx, ny, precision=prec)
    p.exec(d.data_ptr(), d.data_ptr())
    torch.cuda.synchronize()
    p.close()
    got = d.cpu().numpy().astype(np.float64)[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
    err = oracle.rel_l2(got, oracle.r2c_2d(x, ny))
    print(f"  rel-L2 = {err:.3e}", flush=True)
    assert err <= tol, err
    if SHAPES.index((nx, ny)) % 4 == 0:
        rng = np.random.default_rng(nx + ny)
        Y = rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))
        Yd = torch.from_numpy(Y).to(device="cuda", dtype=torch.complex128 if prec == "f64" else torch.complex64)
        Yr = Yd.cpu().numpy().astype(np.complex128)
        adj = hpxfft_amd.rfft2_adjoint(Yd, ny).cpu().numpy().astype(np.float64)
        assert oracle.rel_l2(adj, oracle.adjoint_r2c_2d(Yr, ny)) <= tol
        hpxfft_amd.clear_plan_cache()
