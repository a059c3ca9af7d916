"""Inverse 2-D c2r (unnormalised, FFTW_BACKWARD) on the GPU through the C ABI.

* against the oracle's c2r (itself pinned to numpy.fft.irfft2 on the CPU) on
  arbitrary, non-Hermitian half spectra: FP64 <= 1e-12, FP32 <= 1e-5
  relative L2, over shapes that reach the direct and four-step column
  kernels, short and long rows, and a batch through the torch entry;
* the round trip c2r(r2c(x)) = Nx*Ny*x at the 2^14 x 2^14 FP64 size.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SIZES = [(1, 2), (2, 2), (1, 8), (8, 16), (64, 64), (1024, 512), (256, 4096), (4096, 64), (16384, 64),
         (32768, 256), (64, 32768)]


def spectrum(nx, ny, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_c2r_vs_oracle(gpu, nx, ny, prec, tol):
    import torch
    import hpxfft_amd
    print(f"c2r {prec} {nx}x{ny}", flush=True)
    X = spectrum(nx, ny, nx + 7 * ny)
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    Xd = torch.from_numpy(X).to(device="cuda", dtype=cdt)
    ref = oracle.c2r_2d(Xd.cpu().numpy().astype(np.complex128), ny)   # the same (rounded) input
    x = hpxfft_amd.c2r_2d(Xd, ny)
    torch.cuda.synchronize()
    err = oracle.rel_l2(x.cpu().numpy().astype(np.float64), ref)
    print(f"  rel-L2 = {err:.3e}", flush=True)
    assert err <= tol, err


def test_c2r_batched_through_torch(gpu):
    import torch
    import hpxfft_amd
    nx, ny, batch = 128, 256, 3
    Xs = np.stack([spectrum(nx, ny, 100 + b) for b in range(batch)])
    x = hpxfft_amd.c2r_2d(torch.from_numpy(Xs).cuda(), ny).cpu().numpy()
    for b in range(batch):
        assert oracle.rel_l2(x[b], oracle.c2r_2d(Xs[b], ny)) <= 1e-12


@pytest.mark.slow
def test_round_trip_headline_size(gpu):
    import torch
    import hpxfft_amd
    n = 1 << 14
    print(f"c2r(r2c(x)) f64 {n}x{n}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(2)
    x = torch.empty((n, n), dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0, generator=g)
    back = hpxfft_amd.c2r_2d(hpxfft_amd.rfft2(x), n)
    err = (torch.linalg.vector_norm(back / (n * n) - x) / torch.linalg.vector_norm(x)).item()
    print(f"  rel-L2 = {err:.3e}", flush=True)
    hpxfft_amd.clear_plan_cache()
    assert err <= 1e-12, err
