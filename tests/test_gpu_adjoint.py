"""Adjoint kernels and the torch.autograd.Function, on the GPU through the C ABI.

R x = 2-D r2c of real x (half spectrum); R* its adjoint under Re<A,B>.
Checks:
  * R* Y against the oracle's FP64 adjoint (itself pinned by brute force and
    torch autograd fixtures): FP64 <= 1e-12, FP32 <= 1e-6 relative L2;
  * <R x, Y> = <x, R* Y> at sizes that exercise every kernel family;
  * full Jacobian by central finite differences (FP64) on small grids;
  * loss L = sum_half w(k)|X_k|^2, w = 1/(1+|k|^2): the RFFT2 autograd gradient
    against torch.fft.rfft2 FP64 autograd on the CPU and against FP64 central
    differences (exact for a quadratic up to rounding) along the gradient
    direction and random directions, tolerance 1e-6 (see directional_check
    for the normalisation), including the batch 16 x 4096 x 4096 FP32 shape.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def weights(nx, ny, device, dtype):
    import torch
    kx = torch.arange(nx, device=device, dtype=dtype)
    kx = torch.where(kx >= nx // 2, kx - nx, kx)   # signed frequency
    ky = torch.arange(ny // 2 + 1, device=device, dtype=dtype)
    return 1.0 / (1.0 + kx[:, None] ** 2 + ky[None, :] ** 2)


SIZES = [(1, 2), (2, 4), (4, 4), (8, 16), (16, 8), (64, 128), (256, 512), (1024, 64), (2048, 32), (32, 4096),
         (4096, 16), (16, 16384), (8, 32768), (16384, 8)]


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_adjoint_vs_oracle(gpu, nx, ny, prec):
    import torch
    import hpxfft_amd
    print(f"adjoint {prec} {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx * 7 + ny)
    Y = rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))
    ref = oracle.adjoint_r2c_2d(Y, ny)
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    got = hpxfft_amd.rfft2_adjoint(torch.from_numpy(Y).to(device=gpu, dtype=cdt), ny)
    torch.cuda.synchronize()
    err = oracle.rel_l2(got.cpu().double().numpy(), ref)
    assert err <= (1e-12 if prec == "f64" else 1e-6), err


@pytest.mark.parametrize("nx,ny,batch", [(128, 256, 3), (2048, 64, 2)])
def test_adjoint_batched(gpu, nx, ny, batch):
    """A batch through the torch entry (direct and four-step column passes):
    every grid against the oracle's FP64 adjoint."""
    import torch
    import hpxfft_amd
    rng = np.random.default_rng(nx + batch)
    Y = rng.standard_normal((batch, nx, ny // 2 + 1)) + 1j * rng.standard_normal((batch, nx, ny // 2 + 1))
    x = hpxfft_amd.rfft2_adjoint(torch.from_numpy(Y).cuda(), ny).cpu().numpy()
    for b in range(batch):
        assert oracle.rel_l2(x[b], oracle.adjoint_r2c_2d(Y[b], ny)) <= 1e-12, b


@pytest.mark.parametrize("nx,ny", [(8, 16), (512, 1024), (2048, 4096), (16384, 64)])
def test_inner_product_identity(gpu, nx, ny):
    import torch
    import hpxfft_amd
    g = torch.Generator(device=gpu).manual_seed(nx + ny)
    x = torch.randn(nx, ny, dtype=torch.float64, device=gpu, generator=g)
    Y = torch.randn(nx, ny // 2 + 1, dtype=torch.complex128, device=gpu, generator=g)
    lhs = torch.real(torch.vdot(hpxfft_amd.rfft2(x).flatten(), Y.flatten())).item()
    rhs = torch.dot(x.flatten(), hpxfft_amd.rfft2_adjoint(Y, ny).flatten()).item()
    scale = x.abs().sum().item() * Y.abs().max().item()
    assert abs(lhs - rhs) <= 1e-13 * scale


@pytest.mark.parametrize("nx,ny", [(2, 2), (4, 8), (8, 4), (16, 16)])
def test_full_jacobian_finite_differences(gpu, nx, ny):
    """(R* Y)[m,n] = d/dx[m,n] Re<R x, Y>, by central differences of the FP64 forward."""
    import torch
    import hpxfft_amd
    Y = torch.randn(nx, ny // 2 + 1, dtype=torch.complex128, device=gpu)
    adj = hpxfft_amd.rfft2_adjoint(Y, ny).cpu().numpy()
    x0 = torch.randn(nx, ny, dtype=torch.float64, device=gpu)
    fd = np.empty((nx, ny))
    for m in range(nx):
        for n in range(ny):
            e = torch.zeros_like(x0)
            e[m, n] = 1.0
            lp = torch.real(torch.vdot(hpxfft_amd.rfft2(x0 + e).flatten(), Y.flatten())).item()
            lm = torch.real(torch.vdot(hpxfft_amd.rfft2(x0 - e).flatten(), Y.flatten())).item()
            fd[m, n] = (lp - lm) / 2.0
    assert np.max(np.abs(fd - adj)) <= 1e-12 * max(1.0, np.abs(adj).max())


def loss_fn(X, w):
    import torch
    return torch.sum(w * (X.real ** 2 + X.imag ** 2))


@pytest.mark.parametrize("nx,ny", [(8, 16), (64, 64), (256, 4096)])
def test_autograd_vs_torch_fp64_cpu(gpu, nx, ny):
    import torch
    import hpxfft_amd
    xc = torch.randn(nx, ny, dtype=torch.float64)
    xr = xc.clone().requires_grad_(True)
    loss_fn(torch.fft.rfft2(xr), weights(nx, ny, "cpu", torch.float64)).backward()
    xg = xc.to(gpu).requires_grad_(True)
    loss_fn(hpxfft_amd.rfft2_autograd(xg), weights(nx, ny, gpu, torch.float64)).backward()
    err = oracle.rel_l2(xg.grad.cpu().numpy(), xr.grad.numpy())
    assert err <= 1e-12, err


@pytest.mark.parametrize("nx,ny", [(8, 16), (64, 128)])
def test_full_spectrum_loss_with_hermitian_weights(gpu, nx, ny):
    """The same loss written over the FULL spectrum, sum_k w(k)|X_k|^2 through
    torch.fft.fft2 (FP64, CPU autograd), equals the half-spectrum loss with
    Hermitian weights c(ky) = 1 at ky in {0, Ny/2} and 2 in between (w is even
    in k).  Both the value and the RFFT2 gradient must agree with it."""
    import torch
    import hpxfft_amd
    xc = torch.randn(nx, ny, dtype=torch.float64)
    kx = torch.fft.fftfreq(nx, 1.0 / nx, dtype=torch.float64)
    ky = torch.fft.fftfreq(ny, 1.0 / ny, dtype=torch.float64)
    w_full = 1.0 / (1.0 + kx[:, None] ** 2 + ky[None, :] ** 2)
    xr = xc.clone().requires_grad_(True)
    lf = torch.sum(w_full * torch.abs(torch.fft.fft2(xr)) ** 2)
    lf.backward()
    c = torch.full((ny // 2 + 1,), 2.0, dtype=torch.float64)
    c[0] = c[-1] = 1.0
    wh = weights(nx, ny, gpu, torch.float64) * c.to(gpu)[None, :]
    xg = xc.to(gpu).requires_grad_(True)
    lh = loss_fn(hpxfft_amd.rfft2_autograd(xg), wh)
    lh.backward()
    assert abs(lh.item() - lf.item()) <= 1e-12 * abs(lf.item())
    err = oracle.rel_l2(xg.grad.cpu().numpy(), xr.grad.numpy())
    assert err <= 1e-12, err


def directional_check(x32, h=1.0, ndir=3, seed=0):
    """<grad L(x32), d> from the FP32 adjoint vs FP64 central differences of L.

    Directions: d = g |x|/|g| (the gradient direction, where <g, d> cannot
    cancel; error relative to |fd|), and random Gaussian d, whose inner product
    is a sum of n terms of random sign; there the error is taken relative to
    the sum's natural scale |g * d|_2 (so a direction with an accidentally
    small <g, d> does not inflate it).  Returns the worst of these errors.
    """
    import torch
    import hpxfft_amd
    nx, ny = x32.shape[-2], x32.shape[-1]
    dev = x32.device
    w64 = weights(nx, ny, dev, torch.float64)
    xg = x32.clone().requires_grad_(True)
    loss_fn(hpxfft_amd.rfft2_autograd(xg), weights(nx, ny, dev, torch.float32)).backward()
    grad = xg.grad.double()
    x64 = x32.double()
    g = torch.Generator(device=dev).manual_seed(seed)
    dirs = [grad * (torch.linalg.vector_norm(x64) / torch.linalg.vector_norm(grad))]
    dirs += [torch.randn(x64.shape, dtype=torch.float64, device=dev, generator=g) for _ in range(ndir)]
    worst = 0.0
    for i, d in enumerate(dirs):
        lp = loss_fn(hpxfft_amd.rfft2(x64 + h * d), w64).item()
        lm = loss_fn(hpxfft_amd.rfft2(x64 - h * d), w64).item()
        fd = (lp - lm) / (2 * h)
        an = torch.sum(grad * d).item()
        scale = abs(fd) if i == 0 else torch.linalg.vector_norm(grad * d).item()
        worst = max(worst, abs(an - fd) / scale)
    return worst


@pytest.mark.parametrize("nx,ny", [(16, 16), (128, 256), (1024, 2048)])
def test_fp32_gradient_directional_fd(gpu, nx, ny):
    import torch
    x = torch.randn(nx, ny, dtype=torch.float32, device=gpu)
    assert directional_check(x) <= 1e-6


@pytest.mark.slow
def test_fp32_gradient_batch16_4096(gpu):
    """The differentiable workload: batch 16 x 2^12 x 2^12 N(0,1), FP32."""
    import torch
    print("autograd batch 16 x 4096 x 4096", flush=True)
    g = torch.Generator(device=gpu).manual_seed(5)
    x = torch.randn(16, 4096, 4096, dtype=torch.float32, device=gpu, generator=g)
    assert directional_check(x, ndir=2) <= 1e-6


def test_torch_planning_effort_measure(gpu):
    """Plans made at "measure" through the torch entry points (the column
    split is timed) give the oracle's results; the setting is validated."""
    import torch
    import hpxfft_amd
    with pytest.raises(ValueError, match="Invalid FFTW plan flag string"):
        hpxfft_amd.set_planning_effort("fast")
    nx, ny = 4096, 256
    rng = np.random.default_rng(17)
    Y = rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))
    x = torch.randn(nx, ny, dtype=torch.float64, device=gpu)
    hpxfft_amd.set_planning_effort("measure")
    try:
        X = hpxfft_amd.rfft2(x)
        adj = hpxfft_amd.rfft2_adjoint(torch.from_numpy(Y).to(gpu), ny)
        assert hpxfft_amd.planning_effort() == "measure"
    finally:
        hpxfft_amd.set_planning_effort("estimate")
        hpxfft_amd.clear_plan_cache()
    assert oracle.rel_l2(X.cpu().numpy(), oracle.r2c_2d(x.cpu().numpy())) <= 1e-12
    assert oracle.rel_l2(adj.cpu().numpy(), oracle.adjoint_r2c_2d(Y, ny)) <= 1e-12
