"""torch entry points over the C ABI (torch provides device memory and the
current HIP stream; all arithmetic runs in libhpxfft_hip.so).

* :func:`rfft2`            -- forward 2-D r2c of (..., Nx, Ny) real tensors.
* :func:`rfft2_scaled`     -- 16-bit storage variant with power-of-two scales.
* :func:`rfft2_adjoint`    -- the adjoint R* of the r2c map.
* :func:`c2r_2d`           -- the unnormalised inverse (FFTW_BACKWARD):
                              c2r_2d(rfft2(x), Ny) = Nx*Ny*x.
* :class:`RFFT2`           -- torch.autograd.Function: forward = rfft2,
                              backward = rfft2_adjoint (hand-written kernels).
* :func:`set_planning_effort` -- the planning effort of the cached plans these
                              functions create ("estimate" by default;
                              "measure" times the column-pass choices once
                              per shape, as FFTW's MEASURE does).
"""
from __future__ import annotations

import torch

from .plan import Plan

_DT = {torch.float64: ("f64", torch.complex128), torch.float32: ("f32", torch.complex64)}
_plans: dict = {}
_EFFORTS = ("estimate", "measure", "patient", "exhaustive")
_effort = "estimate"


def set_planning_effort(effort: str) -> None:
    """Planning effort of the plans created from now on (plans already
    cached keep theirs; clear_plan_cache() drops them)."""
    global _effort
    if effort not in _EFFORTS:
        raise ValueError(f"Invalid FFTW plan flag string: {effort!r}")
    _effort = effort


def planning_effort() -> str:
    return _effort


def _get_plan(nx: int, ny: int, batch: int, precision: str, ld_in: int, ld_out: int, device: torch.device) -> Plan:
    stream = torch.cuda.current_stream(device)
    key = (nx, ny, batch, precision, ld_in, ld_out, device.index, stream.cuda_stream, _effort)
    p = _plans.get(key)
    if p is None:
        with torch.cuda.device(device):
            p = Plan(nx, ny, batch=batch, ld_in=ld_in, dist_in=nx * ld_in, ld_out=ld_out, dist_out=nx * ld_out,
                     precision=precision, stream=stream, effort=_effort)
        _plans[key] = p
    return p


def clear_plan_cache() -> None:
    for p in _plans.values():
        p.close()
    _plans.clear()


def _check_real(x: torch.Tensor):
    if not x.is_cuda:
        raise ValueError("hpxfft_amd operates on GPU tensors (no CPU path)")
    if x.dtype not in _DT:
        raise TypeError(f"unsupported dtype {x.dtype}; use float64 or float32 (rfft2_scaled for 16-bit)")
    if x.dim() < 2:
        raise ValueError("need (..., Nx, Ny)")


def _on_complex_element(t: torch.Tensor) -> torch.Tensor:
    """Contiguous, starting on a whole complex element (the kernels' unit):
    a view that starts one real in (storage offset) is copied."""
    t = t.contiguous()
    unit = t.element_size() * (1 if t.is_complex() else 2)
    return t if t.data_ptr() % unit == 0 else t.clone()


def rfft2(x: torch.Tensor) -> torch.Tensor:
    """X[..., kx, ky] = sum_{m,n} x[..., m, n] exp(-2 pi i (kx m/Nx + ky n/Ny)), ky <= Ny/2."""
    _check_real(x)
    x = _on_complex_element(x)
    nx, ny = x.shape[-2], x.shape[-1]
    batch = x.numel() // (nx * ny) if x.numel() else 0
    prec, cdt = _DT[x.dtype]
    out = torch.empty((*x.shape[:-2], nx, ny // 2 + 1), dtype=cdt, device=x.device)
    if batch == 0:
        return out
    p = _get_plan(nx, ny, batch, prec, ny, ny // 2 + 1, x.device)
    p.exec(x.data_ptr(), out.data_ptr())
    return out


def rfft2_adjoint(g: torch.Tensor, ny: int) -> torch.Tensor:
    """(R* g)[..., m, n] = Re sum_{kx, ky<=Ny/2} g[..., kx, ky] exp(+2 pi i (kx m/Nx + ky n/Ny))."""
    return _backward(g, ny, adjoint=True)


def c2r_2d(X: torch.Tensor, ny: int) -> torch.Tensor:
    """Unnormalised 2-D c2r (FFTW_BACKWARD) of a half spectrum (..., Nx, Ny/2+1):
    x = Nx*Ny * numpy.fft.irfft2(X, s=(Nx, Ny)) (imaginary parts of the ky = 0
    and ky = Ny/2 bins ignored after the column transform)."""
    return _backward(X, ny, adjoint=False)


def _backward(g: torch.Tensor, ny: int, adjoint: bool) -> torch.Tensor:
    if not g.is_cuda:
        raise ValueError("hpxfft_amd operates on GPU tensors (no CPU path)")
    rdt = {torch.complex128: torch.float64, torch.complex64: torch.float32}.get(g.dtype)
    if rdt is None:
        raise TypeError(f"unsupported dtype {g.dtype}")
    g = _on_complex_element(g)
    nx = g.shape[-2]
    if g.shape[-1] != ny // 2 + 1:
        raise ValueError("last dim must be Ny/2+1")
    batch = g.numel() // (nx * (ny // 2 + 1)) if g.numel() else 0
    out = torch.empty((*g.shape[:-2], nx, ny), dtype=rdt, device=g.device)
    if batch == 0:
        return out
    p = _get_plan(nx, ny, batch, _DT[rdt][0], ny, ny // 2 + 1, g.device)
    if adjoint:
        p.exec_adjoint(g.data_ptr(), out.data_ptr())
    else:
        p.exec_c2r(g.data_ptr(), out.data_ptr())
    return out


def rfft2_scaled(x16: torch.Tensor, in_row_exp: torch.Tensor | None = None):
    """16-bit storage r2c (float16 or bfloat16 in, same dtype out, FP32 arithmetic).

    ``x16`` is (Nx, 2*(Ny/2+1)) padded (in-place layout); the real value of
    row m is x16[m] * 2**-in_row_exp[m].  Returns (X16, col_exp) with the true
    spectrum X[kx, ky] = X16[kx, ky] * 2**-col_exp[ky]; X16 is a complex view
    as (Nx, Ny/2+1, 2) of the same dtype, computed in place in a copy.
    """
    if x16.dtype not in (torch.float16, torch.bfloat16):
        raise TypeError("rfft2_scaled takes float16 or bfloat16")
    if not x16.is_cuda:
        raise ValueError("hpxfft_amd operates on GPU tensors (no CPU path)")
    if x16.dim() != 2:
        raise ValueError("rfft2_scaled takes one (Nx, 2*(Ny/2+1)) grid")
    nx, ld = x16.shape
    ny = ld - 2
    if in_row_exp is not None:
        # the kernels read Nx int32 exponents from this pointer on x16's device
        if not in_row_exp.is_cuda or in_row_exp.device != x16.device:
            raise ValueError("in_row_exp must be a tensor on the device of x16")
        if in_row_exp.dtype != torch.int32:
            raise TypeError(f"in_row_exp must be int32 (got {in_row_exp.dtype})")
        if in_row_exp.numel() != nx:
            raise ValueError(f"in_row_exp needs Nx = {nx} entries (got {in_row_exp.numel()})")
        in_row_exp = in_row_exp.contiguous()
    buf = x16.contiguous().clone()
    col_exp = torch.empty(ny // 2 + 1, dtype=torch.int32, device=x16.device)
    prec = "f16" if x16.dtype == torch.float16 else "bf16"
    p = _get_plan(nx, ny, 1, prec, ld, ld // 2, x16.device)
    p.exec_scaled(buf.data_ptr(), buf.data_ptr(), None if in_row_exp is None else in_row_exp.data_ptr(),
                  col_exp.data_ptr())
    return buf.view(nx, ld // 2, 2), col_exp


class RFFT2(torch.autograd.Function):
    """Differentiable 2-D r2c: backward is the hand-written adjoint kernel pair.

    With torch's convention for a real loss L and complex output X
    (dL = Re sum conj(G) dX), grad_x = R*(G).
    """

    @staticmethod
    def forward(ctx, x: torch.Tensor) -> torch.Tensor:
        ctx.ny = x.shape[-1]
        return rfft2(x)

    @staticmethod
    def backward(ctx, grad: torch.Tensor):
        return rfft2_adjoint(grad, ctx.ny)


def rfft2_autograd(x: torch.Tensor) -> torch.Tensor:
    return RFFT2.apply(x)
