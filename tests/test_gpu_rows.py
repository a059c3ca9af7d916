"""Rows of 2^13 and 2^14 complex points (Ny = 16384, 32768): each row kernel
kind forced (0: radix-16 stages; 1: radix-32 stages across lane pairs with
DPP exchanges, FP32 arithmetic) against the CPU FP64 oracle.  FP64 <= 1e-12, FP32 <= 1e-5; FP16 / BF16 storage <= 5e-3 / 4e-2 of
the oracle of the unrounded input (as test_gpu_variants), with input rows of
magnitude 1e6 scaled by their exponents.  Shapes include single rows, odd
row counts (the lane pairing is per row, so every row must be whole), a batch
of grids and the in-place layout.
"""
import numpy as np
import pytest

import oracle
from test_gpu_variants import to16

pytestmark = pytest.mark.gpu

KINDS = [0, 1]


@pytest.mark.parametrize("prec,tol,kind", [("f64", 1e-12, 0), ("f32", 1e-5, 0), ("f32", 1e-5, 1)])
@pytest.mark.parametrize("nx,ny", [(64, 16384), (3, 16384), (1, 32768), (40, 32768)])
def test_row_kinds_vs_oracle(gpu, kind, prec, tol, nx, ny):
    import torch
    import hpxfft_amd
    print(f"{prec} {nx}x{ny} row kernel {kind}", flush=True)
    x = oracle.fill(nx, ny, seed=nx * 7 + ny + kind, dist="upm1")
    ref = oracle.r2c_2d(x, ny)
    dt = torch.float64 if prec == "f64" else torch.float32
    d = torch.from_numpy(x).to(device="cuda", dtype=dt)
    p = hpxfft_amd.Plan(nx, ny, precision=prec)
    p.set_row_kernel(kind)
    if kind:
        assert "lane pairs" in p.describe()
    p.exec(d.data_ptr(), d.data_ptr())   # in place on the padded buffer
    torch.cuda.synchronize()
    p.close()
    got = d.cpu().numpy().astype(np.float64).view(np.complex128)
    err = oracle.rel_l2(got, ref)
    print(f"  rel-L2 = {err:.3e}", flush=True)
    assert err <= tol


@pytest.mark.parametrize("kind", KINDS)
def test_row_kinds_batch_out_of_place(gpu, kind):
    """Three grids of 24 x 16384 FP32 in one plan, out of place."""
    import torch
    import hpxfft_amd
    nx, ny, nb = 24, 16384, 3
    xs = [oracle.fill(nx, ny, seed=900 + i, dist="normal") for i in range(nb)]
    d = torch.from_numpy(np.stack(xs)).to(device="cuda", dtype=torch.float32)
    out = torch.empty((nb, nx, ny // 2 + 1), dtype=torch.complex64, device="cuda")
    ld = 2 * (ny // 2 + 1)
    p = hpxfft_amd.Plan(nx, ny, batch=nb, ld_in=ld, precision="f32")
    p.set_row_kernel(kind)
    p.exec(d.data_ptr(), out.data_ptr())
    torch.cuda.synchronize()
    p.close()
    for i in range(nb):
        got = out[i].cpu().numpy().astype(np.complex128)
        assert oracle.rel_l2(got, oracle.r2c_2d(xs[i], ny)) <= 1e-5, i


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dt_name,bound", [("f16", 5e-3), ("bf16", 4e-2)])
@pytest.mark.parametrize("nx,ny,amp", [(64, 16384, 1.0), (16, 32768, 1e6), (5, 32768, 1.0)])
def test_row_kinds_16bit(gpu, kind, dt_name, bound, nx, ny, amp):
    import torch
    import hpxfft_amd
    dt = torch.float16 if dt_name == "f16" else torch.bfloat16
    x = oracle.fill(nx, ny, seed=nx + ny + 17 * kind, dist="normal") * amp
    ref = oracle.r2c_2d(x, ny)
    t, e_dev, _ = to16(x, ny, dt, scale_rows=amp != 1.0)
    ld = t.shape[1]
    col_exp = torch.empty(ny // 2 + 1, dtype=torch.int32, device="cuda")
    p = hpxfft_amd.Plan(nx, ny, ld_in=ld, ld_out=ld // 2, precision=dt_name)
    p.set_row_kernel(kind)
    p.exec_scaled(t.data_ptr(), t.data_ptr(), e_dev.data_ptr() if amp != 1.0 else None, col_exp.data_ptr())
    torch.cuda.synchronize()
    p.close()
    X = t.view(nx, ld // 2, 2).float().cpu().numpy().astype(np.float64)
    Xc = (X[..., 0] + 1j * X[..., 1]) * np.exp2(-col_exp.cpu().numpy().astype(np.float64))[None, :]
    assert np.all(np.isfinite(Xc))
    err = oracle.rel_l2(Xc, ref)
    print(f"{dt_name} {nx}x{ny} kind {kind}: rel-L2 {err:.3e}", flush=True)
    assert err <= bound


def test_row_kernel_kind_rejections(gpu):
    import hpxfft_amd
    p = hpxfft_amd.Plan(64, 16384, precision="f32")
    for bad in (-1, 2):
        with pytest.raises(hpxfft_amd.HpxfftError, match="out of range"):
            p.set_row_kernel(bad)
    p.close()
    for nx, ny, prec in [(64, 4096, "f32"), (64, 16384, "f64")]:
        q = hpxfft_amd.Plan(nx, ny, precision=prec)
        q.set_row_kernel(0)
        with pytest.raises(hpxfft_amd.HpxfftError, match="2\\^13 and 2\\^14"):
            q.set_row_kernel(1)
        q.close()


@pytest.mark.parametrize("effort", ["estimate", "measure", "patient", "exhaustive"])
def test_planning_efforts_widen_the_search(gpu, effort):
    """estimate times nothing; measure times the four-step splits; patient and
    exhaustive also time both row kernels (listed in the plan text) and keep
    the faster; every plan is checked against the oracle."""
    import torch
    import hpxfft_amd
    nx, ny = 2048, 16384
    x = oracle.fill(nx, ny, seed=31, dist="upm1")
    ref = oracle.r2c_2d(x, ny)
    d = torch.from_numpy(x).to(device="cuda", dtype=torch.float32)
    p = hpxfft_amd.Plan(nx, ny, precision="f32", effort=effort)
    desc = p.describe()
    assert ("split autotuned" in desc) == (effort != "estimate")
    assert ("row kernel autotuned" in desc) == (effort in ("patient", "exhaustive"))
    p.exec(d.data_ptr(), d.data_ptr())
    torch.cuda.synchronize()
    p.close()
    err = oracle.rel_l2(d.cpu().numpy().astype(np.float64).view(np.complex128), ref)
    print(f"{effort}: {[ln.strip() for ln in desc.splitlines() if 'autotuned' in ln]} rel-L2 {err:.2e}", flush=True)
    assert err <= 1e-5
