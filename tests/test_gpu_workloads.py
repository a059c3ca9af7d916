"""The benchmarked workloads, at their full sizes, against the CPU FP64 oracle.

The oracle is ``oracle.Loop``, the restatement of
hpxfft::shared::loop::fft_2d_r2c_par (core/src/shared/loop.cpp:56-113), run
on the unrounded FP64 input.  Relative L2 errors are taken over the whole
half spectrum (the norms are formed in FP64 on the device, against the
oracle's spectrum uploaded there).

* 2^14 x 2^14, U[-1,1) seed 2, out of place with the padded input pitch (the
  headline grid as bench.py runs it): FP64 <= 1e-12 and FP32 <= 1e-5 at
  every four-step split the measure effort may keep (P = 64, 128, 256) and
  with the plan made at effort "measure".
* 2^15 x 2^15, N(0,1) seed 3, plans at "measure": FP64 <= 1e-12, FP32 <=
  1e-5, FP16 storage <= 1e-3 and BF16 storage <= 8e-3 (11- and 8-bit
  significands: the input rounding alone gives about 2.8e-4 and 2.3e-3).
  The FP32 scratch of this grid (4.3 GB) is above 2^32 bytes, so this is
  where the four-step tiles' 64-bit offsets are checked.
* The 512 uneven FP32 grids of the batched workload (sides 2^(6+k), k the
  SplitMix64 outputs of seed 4 mod 6, drawn as bench.py draws them), inputs
  U[-1,1) from per-grid seeds, at "estimate" and "measure": every grid
  <= 1e-5.
"""
import numpy as np
import pytest

import oracle

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N14 = 1 << 14
N15 = 1 << 15


def _rel_l2_dev(got, ref):
    import torch
    return (torch.linalg.vector_norm((got.to(torch.complex128) - ref).flatten()) /
            torch.linalg.vector_norm(ref.flatten())).item()


def _oracle_on_device(x_pad, n):
    """The oracle's spectrum of the padded FP64 host array (transformed in
    place), as an n x (n/2+1) complex128 device tensor."""
    import torch
    oracle.Loop(n, n).exec(x_pad)
    return torch.from_numpy(x_pad.view(np.complex128)).cuda()


@pytest.fixture(scope="module")
def grid14(gpu):
    import torch
    print(f"oracle {N14}x{N14} U[-1,1) seed 2", flush=True)
    x = oracle.fill(N14, N14, seed=2, dist="upm1")
    d64 = torch.from_numpy(x).cuda()
    ref = _oracle_on_device(x, N14)
    del x
    yield d64, ref
    del d64, ref
    torch.cuda.empty_cache()


def _run14(d64, prec, split, effort="estimate"):
    import torch
    import hpxfft_amd
    ld = 2 * (N14 // 2 + 1)
    xin = d64 if prec == "f64" else d64.float()
    out = torch.empty((N14, N14 // 2 + 1), device="cuda",
                      dtype=torch.complex128 if prec == "f64" else torch.complex64)
    p = hpxfft_amd.Plan(N14, N14, ld_in=ld, precision=prec, effort=effort)
    if split is not None:
        p.set_column_split(split)
    p.exec(xin.data_ptr(), out.data_ptr())
    torch.cuda.synchronize()
    desc = p.describe()
    p.close()
    return out, desc


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
@pytest.mark.parametrize("split", [6, 7, 8])
def test_headline_grid_every_split(grid14, prec, tol, split):
    d64, ref = grid14
    print(f"{prec} {N14}x{N14} P=2^{split}", flush=True)
    out, desc = _run14(d64, prec, split)
    assert f"P*Q = {1 << split} * {N14 >> split}" in desc
    err = _rel_l2_dev(out, ref)
    print(f"  rel-L2 vs oracle = {err:.3e}", flush=True)
    assert err <= tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_headline_grid_measure_plan(grid14, prec, tol):
    """The plan exactly as the benchmark makes it (effort "measure")."""
    d64, ref = grid14
    out, desc = _run14(d64, prec, None, effort="measure")
    split = [ln.strip() for ln in desc.splitlines() if "four-step" in ln or "autotuned" in ln]
    err = _rel_l2_dev(out, ref)
    print(f"{prec} measure plan {split}: rel-L2 vs oracle = {err:.3e}", flush=True)
    assert err <= tol


def test_column_split_range(grid14):
    import hpxfft_amd
    p = hpxfft_amd.Plan(N14, N14, precision="f64")
    for bad in (5, 9, -1):
        with pytest.raises(hpxfft_amd.HpxfftError, match="out of range"):
            p.set_column_split(bad)
    p.close()
    q = hpxfft_amd.Plan(256, 256, precision="f64")   # one-launch column pass
    with pytest.raises(hpxfft_amd.HpxfftError, match="not a four-step"):
        q.set_column_split(6)
    q.close()


@pytest.fixture(scope="module")
def grid15(gpu):
    import torch
    print(f"oracle {N15}x{N15} N(0,1) seed 3", flush=True)
    x = oracle.fill(N15, N15, seed=3, dist="normal")
    d64 = torch.from_numpy(x).cuda()
    ref = _oracle_on_device(x, N15)
    del x
    yield d64, ref
    del d64, ref
    torch.cuda.empty_cache()


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_2e15_grid(grid15, prec, tol):
    import torch
    import hpxfft_amd
    d64, ref = grid15
    ld = 2 * (N15 // 2 + 1)
    xin = d64 if prec == "f64" else d64.float()
    out = torch.empty((N15, N15 // 2 + 1), device="cuda",
                      dtype=torch.complex128 if prec == "f64" else torch.complex64)
    p = hpxfft_amd.Plan(N15, N15, ld_in=ld, precision=prec, effort="measure")
    p.exec(xin.data_ptr(), out.data_ptr())
    torch.cuda.synchronize()
    p.close()
    del xin
    err = _rel_l2_dev(out, ref)
    print(f"{prec} {N15}x{N15} N(0,1): rel-L2 vs oracle = {err:.3e}", flush=True)
    del out
    torch.cuda.empty_cache()
    assert err <= tol


@pytest.mark.parametrize("prec,tol", [("f16", 1e-3), ("bf16", 8e-3)])
def test_2e15_grid_16bit_storage(grid15, prec, tol):
    """In place, input unscaled (|x| < 2^3), per-row scales after the row
    pass and per-column exponents of the output, as bench.py runs it."""
    import torch
    import hpxfft_amd
    d64, ref = grid15
    ld = 2 * (N15 // 2 + 1)
    buf = d64.to(torch.float16 if prec == "f16" else torch.bfloat16)
    col_exp = torch.empty(N15 // 2 + 1, dtype=torch.int32, device="cuda")
    p = hpxfft_amd.Plan(N15, N15, ld_in=ld, ld_out=ld // 2, precision=prec, effort="measure")
    p.exec_scaled(buf.data_ptr(), buf.data_ptr(), None, col_exp.data_ptr())
    torch.cuda.synchronize()
    p.close()
    Xs = buf.view(N15, ld // 2, 2)
    X = torch.complex(Xs[..., 0].double(), Xs[..., 1].double()) * torch.exp2(-col_exp.double())[None, :]
    del buf, Xs
    assert bool(torch.isfinite(X.real).all()) and bool(torch.isfinite(X.imag).all())
    err = _rel_l2_dev(X, ref)
    print(f"{prec}-storage {N15}x{N15} N(0,1): rel-L2 vs oracle = {err:.3e}", flush=True)
    del X
    torch.cuda.empty_cache()
    assert err <= tol


def _splitmix64(seed, k):
    m = (1 << 64) - 1
    z = (seed + (k + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def batched_sides(count=512, seed=4):
    """Grid sides of the batched workload: 64 << (SplitMix64 output mod 6)."""
    nx = [64 << (_splitmix64(seed, 2 * i) % 6) for i in range(count)]
    ny = [64 << (_splitmix64(seed, 2 * i + 1) % 6) for i in range(count)]
    return nx, ny


@pytest.mark.parametrize("effort", ["estimate", "measure"])
def test_batched_512_grids(gpu, effort):
    import torch
    import hpxfft_amd
    nx, ny = batched_sides()
    assert len(set(zip(nx, ny))) == 36   # every (Nx, Ny) pair of sides occurs
    print(f"batched 512 grids (seed 4), effort {effort}", flush=True)
    xs = [oracle.fill(a, b, seed=4000 + i, dist="upm1") for i, (a, b) in enumerate(zip(nx, ny))]
    ds = [torch.from_numpy(x).to(device="cuda", dtype=torch.float32) for x in xs]
    plan = hpxfft_amd.BatchedPlan(nx, ny, effort=effort)
    plan.exec([d.data_ptr() for d in ds])
    torch.cuda.synchronize()
    plan.close()
    worst, where = 0.0, None
    for i, (a, b) in enumerate(zip(nx, ny)):
        ref = oracle.r2c_2d(xs[i], b)
        got = ds[i].cpu().numpy().astype(np.float64).view(np.complex128)
        e = oracle.rel_l2(got, ref)
        if e > worst:
            worst, where = e, (i, a, b)
    print(f"  worst rel-L2 vs oracle {worst:.2e} at grid {where}", flush=True)
    assert worst <= 1e-5
