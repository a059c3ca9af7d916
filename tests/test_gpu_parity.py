"""GPU parity: the HIP path (through the C ABI) against the CPU FP64 oracle.

Tolerances (relative L2 over the half spectrum):
  FP64  <= 1e-12   (the bar of the FP64 path against FFTW FP64)
  FP32  <= 1e-5    (the FP32 path's stated bound)
The reference's own known-answer test (4x6) is checked bit-exactly, as the
reference does with vector_2d::operator== (test/src/test_shared_loop.cpp:53).
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5

SIZES = [
    (1, 2), (2, 2), (4, 4), (1, 64), (2, 8), (8, 16), (16, 8), (64, 64), (32, 128), (128, 32),
    (256, 512), (1024, 64), (2048, 64), (4096, 32), (8192, 16), (512, 2048), (64, 8192),
    (256, 16384),        # the reference README sizes as exponents: 2^8 x 2^14
    (16384, 64), (32768, 4), (16, 32768), (65536, 2),
    (64, 4096), (1, 4096), (2, 16384), (4, 32768),
    (32768, 256),        # 256-point four-step tiles over several strips
]


def _run_gpu(x_pad: np.ndarray, nx: int, ny: int, precision: str) -> np.ndarray:
    """In-place transform of the padded layout on the device; returns complex128."""
    import torch
    import hpxfft_amd
    dt = torch.float64 if precision == "f64" else torch.float32
    d = torch.from_numpy(x_pad).to(device="cuda", dtype=dt).contiguous()
    p = hpxfft_amd.Plan(nx, ny, precision=precision)
    p.exec(d.data_ptr(), d.data_ptr())
    torch.cuda.synchronize()
    out = d.cpu().numpy().astype(np.float64)
    p.close()
    return out[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)


def test_reference_known_answer_4x6_bit_exact(gpu):
    import hpxfft_amd
    g = np.load(os.path.join(GOLDEN, "rfft2_golden.npz"))
    loop = hpxfft_amd.Loop()
    loop.initialize(g["ref4x6_in"].copy(), "estimate")
    out = loop.fft_2d_r2c_par()
    assert loop.get_measurement("total") >= 0.0
    assert np.array_equal(out, g["ref4x6_out"]), out


@pytest.mark.parametrize("nx,ny", SIZES)
def test_parity_f64(gpu, nx, ny):
    print(f"f64 {nx}x{ny}", flush=True)
    x = oracle.fill(nx, ny, seed=1000 + nx + ny, dist="upm1")
    ref = oracle.r2c_2d(x, ny)
    got = _run_gpu(x, nx, ny, "f64")
    err = oracle.rel_l2(got, ref)
    assert err <= TOL64, err


@pytest.mark.parametrize("prec,dt", [("f64", "float64"), ("f32", "float32")])
def test_pointers_off_a_complex_element(gpu, prec, dt):
    """Device pointers must start on a whole complex element: the C ABI
    rejects one real off with INVALID_ARG (checked on the host, nothing is
    launched), and rfft2 copies such a view before the transform."""
    import torch
    import hpxfft_amd
    from hpxfft_amd import _lib
    nx, ny = 16, 32
    ncy = ny // 2 + 1
    tdt = getattr(torch, dt)
    base = torch.zeros(nx * ny + 1, dtype=tdt, device="cuda")
    out = torch.zeros(nx * ncy * 2 + 2, dtype=tdt, device="cuda")
    p = hpxfft_amd.Plan(nx, ny, ld_in=ny, ld_out=ncy, precision=prec)
    es = base.element_size()
    for d_in, d_out in [(base.data_ptr() + es, out.data_ptr()), (base.data_ptr(), out.data_ptr() + es),
                        (base.data_ptr() + es, base.data_ptr() + es)]:
        with pytest.raises(RuntimeError) as e:
            p.exec(d_in, d_out)
        assert e.value.status == _lib.INVALID_ARG
        assert "aligned" in hpxfft_amd.last_error()
        with pytest.raises(RuntimeError) as e:
            p.exec_c2r(d_out, d_in)
        assert e.value.status == _lib.INVALID_ARG
    p.close()
    assert torch.count_nonzero(out) == 0
    x = oracle.fill(nx, ny, seed=5, dist="upm1")[:, :ny].copy()
    base[1:] = torch.from_numpy(x.reshape(-1)).to(device="cuda", dtype=tdt)
    view = base[1:].view(nx, ny)
    assert view.data_ptr() % (2 * es) != 0
    got = hpxfft_amd.rfft2(view).cpu().numpy().astype(np.complex128)
    err = oracle.rel_l2(got, oracle.r2c_2d(x))
    assert err <= (TOL64 if prec == "f64" else TOL32), err


@pytest.mark.parametrize("nx,ny,batch", [(128, 256, 3), (1024, 64, 2), (4096, 128, 3), (16, 4096, 5),
                                         # general one-level, Bluestein and two-level lengths
                                         (30, 42, 3), (67, 134, 2), (6000, 20, 2), (12, 2 * 6000, 2),
                                         # two levels in both precisions (rows: phase B with the split fused in)
                                         (18000, 20, 2), (12, 2 * 18000, 2)])
@pytest.mark.parametrize("prec,tol", [("f64", TOL64), ("f32", TOL32)])
def test_parity_batched_strided(gpu, nx, ny, batch, prec, tol):
    """Several grids in one plan, out of place, with gaps between grids and
    between rows (dist_in > Nx*ld_in, ld_out > Ny/2+1): every grid against
    the oracle, and the gaps left untouched."""
    import torch
    import hpxfft_amd
    print(f"batched {prec} {batch} x {nx}x{ny}", flush=True)
    ncy = ny // 2 + 1
    ld_in, ld_out = 2 * ncy + 6, ncy + 3
    dist_in, dist_out = nx * ld_in + 10, nx * ld_out + 7
    xs = [oracle.fill(nx, ny, seed=3000 + b, dist="upm1") for b in range(batch)]
    dt = torch.float64 if prec == "f64" else torch.float32
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    host_in = np.zeros(batch * dist_in)
    for b, x in enumerate(xs):
        host_in[b * dist_in: b * dist_in + nx * ld_in].reshape(nx, ld_in)[:, :ny] = x[:, :ny]
    d_in = torch.from_numpy(host_in).to(device="cuda", dtype=dt)
    d_out = torch.full((batch * dist_out,), complex(7.0, -7.0), dtype=cdt, device="cuda")
    p = hpxfft_amd.Plan(nx, ny, batch=batch, ld_in=ld_in, dist_in=dist_in, ld_out=ld_out, dist_out=dist_out,
                        precision=prec)
    p.exec(d_in.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    out = d_out.cpu().numpy().astype(np.complex128)
    p.close()
    for b, x in enumerate(xs):
        grid = out[b * dist_out: b * dist_out + nx * ld_out].reshape(nx, ld_out)
        err = oracle.rel_l2(grid[:, :ncy], oracle.r2c_2d(x, ny))
        assert err <= tol, (b, err)
        assert np.all(grid[:, ncy:] == complex(7.0, -7.0)), b
        assert np.all(out[b * dist_out + nx * ld_out: (b + 1) * dist_out] == complex(7.0, -7.0)), b


@pytest.mark.parametrize("ny", [4096, 16384, 32768])
@pytest.mark.parametrize("prec,tol", [("f64", TOL64), ("f32", TOL32)])
def test_long_rows_every_line_offset(gpu, ny, prec, tol):
    """Long rows (2^11, 2^13, 2^14 complex points) whose starts cycle
    through every offset within a 128-byte line, in and out: input rows one
    complex element longer than Ny reals (ld_in = Ny + 2), output rows Ny/2+1
    complex, out of place.  The FP64 long-row kernel chooses its cache hint
    per lane from that offset; the result must be the oracle's at every one."""
    import torch
    import hpxfft_amd
    nx, ld_in, ncy = 16, ny + 2, ny // 2 + 1
    x = oracle.fill(nx, ny, seed=ny + 11, dist="upm1", ld=ld_in)
    dt = torch.float64 if prec == "f64" else torch.float32
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    d_in = torch.from_numpy(x).to(device="cuda", dtype=dt).contiguous()
    d_out = torch.empty((nx, ncy), dtype=cdt, device="cuda")
    p = hpxfft_amd.Plan(nx, ny, ld_in=ld_in, ld_out=ncy, precision=prec)
    p.exec(d_in.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    p.close()
    err = oracle.rel_l2(d_out.cpu().numpy().astype(np.complex128), oracle.r2c_2d(x, ny))
    print(f"{prec} {nx}x{ny} ld_in {ld_in}: rel-L2 {err:.3e}", flush=True)
    assert err <= tol, err


@pytest.mark.parametrize("prec,tol", [("f64", TOL64), ("f32", TOL32)])
@pytest.mark.parametrize("nx,ny", [(4096, 256), (2048, 64)])
def test_measure_effort_autotunes_split(gpu, nx, ny, prec, tol):
    """effort "measure" times every valid four-step split Nx = P*Q of the
    column pass and keeps the fastest: the plan says so, and the result is
    the oracle's whatever the choice."""
    import torch
    import hpxfft_amd
    x = oracle.fill(nx, ny, seed=77, dist="upm1")
    dt = torch.float64 if prec == "f64" else torch.float32
    d = torch.from_numpy(x).to(device="cuda", dtype=dt)
    p = hpxfft_amd.Plan(nx, ny, precision=prec, effort="measure")
    desc = p.describe()
    assert "split autotuned" in desc, desc
    p.exec(d.data_ptr(), d.data_ptr())
    torch.cuda.synchronize()
    got = d.cpu().numpy().astype(np.float64)[:, : 2 * (ny // 2 + 1)].copy().view(np.complex128)
    p.close()
    assert oracle.rel_l2(got, oracle.r2c_2d(x, ny)) <= tol


@pytest.mark.parametrize("nx,ny", SIZES)
def test_parity_f32(gpu, nx, ny):
    print(f"f32 {nx}x{ny}", flush=True)
    x = oracle.fill(nx, ny, seed=2000 + nx + ny, dist="normal")
    ref = oracle.r2c_2d(x, ny)
    got = _run_gpu(x, nx, ny, "f32")
    err = oracle.rel_l2(got, ref)
    assert err <= TOL32, err


@pytest.mark.parametrize("key", ["np_upm1_8x16", "np_u01_32x64", "np_normal_64x128", "np_upm1_256x256",
                                 "np_upm1_1x32", "np_normal_16x2", "np_u01_128x8"])
def test_golden_numpy_fixtures_f64(gpu, key):
    g = np.load(os.path.join(GOLDEN, "rfft2_golden.npz"))
    a = g[key + "_in"]
    nx, ny = a.shape
    pad = np.zeros((nx, 2 * (ny // 2 + 1)))
    pad[:, :ny] = a
    got = _run_gpu(pad, nx, ny, "f64")
    assert oracle.rel_l2(got, g[key + "_out"]) <= TOL64


@pytest.mark.parametrize("nx,ny", [(8, 16), (256, 16384), (16384, 64)])
def test_ramp_closed_form(gpu, nx, ny):
    """Harness input x[i][j] = j (examples/hpxfft/shared_loop_2d.cpp:33-40):
    X[0][0] = Nx Ny (Ny-1)/2, X[0][ky] = Nx (-Ny/2 + i (Ny/2) cot(pi ky/Ny)), X[kx!=0] = 0."""
    x = oracle.fill(nx, ny, 0, "ramp")
    got = _run_gpu(x, nx, ny, "f64")
    ky = np.arange(1, ny // 2 + 1)
    row0 = np.empty(ny // 2 + 1, dtype=np.complex128)
    row0[0] = nx * ny * (ny - 1) / 2
    row0[1:] = nx * (-ny / 2 + 1j * (ny / 2) / np.tan(np.pi * ky / ny))
    row0[-1] = -nx * ny / 2
    ref = np.zeros((nx, ny // 2 + 1), dtype=np.complex128)
    ref[0] = row0
    assert oracle.rel_l2(got, ref) <= TOL64
