"""The allocation-free entry point, hpxfft_exec_r2c_2d_noalloc, through the C ABI.

It runs the kernels of hpxfft_exec_r2c_2d with a plan made at "estimate",
with the four-step intermediate in the library's static per-device scratch
array instead of plan memory, so its output must equal that entry point's bit
for bit.  Checked on calls in a row -- a four-step grid, then a smaller
four-step grid (the scratch keeps the first call's contents), then a grid
whose column pass is one launch -- for FP64 and FP32, out of place and in
place, and each output also against the CPU FP64 oracle (FP64 <= 1e-12, FP32
<= 1e-5).  Sizes beyond the scratch, 16-bit precisions and lengths that are
not powers of two are refused before any kernel runs.  A call captured into
a HIP graph replays to the eager result.
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20
PREC = {"f64": 0, "f32": 1}


def _noalloc(nx, ny, ld_in, ld_out, prec, d_in, d_out, stream=0):
    import hpxfft_amd
    L = hpxfft_amd.lib()
    return L.hpxfft_exec_r2c_2d_noalloc(nx, ny, ld_in, ld_out, PREC.get(prec, prec), ctypes.c_void_p(d_in),
                                        ctypes.c_void_p(d_out), ctypes.c_void_p(stream))


def _input(nx, ny, prec, seed):
    import torch
    x = oracle.fill(nx, ny, seed=seed, dist="upm1")
    dt = torch.float64 if prec == "f64" else torch.float32
    return x, torch.from_numpy(x).to(device="cuda", dtype=dt)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_noalloc_equals_planned_exec(gpu, prec, tol):
    import torch
    import hpxfft_amd
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    # 4096 x 2048 uses 69 MB of the scratch, 2048 x 512 about 9 MB; 512 x 4096
    # has a one-launch column pass (no scratch)
    for i, (nx, ny) in enumerate([(4096, 2048), (2048, 512), (512, 4096)]):
        print(f"noalloc {prec} {nx}x{ny} (call {i + 1})", flush=True)
        x, d = _input(nx, ny, prec, seed=40 + i)
        ld = d.shape[1]
        got = torch.empty((nx, ny // 2 + 1), dtype=cdt, device="cuda")
        assert _noalloc(nx, ny, ld, ny // 2 + 1, prec, d.data_ptr(), got.data_ptr()) == 0, hpxfft_amd.last_error()
        want = torch.empty_like(got)
        p = hpxfft_amd.Plan(nx, ny, ld_in=ld, precision=prec)
        p.exec(d.data_ptr(), want.data_ptr())
        # in place, on the padded buffer itself
        assert _noalloc(nx, ny, ld, ld // 2, prec, d.data_ptr(), d.data_ptr()) == 0, hpxfft_amd.last_error()
        torch.cuda.synchronize()
        p.close()
        assert torch.equal(got, want)
        assert torch.equal(d.view(cdt)[:, : ny // 2 + 1], want)
        ref = oracle.r2c_2d(x, ny)
        err = oracle.rel_l2(got.cpu().numpy().astype(np.complex128), ref)
        print(f"  bitwise equal to the planned exec; rel-L2 vs oracle {err:.2e}", flush=True)
        assert err <= tol


def test_noalloc_refusals(gpu):
    import torch
    import hpxfft_amd
    from hpxfft_amd import _lib
    d = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = d.data_ptr()
    # 4096 x 4096 FP64 needs 4096 x 4128 x 16 B = 136 MB of intermediate
    assert _noalloc(4096, 4096, 4098, 2049, "f64", p, p) == _lib.UNSUPPORTED_SIZE
    assert f"holds {SCRATCH_BYTES} bytes" in hpxfft_amd.last_error()
    assert _noalloc(64, 64, 66, 33, 2, p, p) == _lib.INVALID_ARG            # FP16 storage
    assert _noalloc(60, 64, 66, 33, "f64", p, p) == _lib.UNSUPPORTED_SIZE   # not a power of two
    assert _noalloc(64, 64, 64, 32, "f64", p, p) == _lib.INVALID_ARG        # ld_out < Ny/2+1
    assert _noalloc(64, 64, 66, 33, "f64", 0, p) == _lib.INVALID_ARG        # null input
    torch.cuda.synchronize()
    assert not bool(d.abs().sum())


def test_noalloc_graph_replays(gpu):
    import torch
    import hpxfft_amd
    nx, ny = 2048, 1024
    x, d = _input(nx, ny, "f64", seed=77)
    out = torch.empty((nx, ny // 2 + 1), dtype=torch.complex128, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        assert _noalloc(nx, ny, d.shape[1], ny // 2 + 1, "f64", d.data_ptr(), out.data_ptr(), s.cuda_stream) == 0
    s.synchronize()
    ref = out.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        rc = _noalloc(nx, ny, d.shape[1], ny // 2 + 1, "f64", d.data_ptr(), out.data_ptr(), s.cuda_stream)
    assert rc == 0, hpxfft_amd.last_error()
    for r in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), r
        d.mul_(1.0)   # eager work between replays
    assert oracle.rel_l2(ref.cpu().numpy(), oracle.r2c_2d(x, ny)) <= 1e-12
