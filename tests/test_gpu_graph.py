"""The transforms under HIP graph capture (torch.cuda.CUDAGraph), on the GPU.

exec launches only kernels and event records on the plan's stream (no
allocation, no memset), so a caller can capture it.  Each captured graph is
replayed several times, with eager work and synchronisation between the
replays, and every replay must equal the eager result bitwise.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,prec", [(2048, 2048, "f64"), (512, 4096, "f32"), (300, 3000, "f64"),
                                        (6000, 16, "f64"),       # one LDS buffer, in place
                                        (67, 134, "f64"),        # Bluestein, in place
                                        (18000, 16, "f64"),      # two-level columns
                                        (8, 2 * 18000, "f64")])  # two-level rows, split fused into phase B
def test_forward_graph_replays(gpu, nx, ny, prec):
    import torch
    import hpxfft_amd
    dt = torch.float64 if prec == "f64" else torch.float32
    x = torch.from_numpy(oracle.fill(nx, ny, seed=nx + ny, dist="upm1")).to(device="cuda", dtype=dt)
    out = torch.empty_like(x)
    s = torch.cuda.Stream()
    p = hpxfft_amd.Plan(nx, ny, precision=prec, stream=s)
    with torch.cuda.stream(s):
        p.exec(x.data_ptr(), out.data_ptr())
    s.synchronize()
    ref = out.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        p.exec(x.data_ptr(), out.data_ptr())
    for r in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), r
        x.mul_(1.0)          # eager work between replays
    p.close()


def test_c2r_graph_replays(gpu):
    import torch
    import hpxfft_amd
    nx, ny = 1024, 2048
    rng = np.random.default_rng(3)
    Y = torch.from_numpy(rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))).cuda()
    out = torch.empty((nx, ny), dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    p = hpxfft_amd.Plan(nx, ny, ld_in=ny, ld_out=ny // 2 + 1, stream=s)
    with torch.cuda.stream(s):
        p.exec_c2r(Y.data_ptr(), out.data_ptr())
    s.synchronize()
    ref = out.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        p.exec_c2r(Y.data_ptr(), out.data_ptr())
    for r in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), r
    p.close()
