"""Batched uneven grids (FP32, one persistent launch per phase) against the oracle.

Each grid of the batch is checked against the CPU FP64 oracle (rel-L2 <=
1e-5, the FP32 bound) and against the single-grid path (<= 1e-5).  Both the
default task shapes and the autotuned ones ("measure") are exercised.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SIDES = [64, 128, 256, 512, 1024, 2048]


def make_sizes(count, seed):
    rng = np.random.default_rng(seed)
    return [int(SIDES[i]) for i in rng.integers(0, 6, count)], [int(SIDES[i]) for i in rng.integers(0, 6, count)]


@pytest.mark.parametrize("effort", ["estimate", "measure"])
def test_batched_vs_oracle(gpu, effort):
    import torch
    import hpxfft_amd
    nx, ny = make_sizes(40, 4)
    nx += [64, 2048, 2048, 64]          # extremes of both sides
    ny += [64, 2048, 64, 2048]
    print(f"batched {len(nx)} grids, effort {effort}", flush=True)
    xs = [oracle.fill(a, b, seed=1000 + i, dist="upm1") for i, (a, b) in enumerate(zip(nx, ny))]
    ds = [torch.from_numpy(x).to(device=gpu, dtype=torch.float32) for x in xs]
    plan = hpxfft_amd.BatchedPlan(nx, ny, effort=effort)
    plan.exec([d.data_ptr() for d in ds])
    torch.cuda.synchronize()
    assert plan.measurement("total") > 0.0
    assert "persistent launch" in plan.describe()
    worst = 0.0
    for i, (a, b) in enumerate(zip(nx, ny)):
        ref = oracle.r2c_2d(xs[i], b)
        got = ds[i].cpu().numpy().astype(np.float64).view(np.complex128)
        worst = max(worst, oracle.rel_l2(got, ref))
    print(f"  worst rel-L2 vs oracle {worst:.2e}", flush=True)
    assert worst <= 1e-5
    plan.close()


def test_batched_matches_single_grid_plans(gpu):
    import torch
    import hpxfft_amd
    nx, ny = make_sizes(12, 9)
    xs = [torch.rand(a, 2 * (b // 2 + 1), device=gpu) * 2 - 1 for a, b in zip(nx, ny)]
    for x, b in zip(xs, ny):
        x[:, b:] = 0
    ys = [x.clone() for x in xs]
    plan = hpxfft_amd.BatchedPlan(nx, ny)
    plan.exec([x.data_ptr() for x in xs])
    for y, a, b in zip(ys, nx, ny):
        p = hpxfft_amd.Plan(a, b, precision="f32")
        p.exec(y.data_ptr(), y.data_ptr())
        p.close()
    torch.cuda.synchronize()
    for x, y in zip(xs, ys):
        assert (torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(y)).item() <= 1e-5


@pytest.mark.parametrize("cfg", [0, 1, 2, 3])
def test_batched_every_configuration(gpu, cfg):
    """Each candidate task shape (row and column phase), forced, against the
    oracle on grids that include both extremes of each side."""
    import torch
    import hpxfft_amd
    nx, ny = make_sizes(10, 11)
    nx += [64, 2048, 2048, 64]
    ny += [64, 2048, 64, 2048]
    xs = [oracle.fill(a, b, seed=3000 + i, dist="normal") for i, (a, b) in enumerate(zip(nx, ny))]
    ds = [torch.from_numpy(x).to(device=gpu, dtype=torch.float32) for x in xs]
    plan = hpxfft_amd.BatchedPlan(nx, ny)
    plan.set_config(cfg, cfg)
    assert f"configuration {cfg} " in plan.describe()
    plan.exec([d.data_ptr() for d in ds])
    torch.cuda.synchronize()
    plan.close()
    for i, (a, b) in enumerate(zip(nx, ny)):
        got = ds[i].cpu().numpy().astype(np.float64).view(np.complex128)
        assert oracle.rel_l2(got, oracle.r2c_2d(xs[i], b)) <= 1e-5, (cfg, a, b)


def test_batched_set_config_rejections(gpu):
    import hpxfft_amd
    plan = hpxfft_amd.BatchedPlan([64], [64])
    for bad in [(-1, 0), (0, 4), (4, 4)]:
        with pytest.raises(RuntimeError, match="out of range"):
            plan.set_config(*bad)
    plan.close()
    p = hpxfft_amd.Plan(64, 64)
    h = p.handle
    assert hpxfft_amd.lib().hpxfft_plan_batched_set_config(h, 0, 0) != 0
    p.close()
