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


@pytest.mark.parametrize("nx,ny,tasks", [([64], [64], 1), ([64, 128, 64], [64, 64, 256], 3), ([2048], [128], 9)])
def test_batched_fewer_column_tasks_than_ranges(gpu, nx, ny, tasks):
    """The column tasks are cut into 8 contiguous ranges, one per XCD, each
    with its own ticket counter; with fewer tasks than ranges most ranges are
    empty and every workgroup takes its tasks from the others.  1, 3 and 9
    column tasks (16384-point task shape), against the oracle."""
    import re
    import torch
    import hpxfft_amd
    xs = [oracle.fill(a, b, seed=5000 + i, dist="upm1") for i, (a, b) in enumerate(zip(nx, ny))]
    ds = [torch.from_numpy(x).to(device=gpu, dtype=torch.float32) for x in xs]
    plan = hpxfft_amd.BatchedPlan(nx, ny)
    plan.set_config(2, 3)
    col = [ln for ln in plan.describe().splitlines() if "column phase" in ln][0]
    assert int(re.search(r"(\d+) tasks", col).group(1)) == tasks, col
    for _ in range(2):   # the counters are cleared for every exec
        for d, x in zip(ds, xs):
            d.copy_(torch.from_numpy(x))
        plan.exec([d.data_ptr() for d in ds])
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(nx, ny)):
            got = ds[i].cpu().numpy().astype(np.float64).view(np.complex128)
            assert oracle.rel_l2(got, oracle.r2c_2d(xs[i], b)) <= 1e-5, (a, b)
    plan.close()


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


def test_batched_alternating_pointer_sets(gpu):
    """Calls that switch between two pointer sets without host waits (the
    staging slots are reused in turn) transform every grid correctly."""
    import torch
    import hpxfft_amd
    nx, ny = make_sizes(24, 21)
    sets = []
    for k in range(2):
        xs = [oracle.fill(a, b, seed=5000 + 100 * k + i, dist="upm1") for i, (a, b) in enumerate(zip(nx, ny))]
        sets.append((xs, [torch.from_numpy(x).to(device=gpu, dtype=torch.float32) for x in xs]))
    plan = hpxfft_amd.BatchedPlan(nx, ny)
    # set 0, set 1, then set 0 again on freshly refilled data: four uploads, no synchronisation between
    plan.exec([d.data_ptr() for d in sets[0][1]])
    plan.exec([d.data_ptr() for d in sets[1][1]])
    for d, x in zip(sets[0][1], sets[0][0]):
        d.copy_(torch.from_numpy(x))
    plan.exec([d.data_ptr() for d in sets[0][1]])
    torch.cuda.synchronize()
    plan.close()
    for xs, ds in sets:
        for i, (a, b) in enumerate(zip(nx, ny)):
            got = ds[i].cpu().numpy().astype(np.float64).view(np.complex128)
            assert oracle.rel_l2(got, oracle.r2c_2d(xs[i], b)) <= 1e-5, (i, a, b)


def test_batched_graph_replays(gpu):
    """With the pointer set of the previous call, a batched exec enqueues only
    kernels: captured once, each replay (input restored eagerly in between)
    equals the eager result bitwise.  Capturing a call with a new pointer set
    is refused before anything is enqueued."""
    import torch
    import hpxfft_amd
    nx, ny = make_sizes(16, 31)
    x0 = [torch.from_numpy(oracle.fill(a, b, seed=6000 + i, dist="normal")).to(device=gpu, dtype=torch.float32)
          for i, (a, b) in enumerate(zip(nx, ny))]
    ds = [x.clone() for x in x0]
    s = torch.cuda.Stream()
    plan = hpxfft_amd.BatchedPlan(nx, ny, stream=s)
    ptrs = [d.data_ptr() for d in ds]
    with torch.cuda.stream(s):
        plan.exec(ptrs)
    s.synchronize()
    ref = [d.clone() for d in ds]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.exec(ptrs)
    for r in range(3):
        for d, x in zip(ds, x0):
            d.copy_(x)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(d, e) for d, e in zip(ds, ref)), r
    others = [x.clone() for x in x0]
    g2 = torch.cuda.CUDAGraph()
    refused = None
    with torch.cuda.graph(g2, stream=s):
        try:
            plan.exec([o.data_ptr() for o in others])
        except hpxfft_amd.HpxfftError as e:
            refused = e
    assert refused is not None and "previous exec" in str(refused)
    assert all(torch.equal(o, x) for o, x in zip(others, x0))   # nothing ran
    plan.close()
