#!/usr/bin/env python3
"""Benchmark: 2-D r2c FFT throughput on one MI355X (gfx950).

Workload (the headline configuration): Nx = Ny = 2^14 real input, uniform
[-1,1) synthetic data, FP64 storage and arithmetic, forward transform, data
resident in HBM.  One step = one full 2-D r2c (row pass + column pass) of the
grid.  Also measured: the same grid in FP32 (speed-up and rel-L2 vs FP64).

value = algorithmic bytes / time, 2-pass model (transposes not credited):
    B = [Nx*Ny*s + H] + [2H],  H = Nx*(Ny/2+1)*2*s,  s = bytes per real.

Prints ONE JSON line.  Flags: --steps K --warmup W --gpus N (N must be 1)
--no-cpu (skip the CPU baseline leg) --no-f32 --no-extra (skip the other
legs) --nx/--ny (other sizes) --copy-probe BYTES (only a device copy, for
calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE) --general-only (only the
general-length legs, one JSON line) --only LEG (one leg by name, for
profiling it alone).  The vendor_rocfft leg times hipFFT (over rocFFT),
called directly and through torch.fft.rfft2, on the same transforms beside
this build's plans; cpu_library times scipy.fft.rfft2 on the host cores beside
the oracle's CPU port (cpu_baseline).

Every leg reports its launches (event-timed seconds, algorithmic bytes and,
where a committed counter summary covers the workload, HBM traffic) and a
roofline entry for its longest launch.

roofline: the dominant kernel, the pass with the longest event-timed
duration on the plan's stream (a four-step column pass is one kernel in two
launches, phases A and B); achieved = its algorithmic bytes per launch / its
average launch duration; traffic = its HBM bytes per launch from
profiles/traffic.json (rocprofv3 counters scaled by the copy-kernel factor),
null when no summary covers the workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (REPO, os.path.join(REPO, "hpx-fft_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
FP64_VEC_TFLOPS = 78.6         # MI355X, published
FP32_VEC_TFLOPS = 157.3


def alg_bytes(nx: int, ny: int, s: int) -> dict:
    H = nx * (ny // 2 + 1) * 2 * s
    row = nx * ny * s + H
    col = 2 * H
    return {"row": row, "col": col, "total": row + col}


def nominal_flops(nx: int, ny: int) -> float:
    n = nx * ny
    return 5.0 * n * math.log2(n)


def time_gpu(nx: int, ny: int, precision: str, steps: int, warmup: int, seed: int):
    import torch
    import hpxfft_amd

    dt = torch.float64 if precision == "f64" else torch.float32
    ld = 2 * (ny // 2 + 1)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.empty((nx, ld), dtype=dt, device="cuda")
    x[:, :ny].uniform_(-1.0, 1.0, generator=g)
    x[:, ny:].zero_()
    out = torch.empty((nx, ny // 2 + 1), dtype=torch.complex128 if precision == "f64" else torch.complex64,
                      device="cuda")
    stream = torch.cuda.current_stream()
    # effort "measure": the plan times the valid four-step splits of the column pass (outside the timed region)
    plan = hpxfft_amd.Plan(nx, ny, ld_in=ld, precision=precision, effort="measure", stream=stream)
    for _ in range(warmup):
        plan.exec(x.data_ptr(), out.data_ptr())
    torch.cuda.synchronize()
    # timed region: K steps back to back on the plan's stream
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.exec(x.data_ptr(), out.data_ptr())
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / steps
    # per-phase times from the plan's own HIP events (row pass, column pass)
    med = phase_medians(plan, lambda: plan.exec(x.data_ptr(), out.data_ptr()), max(3, min(steps, 10)))
    desc = plan.describe()
    plan_s = plan.measurement("plan")
    plan.close()
    return {"ms_per_step": ms, "wall_s": wall, "row_s": med["first_fftw"], "col_s": med["second_fftw"],
            "col_a_s": med["column_a"], "col_b_s": med["column_b"], "out": out, "x": x, "plan": desc,
            "plan_s": plan_s}


def launches(nx: int, ny: int, s: int, r: dict) -> dict:
    """Per-launch algorithmic bytes and event-timed duration.  A four-step
    column pass is two launches; each is credited half of the column pass's
    2H (phase A reads the spectrum, phase B writes it; the scratch between
    them is not algorithmic)."""
    b = alg_bytes(nx, ny, s)
    out = {"row_r2c": (b["row"], r["row_s"])}
    if r["col_a_s"] > 0.0:
        out["column_phase_a"] = (b["col"] // 2, r["col_a_s"])
        out["column_phase_b"] = (b["col"] // 2, r["col_b_s"])
    else:
        out["column_c2c"] = (b["col"], r["col_s"])
    return out


def measured_traffic(workload: str) -> dict:
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (profiles/traffic.json), already scaled by the copy-kernel factor."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(workload, {})
    except (OSError, ValueError):
        return {}


def phase_medians(plan, run, reps: int = 5) -> dict:
    """Median event-timed seconds of each phase key over `reps` runs of `run()`
    (one exec of `plan` on the plan's stream)."""
    keys = ("first_fftw", "second_fftw", "column_a", "column_b")
    ph = {k: [] for k in keys}
    for _ in range(reps):
        run()
        for k in keys:
            ph[k].append(plan.measurement(k))
    return {k: sorted(v)[len(v) // 2] for k, v in ph.items()}


def pass_launches(med: dict, row_bytes: int, col_bytes: int, backward: bool = False, row="row_r2c") -> dict:
    """Per-launch (algorithmic bytes, seconds) of a forward (row pass, then
    column pass) or backward (column pass, then rows) exec.  A four-step
    column pass is two launches, phases A and B, credited half the column
    pass's bytes each."""
    row_s, col_s = (med["second_fftw"], med["first_fftw"]) if backward else (med["first_fftw"], med["second_fftw"])
    out = {row: (row_bytes, row_s)}
    if med["column_a"] > 0.0:
        out["column_phase_a"] = (col_bytes // 2, med["column_a"])
        out["column_phase_b"] = (col_bytes // 2, med["column_b"])
    else:
        out["column_c2c"] = (col_bytes, col_s)
    return out


def launch_table(launches: dict, workload: str) -> dict:
    tr = measured_traffic(workload)
    return {k: {"s": v[1], "alg_bytes": v[0], "alg_GBps": round(v[0] / v[1] / 1e9, 1) if v[1] > 0 else None,
                "traffic": tr.get(k),
                "traffic_GBps": round(tr[k] / v[1] / 1e9, 1) if tr.get(k) and v[1] > 0 else None}
            for k, v in launches.items()}


def leg_roofline(launches: dict, workload: str) -> dict:
    """Roofline entry of a leg's dominant launch (the longest event-timed one):
    its algorithmic bytes / its duration against the HBM peak, and its HBM
    traffic from the committed counter summary when one covers the workload."""
    name = max(launches, key=lambda k: launches[k][1])
    b, sec = launches[name]
    ach = b / sec / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(workload).get(name), "kernel": name,
            "alg_bytes_per_launch": b, "avg_launch_s": sec, "workload": workload}


def progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_probe(nbytes: int, reps: int = 3) -> int:
    """A plain device copy of known size (counter calibration under rocprofv3 --pmc)."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    ms = _events_ms(lambda: dst.copy_(src), reps, warmup=1)
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps, "ms": round(ms, 4),
                      "GBps_read_plus_write": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)
    return 0


def _sm64(seed: int, k: int) -> int:
    """SplitMix64 output k of a generator seeded with `seed` (counter form)."""
    m = (1 << 64) - 1
    z = (seed + (k + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def _events_ms(fn, reps: int, warmup: int = 2) -> float:
    import torch
    for _ in range(warmup):
        fn()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def bench_batched(count: int = 512, seed: int = 4, reps: int = 5, per_grid: bool = True) -> dict:
    """512 uneven FP32 grids, sides 2^(6 + k), k = SplitMix64 output mod 6.
    per_grid=False skips the comparison with one plan per grid (512 plans, a
    few thousand small launches), leaving the batched plans alone."""
    import torch
    import hpxfft_amd

    sides = [64 << (_sm64(seed, 2 * i) % 6) for i in range(count)], [64 << (_sm64(seed, 2 * i + 1) % 6) for i in
                                                                     range(count)]
    nx, ny = sides
    g = torch.Generator(device="cuda").manual_seed(seed)
    xs = []
    for a, b in zip(nx, ny):
        x = torch.empty((a, 2 * (b // 2 + 1)), dtype=torch.float32, device="cuda")
        x[:, :b].uniform_(-1.0, 1.0, generator=g)
        x[:, b:].zero_()
        xs.append(x)
    ptrs = [x.data_ptr() for x in xs]
    stream = torch.cuda.current_stream()
    tuned = hpxfft_amd.BatchedPlan(nx, ny, effort="measure", stream=stream)
    default = hpxfft_amd.BatchedPlan(nx, ny, effort="estimate", stream=stream)
    t_tuned = _events_ms(lambda: tuned.exec(ptrs), reps)
    t_default = _events_ms(lambda: default.exec(ptrs), reps)
    med = phase_medians(tuned, lambda: tuned.exec(ptrs), reps)
    singles = [hpxfft_amd.Plan(a, b, precision="f32", stream=stream) for a, b in zip(nx, ny)] if per_grid else []

    def run_per_grid():
        for p, x in zip(singles, xs):
            p.exec(x.data_ptr(), x.data_ptr())

    t_single = _events_ms(run_per_grid, reps) if per_grid else float("nan")
    alg = sum(a * b * 4 + a * (b // 2 + 1) * 8 for a, b in zip(nx, ny))   # 1-pass compulsory model
    H = sum(a * (b // 2 + 1) * 8 for a, b in zip(nx, ny))
    launches = {"batched_rows": (alg, med["first_fftw"]), "batched_columns": (2 * H, med["second_fftw"])}
    desc = tuned.describe()
    for p in singles:
        p.close()
    tuned.close()
    default.close()
    wl = f"batched_{count}_seed{seed}"
    return {"grids": count, "seed": seed, "alg_bytes_1pass": alg, "ms_autotuned": round(t_tuned, 4),
            "ms_default_tiles": round(t_default, 4), "ms_per_grid_launches": round(t_single, 4) if per_grid else None,
            "speedup_vs_per_grid": round(t_single / t_tuned, 3) if per_grid else None,
            "GBps_1pass": round(alg / (t_tuned * 1e-3) / 1e9, 1),
            "hbm_frac_1pass": round(alg / (t_tuned * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "launches": launch_table(launches, wl), "roofline": leg_roofline(launches, wl),
            "plan": desc.strip().splitlines()[1:3]}


def bench_adjoint(batch: int = 16, n: int = 4096, reps: int = 5) -> dict:
    """Differentiable workload: batch 16 x 2^12 x 2^12 FP32, forward and
    adjoint, through one plan of 16 grids (the torch binding's layout)."""
    import torch
    import hpxfft_amd

    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(batch, n, n, dtype=torch.float32, device="cuda", generator=g)
    X = torch.empty(batch, n, n // 2 + 1, dtype=torch.complex64, device="cuda")
    r = torch.empty_like(x)
    p = hpxfft_amd.Plan(n, n, batch=batch, ld_in=n, ld_out=n // 2 + 1, precision="f32", effort="measure",
                        stream=torch.cuda.current_stream())
    fwd = lambda: p.exec(x.data_ptr(), X.data_ptr())   # noqa: E731
    adj = lambda: p.exec_adjoint(X.data_ptr(), r.data_ptr())   # noqa: E731
    fwd()
    adj()
    t_f = _events_ms(fwd, reps)
    med_f = phase_medians(p, fwd, reps)
    t_a = _events_ms(adj, reps)
    med_a = phase_medians(p, adj, reps)
    p.close()
    b1 = alg_bytes(n, n, 4)
    B = batch * b1["total"]
    lf = pass_launches(med_f, batch * b1["row"], batch * b1["col"])
    la = pass_launches(med_a, batch * b1["row"], batch * b1["col"], backward=True, row="row_c2r")
    wl = f"adjoint_{batch}x{n}x{n}"
    return {"config": f"{batch} x {n} x {n} FP32", "fwd_ms": round(t_f, 4), "adjoint_ms": round(t_a, 4),
            "fwd_GBps": round(B / (t_f * 1e-3) / 1e9, 1), "adjoint_GBps": round(B / (t_a * 1e-3) / 1e9, 1),
            "launches_forward": launch_table(lf, wl + "_fwd"), "launches_adjoint": launch_table(la, wl + "_adj"),
            "roofline_forward": leg_roofline(lf, wl + "_fwd"), "roofline": leg_roofline(la, wl + "_adj")}


def bench_mixed(n: int = 1 << 15, reps: int = 3) -> dict:
    """2^15 x 2^15 N(0,1): FP64, FP32, FP16-storage and BF16-storage times,
    per-launch event timings and roofline entries, and the rel-L2 of each
    against the FP64 path (which the oracle pins at this size)."""
    import torch
    import hpxfft_amd

    ld = 2 * (n // 2 + 1)
    g = torch.Generator(device="cuda").manual_seed(3)
    x64 = torch.empty((n, ld), dtype=torch.float64, device="cuda")
    x64[:, :n].normal_(generator=g)
    x64[:, n:].zero_()
    out = {}
    ref = torch.empty((n, n // 2 + 1), dtype=torch.complex128, device="cuda")
    p64 = hpxfft_amd.Plan(n, n, ld_in=ld, precision="f64", effort="measure")
    run64 = lambda: p64.exec(x64.data_ptr(), ref.data_ptr())   # noqa: E731
    run64()
    t64 = _events_ms(run64, reps)
    med = phase_medians(p64, run64, reps)
    b = alg_bytes(n, n, 8)
    l64 = pass_launches(med, b["row"], b["col"])
    out["f64"] = {"ms": round(t64, 3), "GBps": round(b["total"] / (t64 * 1e-3) / 1e9, 1), "column_pass": _split_line(p64),
                  "launches": launch_table(l64, f"f64_{n}x{n}"), "roofline": leg_roofline(l64, f"f64_{n}x{n}")}
    p64.close()
    nref = torch.linalg.vector_norm(ref)
    x32 = x64.float()
    y32 = torch.empty((n, n // 2 + 1), dtype=torch.complex64, device="cuda")
    p32 = hpxfft_amd.Plan(n, n, ld_in=ld, precision="f32", effort="measure")
    run32 = lambda: p32.exec(x32.data_ptr(), y32.data_ptr())   # noqa: E731
    t32 = _events_ms(run32, reps)
    med = phase_medians(p32, run32, reps)
    b = alg_bytes(n, n, 4)
    l32 = pass_launches(med, b["row"], b["col"])
    split32 = _split_line(p32)
    p32.close()
    err = (torch.linalg.vector_norm(y32.to(torch.complex128) - ref) / nref).item()
    out["f32"] = {"ms": round(t32, 3), "rel_l2_vs_f64": err, "GBps": round(b["total"] / (t32 * 1e-3) / 1e9, 1),
                  "column_pass": split32, "launches": launch_table(l32, f"f32_{n}x{n}"),
                  "roofline": leg_roofline(l32, f"f32_{n}x{n}")}
    del x32, y32
    for name, dt, prec in (("f16", torch.float16, "f16"), ("bf16", torch.bfloat16, "bf16")):
        x16 = x64.to(dt)
        buf = torch.empty_like(x16)
        col_exp = torch.empty(n // 2 + 1, dtype=torch.int32, device="cuda")
        p = hpxfft_amd.Plan(n, n, ld_in=ld, ld_out=ld // 2, precision=prec, effort="measure")
        run = lambda: (buf.copy_(x16), p.exec_scaled(buf.data_ptr(), buf.data_ptr(), None,   # noqa: E731
                                                     col_exp.data_ptr()))
        # the copy alone first: the timed loop must be the last writer of buf
        tc = _events_ms(lambda: buf.copy_(x16), reps)
        t = _events_ms(run, reps)
        med = phase_medians(p, run, reps)
        b = alg_bytes(n, n, 2)
        l16 = pass_launches(med, b["row"], b["col"])
        split = _split_line(p)
        p.close()
        Xs = buf.view(n, ld // 2, 2).float()
        X = torch.complex(Xs[..., 0], Xs[..., 1]).to(torch.complex128) * torch.exp2(-col_exp.double())[None, :]
        e = (torch.linalg.vector_norm(X - ref) / nref).item()
        out[name] = {"ms": round(t - tc, 3), "rel_l2_vs_f64": e,
                     "GBps": round(b["total"] / ((t - tc) * 1e-3) / 1e9, 1),
                     "vs_f32_time": round((t - tc) / t32, 3),
                     "scales": "per input row after the row pass, per output column after the column pass",
                     "column_pass": split, "launches": launch_table(l16, f"{name}_{n}x{n}"),
                     "roofline": leg_roofline(l16, f"{name}_{n}x{n}")}
        del x16, buf, Xs, X
    return out


def _split_line(p) -> list:
    return [ln.strip() for ln in p.describe().splitlines() if "four-step" in ln or "autotuned" in ln]


def bench_c1(nx: int = 1 << 8, ny: int = 1 << 14, seed: int = 1, reps: int = 50, per_graph: int = 20) -> dict:
    """The reference harness's shape family read as exponents
    (shared_loop_2d.cpp:142-143, README sizes 8 and 14): 2^8 x 2^14 FP64,
    U[0,1), plan at the harness default "estimate".  34 MB in and out, so
    launch latency is a visible share: timed eagerly (one exec per step, back
    to back) and as a captured HIP graph of `per_graph` execs replayed."""
    import torch
    import hpxfft_amd

    ld = 2 * (ny // 2 + 1)
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.empty((nx, ld), dtype=torch.float64, device="cuda")
    x[:, :ny].uniform_(0.0, 1.0, generator=g)
    x[:, ny:].zero_()
    out = torch.empty((nx, ny // 2 + 1), dtype=torch.complex128, device="cuda")
    s = torch.cuda.Stream()
    p = hpxfft_amd.Plan(nx, ny, ld_in=ld, precision="f64", effort="estimate", stream=s)
    run = lambda: p.exec(x.data_ptr(), out.data_ptr())   # noqa: E731
    with torch.cuda.stream(s):
        eager = _events_ms(run, reps, warmup=5)
        med = phase_medians(p, run, 9)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            for _ in range(per_graph):
                run()
        graphed = _events_ms(graph.replay, max(3, reps // per_graph), warmup=2) / per_graph
    s.synchronize()
    desc = [ln.strip() for ln in p.describe().splitlines() if "kernel" in ln]
    p.close()
    b = alg_bytes(nx, ny, 8)
    launches = pass_launches(med, b["row"], b["col"])
    wl = f"f64_{nx}x{ny}"
    return {"config": f"{nx}x{ny} FP64 U[0,1) seed {seed}, plan estimate", "alg_bytes": b["total"],
            "ms_eager": round(eager, 5), "ms_graph": round(graphed, 5),
            "GBps_eager": round(b["total"] / (eager * 1e-3) / 1e9, 1),
            "GBps_graph": round(b["total"] / (graphed * 1e-3) / 1e9, 1),
            "hbm_frac_graph": round(b["total"] / (graphed * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "launches": launch_table(launches, wl), "roofline": leg_roofline(launches, wl), "kernels": desc}


def bench_general(nx: int = 3000, ny: int = 3000, reps: int = 10) -> dict:
    """Lengths that are not powers of two, FP64, on the general kernels
    (2-pass byte model): by default 3000 x 3000 (3000 = 2^3 3 5^3); with
    prime lengths such as 2039 x 4078 the Bluestein path; 6000 x 12000 runs
    its long lengths (radices 2..7) in place in one LDS buffer, and lengths
    beyond that such as 9000 x 18000 take the two-level path (two launches
    per pass: the roofline entry is then per pass)."""
    import torch
    import hpxfft_amd

    ld = 2 * (ny // 2 + 1)
    g = torch.Generator(device="cuda").manual_seed(6)
    x = torch.empty((nx, ld), dtype=torch.float64, device="cuda")
    x[:, :ny].uniform_(-1.0, 1.0, generator=g)
    x[:, ny:].zero_()
    out = torch.empty((nx, ny // 2 + 1), dtype=torch.complex128, device="cuda")
    p = hpxfft_amd.Plan(nx, ny, ld_in=ld, precision="f64", effort="measure")
    run = lambda: p.exec(x.data_ptr(), out.data_ptr())   # noqa: E731
    ms = _events_ms(run, reps)
    med = phase_medians(p, run, 5)
    desc = [ln.strip() for ln in p.describe().splitlines() if "general" in ln]
    p.close()
    b = alg_bytes(nx, ny, 8)
    launches = {"row_pass": (b["row"], med["first_fftw"]), "column_pass": (b["col"], med["second_fftw"])}
    wl = f"f64_{nx}x{ny}"
    return {"config": f"{nx}x{ny} FP64", "ms": round(ms, 4), "GBps": round(b["total"] / (ms * 1e-3) / 1e9, 1),
            "launches": launch_table(launches, wl), "roofline": leg_roofline(launches, wl), "kernels": desc}


def general_legs() -> dict:
    """The general-length legs: mixed radix, Bluestein, long lengths in one
    LDS buffer, and two levels."""
    return {"general_length": bench_general(),
            "general_length_bluestein": bench_general(2039, 2 * 2039),
            "general_length_long": bench_general(6000, 12000),
            "general_length_two_level": bench_general(9000, 18000)}


def cpu_baseline(nx: int, ny: int, budget_s: float = 20.0):
    """The oracle's restatement of shared::loop::fft_2d_r2c_par on host cores."""
    import oracle

    threads = len(os.sched_getaffinity(0))   # every core this process may run on
    x = oracle.fill(nx, ny, seed=2, dist="upm1")
    loop = oracle.Loop(nx, ny)
    t0 = time.perf_counter()
    loop.exec(x, nthreads=threads)   # warm-up (also first touch of the transpose buffer)
    first = time.perf_counter() - t0
    reps = max(1, min(3, int(budget_s / max(first, 1e-3)) - 1))
    ts = []
    for _ in range(reps):
        x = oracle.fill(nx, ny, seed=2, dist="upm1")
        loop.exec(x, nthreads=threads)
        ts.append(loop.measurements["total"])
    t = sorted(ts)[len(ts) // 2]
    B = alg_bytes(nx, ny, 8)["total"]
    return {"value": B / t / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"full {nx}x{ny} FP64 grid, oracle restatement of shared::loop::fft_2d_r2c_par "
                      f"(OpenMP over all {threads} affinity cores, radix-2 FFT), median of {reps} after 1 warm-up",
            "seconds": t, "phases": dict(loop.measurements)}


def cpu_library(nx: int, ny: int, reps: int = 3) -> dict:
    """A tuned CPU FFT library on the same host cores, beside the oracle's
    port: scipy.fft.rfft2 (pocketfft) with one worker per affinity core, on
    the same FP64 grid (unpadded Nx x Ny input, Nx x (Ny/2+1) output; bytes on
    the 2-pass model).  Context for the port's figure, not the baseline."""
    import numpy as np
    import scipy
    import scipy.fft

    threads = len(os.sched_getaffinity(0))
    x = np.random.default_rng(2).uniform(-1.0, 1.0, (nx, ny))
    scipy.fft.rfft2(x, workers=threads)   # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        scipy.fft.rfft2(x, workers=threads)
        ts.append(time.perf_counter() - t0)
    t = sorted(ts)[len(ts) // 2]
    B = alg_bytes(nx, ny, 8)["total"]
    return {"value": B / t / 1e9, "unit": "GB/s", "cores": threads, "seconds": t,
            "library": f"scipy.fft.rfft2 (pocketfft, scipy {scipy.__version__}), workers={threads}",
            "sample": f"full {nx}x{ny} FP64 grid, median of {reps} after 1 warm-up"}


def _hipfft_lib():
    """torch's own hipFFT (the library torch.fft calls, over rocFFT), loaded
    for direct calls: plans and executes without torch's tensor glue."""
    import ctypes
    import torch
    lib = ctypes.CDLL(os.path.join(os.path.dirname(torch.__file__), "lib", "libhipfft.so"))
    vp = ctypes.c_void_p
    lib.hipfftPlan2d.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.hipfftSetStream.argtypes = [vp, vp]
    lib.hipfftExecD2Z.argtypes = [vp, vp, vp]
    lib.hipfftExecR2C.argtypes = [vp, vp, vp]
    lib.hipfftDestroy.argtypes = [vp]
    for f in (lib.hipfftPlan2d, lib.hipfftSetStream, lib.hipfftExecD2Z, lib.hipfftExecR2C, lib.hipfftDestroy):
        f.restype = ctypes.c_int
    return lib


def bench_vendor(reps: int = 10) -> dict:
    """The same transforms through the vendor FFT library on the same GPU and
    stream, timed beside this build's plans: hipFFT (over rocFFT) called
    directly -- hipfftPlan2d + hipfftExecD2Z / hipfftExecR2C -- and, as a
    user would call it, torch.fft.rfft2 (the same library plus torch's copies).
    A runtime comparator only: neither the implementation nor the parity
    oracle (the rel-L2 between the outputs is reported, not tested).
    Layout: the unpadded one both libraries use (input Nx x Ny real, output
    Nx x (Ny/2+1) complex, out of place); bytes on the 2-pass model.  Ours
    plans at effort measure, as the headline does; the vendor plans are made
    before the timed calls."""
    import ctypes
    import torch
    import hpxfft_amd

    HIPFFT_R2C, HIPFFT_D2Z = 0x2A, 0x6A
    hf = _hipfft_lib()
    res = {"library": "hipFFT / rocFFT bundled with torch " + torch.__version__}
    stream = torch.cuda.current_stream()
    for nx, ny, prec in [(1 << 14, 1 << 14, "f64"), (1 << 14, 1 << 14, "f32"), (1 << 15, 1 << 15, "f32")]:
        dt = torch.float64 if prec == "f64" else torch.float32
        g = torch.Generator(device="cuda")
        g.manual_seed(2)
        x = torch.empty((nx, ny), dtype=dt, device="cuda").uniform_(-1.0, 1.0, generator=g)
        yv = torch.empty((nx, ny // 2 + 1), dtype=torch.complex128 if prec == "f64" else torch.complex64,
                         device="cuda")
        yo = torch.empty_like(yv)
        plan = hpxfft_amd.Plan(nx, ny, ld_in=ny, precision=prec, effort="measure", stream=stream)
        oms = _events_ms(lambda: plan.exec(x.data_ptr(), yo.data_ptr()), reps)
        plan.close()
        h = ctypes.c_void_p()
        t0 = time.perf_counter()
        if hf.hipfftPlan2d(ctypes.byref(h), nx, ny, HIPFFT_D2Z if prec == "f64" else HIPFFT_R2C) != 0:
            raise RuntimeError("hipfftPlan2d failed")
        if hf.hipfftSetStream(h, ctypes.c_void_p(stream.cuda_stream)) != 0:
            raise RuntimeError("hipfftSetStream failed")
        ex = hf.hipfftExecD2Z if prec == "f64" else hf.hipfftExecR2C

        def run_vendor():
            if ex(h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(yv.data_ptr())) != 0:
                raise RuntimeError("hipfftExec failed")

        run_vendor()
        torch.cuda.synchronize()
        vplan_s = time.perf_counter() - t0
        diff = (torch.linalg.vector_norm((yo - yv).to(torch.complex128)) /
                torch.linalg.vector_norm(yo.to(torch.complex128))).item()
        vms = _events_ms(run_vendor, reps)
        hf.hipfftDestroy(h)
        torch.fft.rfft2(x, out=yv)
        tms = _events_ms(lambda: torch.fft.rfft2(x, out=yv), reps)
        b = alg_bytes(nx, ny, 8 if prec == "f64" else 4)["total"]
        res[f"{prec}_{nx}x{ny}"] = {
            "ours_ms": round(oms, 4), "hipfft_ms": round(vms, 4), "torch_rfft2_ms": round(tms, 4),
            "ours_GBps": round(b / (oms * 1e-3) / 1e9, 1), "hipfft_GBps": round(b / (vms * 1e-3) / 1e9, 1),
            "speedup_vs_hipfft": round(vms / oms, 3), "speedup_vs_torch_rfft2": round(tms / oms, 3),
            "rel_l2_between": diff, "hipfft_plan_and_first_call_s": round(vplan_s, 3)}
        progress(f"vendor {prec} {nx}x{ny}: hipFFT {vms:.4f} ms, torch.fft.rfft2 {tms:.4f} ms, ours {oms:.4f} ms")
        del x, yv, yo
        torch.cuda.empty_cache()
    return res


def _headline_leg(precision: str, steps: int = 5, warmup: int = 2) -> dict:
    n = 1 << 14
    r = time_gpu(n, n, precision, steps, warmup, seed=2)
    s = 8 if precision == "f64" else 4
    lt = launches(n, n, s, r)
    wl = f"{precision}_{n}x{n}"
    return {"ms_per_step": round(r["ms_per_step"], 4), "launches": launch_table(lt, wl),
            "roofline": leg_roofline(lt, wl)}


# the legs --only can run on their own (each prints the object the full run nests under its name)
LEGS = {
    "f64_16384x16384": lambda: _headline_leg("f64"),
    "f32_16384x16384": lambda: _headline_leg("f32"),
    "readme_shape_2e8x2e14": bench_c1,
    "batched_uneven": bench_batched,
    "batched_uneven_plans_only": lambda: bench_batched(per_grid=False),
    "adjoint": bench_adjoint,
    "mixed_2e15": bench_mixed,
    "general_length": bench_general,
    "general_length_bluestein": lambda: bench_general(2039, 2 * 2039),
    "general_length_long": lambda: bench_general(6000, 12000),
    "general_length_two_level": lambda: bench_general(9000, 18000),
    "vendor_rocfft": bench_vendor,
    "cpu_library": lambda: cpu_library(1 << 14, 1 << 14),
}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=1 << 14)
    ap.add_argument("--ny", type=int, default=1 << 14)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f32", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched / adjoint / 2^15 mixed-precision legs")
    ap.add_argument("--general-only", action="store_true",
                    help="only the lengths that are not powers of two (FP64, plans at measure): one JSON line")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (reads BYTES, writes BYTES) three times")
    ap.add_argument("--only", choices=sorted(LEGS), default=None,
                    help="run one leg and print its JSON object (for profiling a leg on its own)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the 2-D r2c path runs on a single MI355X (no multi-GPU path is built)",
              file=sys.stderr)
        return 2
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 2
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.general_only:
        print(json.dumps(general_legs()), flush=True)
        return 0
    if args.only:
        print(json.dumps({args.only: LEGS[args.only]()}), flush=True)
        return 0
    nx, ny = args.nx, args.ny
    r64 = time_gpu(nx, ny, "f64", args.steps, args.warmup, seed=2)
    progress(f"f64 {nx}x{ny}: {r64['ms_per_step']:.4f} ms/step")
    b64 = alg_bytes(nx, ny, 8)
    gbs = b64["total"] / (r64["ms_per_step"] * 1e-3) / 1e9
    # dominant kernel: the pass with the longest event-timed duration on the
    # plan's stream.  A four-step column pass is one kernel (tile_kernel) in
    # two launches, phases A and B, credited H each: per launch, its
    # algorithmic bytes, duration and traffic are the two launches' means.
    per_launch = launches(nx, ny, 8, r64)
    workload = f"f64_{nx}x{ny}"
    tr = measured_traffic(workload)
    groups = {"row_r2c": ["row_r2c"],
              "column_pass": [k for k in per_launch if k.startswith("column_")]}
    dom = max(groups, key=lambda g: sum(per_launch[k][1] for k in groups[g]))
    keys = groups[dom]
    dom_bytes = sum(per_launch[k][0] for k in keys) // len(keys)
    dom_s = sum(per_launch[k][1] for k in keys) / len(keys)
    ach = dom_bytes / dom_s / 1e9
    traffic = sum(tr[k] for k in keys) // len(keys) if all(k in tr for k in keys) else None
    dom_name = dom if len(keys) == 1 else f"{dom} ({' + '.join(keys)}, mean per launch)"
    result = {
        "metric": "2D r2c FFT GB/s & %HBM peak at 2^14^2, FP32 vs FP64 at stated L2 error",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: uniform [-1,1) from a seeded generator, resident in HBM",
        "config": {"workload": f"2-D r2c FFT {nx}x{ny} FP64 (out of place, padded input layout)",
                   "nx": nx, "ny": ny, "parallelism": "single GPU"},
        "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
        "gflops": round(nominal_flops(nx, ny) / (r64["ms_per_step"] * 1e-3) / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom_name,
                     "alg_bytes_per_launch": dom_bytes, "avg_launch_s": dom_s},
        "launches_s": {k: v[1] for k, v in per_launch.items()},
        # every launch: algorithmic bytes, HBM traffic (committed counter summary) and both rates
        "launches": {k: {"s": v[1], "alg_bytes": v[0], "alg_GBps": round(v[0] / v[1] / 1e9, 1),
                         "traffic": measured_traffic(workload).get(k),
                         "traffic_GBps": (round(measured_traffic(workload)[k] / v[1] / 1e9, 1)
                                          if measured_traffic(workload).get(k) else None)}
                     for k, v in per_launch.items()},
        "plan": {"effort": "measure", "seconds": round(r64["plan_s"], 4),
                 "column_pass": [ln.strip() for ln in r64["plan"].splitlines() if "four-step" in ln or "autotuned" in ln]},
        "phases_s": {"row_r2c": r64["row_s"], "column_c2c": r64["col_s"]},
    }
    result["fp64_vector_frac"] = round(result["gflops"] / (FP64_VEC_TFLOPS * 1e3), 4)
    if not args.no_f32:
        r32 = time_gpu(nx, ny, "f32", args.steps, args.warmup, seed=2)
        progress(f"f32 {nx}x{ny}: {r32['ms_per_step']:.4f} ms/step")
        b32 = alg_bytes(nx, ny, 4)
        # FP32 error against the FP64 result of the same input (the FP64 path matches the oracle to 1e-12)
        x64 = r64["x"]
        import hpxfft_amd
        x32 = x64.float()
        out32 = torch.empty((nx, ny // 2 + 1), dtype=torch.complex64, device="cuda")
        p32 = hpxfft_amd.Plan(nx, ny, ld_in=2 * (ny // 2 + 1), precision="f32")
        p32.exec(x32.data_ptr(), out32.data_ptr())
        ref = r64["out"]
        x64d = x32.double()
        p64 = hpxfft_amd.Plan(nx, ny, ld_in=2 * (ny // 2 + 1), precision="f64")
        p64.exec(x64d.data_ptr(), ref.data_ptr())   # FP64 transform of the FP32-rounded input
        torch.cuda.synchronize()
        err = (torch.linalg.vector_norm(out32.to(torch.complex128) - ref) / torch.linalg.vector_norm(ref)).item()
        p32.close()
        p64.close()
        l32 = launches(nx, ny, 4, r32)
        result["f32"] = {"ms_per_step": round(r32["ms_per_step"], 4),
                         "GBps": round(b32["total"] / (r32["ms_per_step"] * 1e-3) / 1e9, 2),
                         "speedup_vs_f64": round(r64["ms_per_step"] / r32["ms_per_step"], 3),
                         "rel_l2_vs_f64": err,
                         "gflops": round(nominal_flops(nx, ny) / (r32["ms_per_step"] * 1e-3) / 1e9, 1),
                         "phases_s": {"row_r2c": r32["row_s"], "column_c2c": r32["col_s"]},
                         "launches": launch_table(l32, f"f32_{nx}x{ny}"),
                         "roofline": leg_roofline(l32, f"f32_{nx}x{ny}")}
    if not args.no_extra:
        torch.cuda.empty_cache()
        result["readme_shape_2e8x2e14"] = bench_c1()
        progress("2^8 x 2^14 done")
        result["batched_uneven"] = bench_batched()
        progress("batched uneven grids done")
        torch.cuda.empty_cache()
        result["adjoint"] = bench_adjoint()
        progress("adjoint done")
        torch.cuda.empty_cache()
        result["mixed_2e15"] = bench_mixed()
        progress("2^15 mixed precision done")
        torch.cuda.empty_cache()
        result.update(general_legs())
        progress("general length done")
        torch.cuda.empty_cache()
        try:
            result["vendor_rocfft"] = bench_vendor()
        except Exception as e:   # an external library: its absence must not cost the bench line
            result["vendor_rocfft"] = {"error": f"{type(e).__name__}: {e}"}
        progress("vendor comparator done")
        torch.cuda.empty_cache()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(nx, ny)
        progress("cpu baseline done")
        try:
            result["cpu_library"] = cpu_library(nx, ny)
        except Exception as e:   # an external library: its absence must not cost the bench line
            result["cpu_library"] = {"error": f"{type(e).__name__}: {e}"}
        progress("cpu library comparator done")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
