"""Distinct plans used from several host threads at once (C ABI contract:
any host thread may call; calls on one plan are serialised by the caller;
distinct plans are independent).  Each thread plans its own size, precision
and stream, runs it repeatedly, and must match a result computed alone.
Also: the thread-local last-error message stays per thread.
"""
import threading

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

CASES = [(1024, 2048, "f64"), (2048, 512, "f32"), (300, 3000, "f64"), (4096, 64, "f32"), (96, 134, "f64"),
         (6000, 10, "f64")]


def _run(nx, ny, prec, x, stream):
    import torch
    import hpxfft_amd
    d = torch.from_numpy(x).to(device="cuda", dtype=torch.float64 if prec == "f64" else torch.float32)
    out = torch.empty_like(d)
    p = hpxfft_amd.Plan(nx, ny, precision=prec, stream=stream)
    with torch.cuda.stream(stream):
        for _ in range(3):
            p.exec(d.data_ptr(), out.data_ptr())
    stream.synchronize()
    p.close()
    return out.cpu().numpy()


def test_plans_on_concurrent_threads(gpu):
    import torch
    xs = [oracle.fill(nx, ny, seed=i, dist="upm1") for i, (nx, ny, _) in enumerate(CASES)]
    alone = [_run(nx, ny, prec, x, torch.cuda.Stream()) for (nx, ny, prec), x in zip(CASES, xs)]
    results = [None] * len(CASES)
    errors = []

    def work(i):
        try:
            nx, ny, prec = CASES[i]
            results[i] = _run(nx, ny, prec, xs[i], torch.cuda.Stream())
        except Exception as e:   # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(CASES))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for i, (got, ref) in enumerate(zip(results, alone)):
        assert np.array_equal(got, ref), CASES[i]


def test_last_error_is_per_thread(gpu):
    import hpxfft_amd
    msgs = {}

    def bad(name, nx, ny):
        try:
            hpxfft_amd.Plan(nx, ny, precision="f64")
        except RuntimeError:
            pass
        msgs[name] = hpxfft_amd.last_error()

    bad("main", 0, 16)
    t1 = threading.Thread(target=bad, args=("odd", 16, 15))
    t2 = threading.Thread(target=bad, args=("prime", 1, 2 * 4194319))   # a prime above 2^22
    t1.start()
    t1.join()
    t2.start()
    t2.join()
    assert "even Ny" in msgs["odd"], msgs
    assert "4194319" in msgs["prime"], msgs
    assert hpxfft_amd.last_error() == msgs["main"] and "Nx >= 1" in msgs["main"], msgs
