"""Writes tests/golden/rfft2_golden.npz (run from the repo root).

Fixtures (inputs together with expected outputs, all float64/complex128):

* ``ref4x6_*``: the reference's only known-answer test,
  test/src/test_shared_loop.cpp:15-34 (the same values appear in
  test_shared_sync.cpp, test_shared_opt.cpp, test_naive.cpp,
  test_shared_agas.cpp): four rows [1,2,3,4] plus two padding doubles;
  expected row 0 = (40,0) (-8,8) (-8,0), all other rows 0; compared with
  exact equality there (vector_2d operator==).
* ``np_<dist>_<nx>x<ny>_*``: SplitMix64-seeded inputs (the oracle's
  generator) and numpy.fft.rfft2 (pocketfft, FP64) spectra -- a library
  routine independent of the oracle, standing in for FFTW, which is absent.
* ``adj_<nx>x<ny>_*``: random half-spectra Y and the gradient of
  Re<rfft2(x), Y> w.r.t. x from torch.fft.rfft2 autograd in FP64 (CPU), i.e.
  the adjoint R*Y of an independent implementation.
"""
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)
import oracle  # noqa: E402

CASES = [("upm1", 8, 16, 11), ("u01", 32, 64, 12), ("normal", 64, 128, 13), ("upm1", 256, 256, 14),
         ("upm1", 1, 32, 15), ("normal", 16, 2, 16), ("u01", 128, 8, 17)]
ADJ = [(4, 8, 21), (8, 16, 22), (16, 4, 23)]


def main():
    out = {}
    x = np.zeros((4, 6))
    x[:, :4] = [1.0, 2.0, 3.0, 4.0]
    exp = np.zeros((4, 6))
    exp[0, 0], exp[0, 2], exp[0, 3], exp[0, 4] = 40.0, -8.0, 8.0, -8.0
    out["ref4x6_in"] = x
    out["ref4x6_out"] = exp
    for dist, nx, ny, seed in CASES:
        a = oracle.fill(nx, ny, seed, dist)[:, :ny].copy()
        key = f"np_{dist}_{nx}x{ny}"
        out[key + "_seed"] = np.array([seed], dtype=np.int64)
        out[key + "_in"] = a
        out[key + "_out"] = np.fft.rfft2(a)
    import torch
    for nx, ny, seed in ADJ:
        rng = np.random.default_rng(seed)
        Y = rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))
        xt = torch.zeros((nx, ny), dtype=torch.float64, requires_grad=True)
        Yt = torch.from_numpy(Y)
        loss = torch.real(torch.sum(torch.conj(torch.fft.rfft2(xt)) * Yt))
        loss.backward()
        key = f"adj_{nx}x{ny}"
        out[key + "_Y"] = Y
        out[key + "_RstarY"] = xt.grad.numpy().copy()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rfft2_golden.npz")
    np.savez_compressed(path, **out)
    print(path, sorted(out))


if __name__ == "__main__":
    main()
