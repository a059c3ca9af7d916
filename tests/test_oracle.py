"""Pins the CPU FP64 oracle (oracle/oracle.c) before anything is compared to it.

Each check is chosen so that a plausible mistake fails it: a wrong sign
(conjugate spectrum), a transposed index, a dropped term, a wrong twiddle
stride or a wrong r2c split.  test_planted_mistakes_are_caught applies such
mistakes to correct outputs and asserts the checks reject them.

Pins, per oracle function:
  r2c_2d / Loop  -- the reference's 4x6 known answer (exact, as
                    test/src/test_shared_loop.cpp:53), long-double brute force,
                    numpy.fft.rfft2 fixtures, ramp closed form, deltas,
                    constants, Parseval, linearity, shift theorem, real-valued
                    self-conjugate bins.
  fft_c2c_1d / fft_r2c_1d -- brute force and numpy.
  adjoint_r2c_2d -- brute force, <Rx, Y> = <x, R*Y>, torch autograd fixtures.
  fill           -- SplitMix64 reference values and distribution moments.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

G = np.load(os.path.join(GOLDEN, "rfft2_golden.npz"))


def rl2(a, b):
    return oracle.rel_l2(a, b)


# --------------------------------------------------------------------------- generator
def test_splitmix64_known_values():
    # SplitMix64 seeded with 0: first outputs 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4 (published test vector)
    a = oracle.fill(1, 2, seed=0, dist="u01")
    expect = np.array([0xe220a8397b1dcdaf >> 11, 0x6e789e6aa1b965f4 >> 11], dtype=np.float64) * 2.0 ** -53
    assert np.array_equal(a[0, :2], expect)
    assert np.all(a[0, 2:] == 0.0)   # padding


def test_distributions_moments():
    u = oracle.fill(256, 256, 5, "upm1")[:, :256]
    assert -1.0 <= u.min() and u.max() < 1.0 and abs(u.mean()) < 0.01 and abs(u.var() - 1 / 3) < 0.01
    n = oracle.fill(256, 256, 6, "normal")[:, :256]
    assert abs(n.mean()) < 0.01 and abs(n.std() - 1.0) < 0.01
    r = oracle.fill(4, 8, 0, "ramp")
    assert np.array_equal(r[:, :8], np.tile(np.arange(8.0), (4, 1)))


# --------------------------------------------------------------------------- 2-D r2c
def test_reference_known_answer_4x6_exact():
    """test/src/test_shared_loop.cpp:15-34,53 -- exact equality."""
    x = G["ref4x6_in"].copy()
    loop = oracle.Loop(4, 4)
    loop.exec(x)
    assert np.array_equal(x, G["ref4x6_out"])
    assert set(loop.measurements) == {"total", "first_fftw", "first_trans", "second_fftw", "second_trans"}
    assert all(v >= 0 for v in loop.measurements.values())


@pytest.mark.parametrize("nx,ny", [(1, 2), (2, 2), (2, 4), (4, 8), (8, 4), (8, 16), (16, 8), (16, 16)])
def test_brute_force_long_double(nx, ny):
    x = oracle.fill(nx, ny, 100 + nx * ny, "normal")[:, :ny].copy()
    assert rl2(oracle.r2c_2d(x), oracle.dft_r2c_2d_brute(x)) < 1e-14


@pytest.mark.parametrize("key", [k[:-3] for k in G.files if k.startswith("np_") and k.endswith("_in")])
def test_numpy_fixtures(key):
    x = G[key + "_in"]
    assert rl2(oracle.r2c_2d(x), G[key + "_out"]) < 1e-14


def ramp_closed_form(nx, ny):
    ky = np.arange(1, ny // 2 + 1)
    row0 = np.empty(ny // 2 + 1, dtype=np.complex128)
    row0[0] = nx * ny * (ny - 1) / 2
    row0[1:] = nx * (-ny / 2 + 1j * (ny / 2) / np.tan(np.pi * ky / ny))
    row0[-1] = -nx * ny / 2
    ref = np.zeros((nx, ny // 2 + 1), dtype=np.complex128)
    ref[0] = row0
    return ref


@pytest.mark.parametrize("nx,ny", [(4, 4), (8, 16), (64, 1024), (256, 4096)])
def test_ramp_closed_form(nx, ny):
    """examples/hpxfft/shared_loop_2d.cpp:33-40 input x[i][j] = j."""
    got = oracle.r2c_2d(oracle.fill(nx, ny, 0, "ramp"), ny)
    assert rl2(got, ramp_closed_form(nx, ny)) < 1e-14


@pytest.mark.parametrize("nx,ny,a,b", [(8, 16, 3, 5), (32, 64, 31, 63), (16, 128, 0, 7)])
def test_delta(nx, ny, a, b):
    x = np.zeros((nx, ny))
    x[a, b] = 1.0
    kx = np.arange(nx)[:, None]
    ky = np.arange(ny // 2 + 1)[None, :]
    ph = ((kx * a) % nx) / nx + ((ky * b) % ny) / ny   # phase reduced exactly before exp
    ref = np.exp(-2j * np.pi * ph)
    assert np.max(np.abs(oracle.r2c_2d(x) - ref)) < 5e-14


def test_constant_gives_dc_spike():
    x = np.full((32, 64), 2.5)
    X = oracle.r2c_2d(x)
    assert abs(X[0, 0] - 2.5 * 32 * 64) < 1e-10
    X[0, 0] = 0
    assert np.max(np.abs(X)) < 1e-10


def parseval_ok(x, X, ny):
    w = np.full(X.shape[1], 2.0)
    w[0] = 1.0
    w[-1] = 1.0 if ny % 2 == 0 else 2.0
    lhs = np.sum(x * x)
    rhs = np.sum(w[None, :] * np.abs(X) ** 2) / x.size
    return abs(lhs - rhs) <= 1e-12 * lhs


@pytest.mark.parametrize("nx,ny", [(16, 32), (128, 256), (2, 2048)])
def test_parseval_linearity_shift(nx, ny):
    x = oracle.fill(nx, ny, 7, "normal")[:, :ny].copy()
    y = oracle.fill(nx, ny, 8, "upm1")[:, :ny].copy()
    X, Y = oracle.r2c_2d(x), oracle.r2c_2d(y)
    assert parseval_ok(x, X, ny)
    assert rl2(oracle.r2c_2d(2.0 * x - 3.0 * y), 2.0 * X - 3.0 * Y) < 1e-14
    # circular shift by (s, t): multiply by exp(-2 pi i (kx s/nx + ky t/ny))
    s, t = 3 % nx, 5 % ny
    kx = np.arange(nx)[:, None]
    ky = np.arange(ny // 2 + 1)[None, :]
    ph = np.exp(-2j * np.pi * (kx * s / nx + ky * t / ny))
    assert rl2(oracle.r2c_2d(np.roll(x, (s, t), axis=(0, 1))), X * ph) < 1e-13


def test_self_conjugate_bins_are_real():
    nx, ny = 64, 128
    X = oracle.r2c_2d(oracle.fill(nx, ny, 9, "normal"), ny)
    assert X[0, 0].imag == 0.0
    for kx, ky in ((nx // 2, 0), (0, ny // 2), (nx // 2, ny // 2)):
        assert abs(X[kx, ky].imag) < 1e-12 * np.abs(X).max()


def test_1d_transforms_match_brute_force_and_numpy():
    rng = np.random.default_rng(3)
    for n in (1, 2, 4, 8, 64, 1024):
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        k = np.arange(n)
        W = np.exp(-2j * np.pi * (np.outer(k, k) % n) / n)   # exponent reduced exactly
        assert rl2(oracle.fft_c2c_1d(z, -1), W @ z) < 1e-13
        assert rl2(oracle.fft_c2c_1d(z, +1), np.conj(W) @ z) < 1e-13
    for n in (2, 4, 16, 4096):
        x = rng.standard_normal(n)
        assert rl2(oracle.fft_r2c_1d(x), np.fft.rfft(x)) < 1e-14


def test_twiddles_exact_on_axes():
    assert oracle.twiddle(0, 8) == 1 + 0j
    assert oracle.twiddle(2, 8) == 0 - 1j
    assert oracle.twiddle(4, 8) == -1 + 0j
    assert oracle.twiddle(6, 8) == 0 + 1j
    for m, n in ((1, 12), (5, 16), (1000, 65536)):
        assert abs(oracle.twiddle(m, n) - np.exp(-2j * np.pi * m / n)) < 2e-16


# --------------------------------------------------------------------------- adjoint
@pytest.mark.parametrize("nx,ny", [(4, 8), (8, 16), (16, 4)])
def test_adjoint_torch_fixtures_and_brute(nx, ny):
    Y = G[f"adj_{nx}x{ny}_Y"]
    ref = G[f"adj_{nx}x{ny}_RstarY"]
    assert np.max(np.abs(oracle.adjoint_r2c_2d(Y, ny) - ref)) < 1e-12 * np.abs(ref).max()
    assert np.max(np.abs(oracle.adjoint_brute(Y, ny) - ref)) < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n", [3, 5, 6, 7, 9, 12, 14, 15, 30, 49, 97, 210, 1155, 1031, 2039, 4099, 6000, 8186,
                               131072])
def test_general_length_c2c_matches_numpy(n):
    """Mixed-radix lengths (any n, primes up to 4099 among them, the lengths
    the GPU's Bluestein and two-level paths take): forward and backward
    against numpy.fft."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    tol = 1e-14 if n < 1024 else 1e-13
    assert np.max(np.abs(oracle.fft_c2c_1d(x) - ref)) <= tol * np.max(np.abs(ref))
    assert np.max(np.abs(oracle.fft_c2c_1d(x, +1) - n * np.fft.ifft(x))) <= tol * np.max(np.abs(ref))


@pytest.mark.parametrize("nx,ny", [(8, 14), (3, 10), (5, 6), (7, 30), (6, 4), (1, 2), (9, 18)])
def test_general_size_r2c_brute_and_numpy(nx, ny):
    """Non-power-of-two grids (the reference harness's default 8 x 14 among
    them) against the long-double brute force and numpy.fft.rfft2."""
    x = oracle.fill(nx, ny, seed=nx * 31 + ny, dist="normal")[:, :ny].copy()
    got = oracle.r2c_2d(x)
    assert rl2(got, oracle.dft_r2c_2d_brute(x)) < 1e-14
    assert rl2(got, np.fft.rfft2(x)) < 1e-14


@pytest.mark.parametrize("nx,ny", [(1, 2), (2, 2), (1, 8), (4, 8), (16, 4), (8, 16), (64, 128), (8, 14), (5, 6),
                                   (12, 30)])
def test_c2r_matches_numpy_irfft2(nx, ny):
    """Inverse c2r of an arbitrary (non-Hermitian) half spectrum: Nx*Ny *
    numpy.fft.irfft2, which also drops the imaginary parts of the ky = 0 and
    ky = Ny/2 bins after the column transform; and c2r(r2c(x)) = Nx*Ny*x."""
    rng = np.random.default_rng(nx * 1000 + ny)
    X = rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))
    got = oracle.c2r_2d(X, ny)
    ref = nx * ny * np.fft.irfft2(X, s=(nx, ny))
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    # the adjoint (interior bins not doubled) is a different map: the check can fail
    if ny >= 4:
        assert np.max(np.abs(oracle.adjoint_r2c_2d(X, ny) - ref)) > 1e-3 * np.max(np.abs(ref))
    x = oracle.fill(nx, ny, 9, "normal")[:, :ny].copy()
    back = oracle.c2r_2d(oracle.r2c_2d(x), ny)
    assert np.max(np.abs(back - nx * ny * x)) <= 1e-12 * nx * ny * np.max(np.abs(x))


@pytest.mark.parametrize("nx,ny", [(2, 2), (32, 64), (256, 512)])
def test_adjoint_inner_product_identity(nx, ny):
    rng = np.random.default_rng(nx + ny)
    x = rng.standard_normal((nx, ny))
    Y = rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))
    lhs = np.real(np.vdot(oracle.r2c_2d(x), Y))
    rhs = np.dot(x.ravel(), oracle.adjoint_r2c_2d(Y, ny).ravel())
    assert abs(lhs - rhs) <= 1e-12 * (np.abs(x).sum() * np.abs(Y).max())


# --------------------------------------------------------------------------- planted mistakes
def checks_pass(x, X):
    """The size-independent checks above, applied to a candidate spectrum X of real x."""
    nx, ny = x.shape
    ok = rl2(X, oracle.dft_r2c_2d_brute(x)) < 1e-12
    ok = ok and parseval_ok(x, X, ny)
    ok = ok and rl2(X, np.fft.rfft2(x)) < 1e-12
    return ok


def wrong_split(x):
    """r2c via the half-length trick with w_N^{+k} instead of w_N^{-k} in the split."""
    nx, ny = x.shape
    L = ny // 2
    z = x[:, 0::2] + 1j * x[:, 1::2]
    Z = np.fft.fft(z, axis=1)
    k = np.arange(L + 1)
    A = Z[:, k % L]
    B = np.conj(Z[:, (L - k) % L])
    Fe, Fo = (A + B) / 2, (A - B) / 2j
    rows = Fe + np.exp(+2j * np.pi * k / ny)[None, :] * Fo
    return np.fft.fft(rows, axis=0)


def test_planted_mistakes_are_caught():
    nx, ny = 8, 16
    x = oracle.fill(nx, ny, 77, "normal")[:, :ny].copy()
    X = oracle.r2c_2d(x)
    assert checks_pass(x, X)
    mistakes = {
        "conjugate (wrong sign)": np.conj(X),
        "kx reversed (transposed/flipped index)": np.roll(X[::-1], 1, axis=0),
        "dropped imaginary part": X.real.astype(np.complex128),
        "missing Nyquist column": np.concatenate([X[:, :-1], np.zeros((nx, 1))], axis=1),
        "r2c split with conjugated twiddle": wrong_split(x),
        "ramp row-major/col-major swap": X * np.exp(-2j * np.pi * np.arange(ny // 2 + 1)[None, :] / ny),
    }
    for name, bad in mistakes.items():
        assert not checks_pass(x, bad), name
    # a closed-form check catches a dropped term in the ramp case
    R = oracle.r2c_2d(oracle.fill(8, 16, 0, "ramp"), 16)
    bad = R.copy()
    bad[0, 1:-1] = bad[0, 1:-1].real   # dropped cot term
    assert rl2(bad, ramp_closed_form(8, 16)) > 1e-3 and rl2(R, ramp_closed_form(8, 16)) < 1e-14
