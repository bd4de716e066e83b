"""Pins the oracle's 3-D real FFT (A2a_k / A_k2a, fft_wrapper.h:432-631).

Three independent references:
  - the reference's own Ooura rdft3d + rdft3dsort, compiled from its
    sources into oracle/_ref (skipped where that build is absent);
  - numpy.fft.rfftn mapped to the Ooura sign: Re = Re X, Im = -Im X;
  - a brute-force sum over exp(+i k.x) on an 8^3 grid.
"""
import numpy as np
import pytest

from kapsel_old_amd import abi
from oracle.oracle import Oracle, ref_A2a_k, ref_A_k2a, ref_available
from tests.helpers import real_field

GRIDS = [(8, 8, 8), (16, 8, 32), (32, 32, 32), (64, 16, 8)]


def numpy_spectrum(a, nz):
    X = np.fft.rfftn(a[:, :, :nz])
    out = np.zeros(a.shape)
    out[:, :, 0::2] = X.real
    out[:, :, 1::2] = -X.imag
    return out


@pytest.mark.parametrize("grid", GRIDS)
def test_forward_matches_numpy(grid):
    o = Oracle(abi.make_params(*grid))
    a = real_field(o.shape, grid[2], seed=1)
    k = o.A2a_k(a)
    ref = numpy_spectrum(a, grid[2])
    assert np.abs(k - ref).max() / np.abs(ref).max() < 1e-14


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref (reference Ooura build) absent")
@pytest.mark.parametrize("grid", GRIDS)
def test_forward_and_inverse_match_reference_ooura(grid):
    nz = grid[2]
    o = Oracle(abi.make_params(*grid))
    a = real_field(o.shape, nz, seed=2)
    k_ref = ref_A2a_k(a)
    k = o.A2a_k(a)
    assert np.abs(k - k_ref).max() / np.abs(k_ref).max() < 1e-14
    # inverse on a spectrum the reference produced (Hermitian-consistent)
    b_ref = ref_A_k2a(k_ref)[:, :, :nz]
    b = o.A_k2a(k_ref)[:, :, :nz]
    assert np.abs(b - b_ref).max() / np.abs(b_ref).max() < 1e-14
    assert np.abs(b - a[:, :, :nz]).max() < 1e-13


def test_brute_force_dft_8():
    n = 8
    o = Oracle(abi.make_params(n))
    a = real_field(o.shape, n, seed=3)
    x = np.arange(n)
    k = o.A2a_k(a)
    rng = np.random.default_rng(0)
    for _ in range(40):
        k1, k2, k3 = rng.integers(0, n), rng.integers(0, n), rng.integers(0, n // 2 + 1)
        ph = np.exp(2j * np.pi * (k1 * x[:, None, None] + k2 * x[None, :, None] + k3 * x[None, None, :]) / n)
        X = (a[:, :, :n] * ph).sum()
        assert abs(k[k1, k2, 2 * k3] - X.real) < 1e-12
        assert abs(k[k1, k2, 2 * k3 + 1] - X.imag) < 1e-12


@pytest.mark.parametrize("grid", GRIDS)
def test_round_trip_and_parseval(grid):
    nz = grid[2]
    o = Oracle(abi.make_params(*grid))
    a = real_field(o.shape, nz, seed=4)
    k = o.A2a_k(a)
    b = o.A_k2a(k)
    assert np.abs(b[:, :, :nz] - a[:, :, :nz]).max() < 1e-13
    assert np.all(b[:, :, nz:] == 0.0)
    # Parseval on the half spectrum: weight 2 except the kz = 0 and NZ/2 planes
    c = k[:, :, 0::2] + 1j * k[:, :, 1::2]
    w = np.full(c.shape[2], 2.0)
    w[0] = w[-1] = 1.0
    e_k = (np.abs(c) ** 2 * w).sum() / np.prod(grid)
    assert abs(e_k - (a[:, :, :nz] ** 2).sum()) / e_k < 1e-12


def test_single_mode_sign_convention():
    """cos + sin mode lands in the slot the +i convention predicts."""
    n = 16
    o = Oracle(abi.make_params(n))
    x = np.arange(n)
    a = np.zeros(o.shape)
    a[:, :, :n] = np.sin(2 * np.pi * 3 * x[None, None, :] / n) * np.ones((n, n, 1))
    k = o.A2a_k(a)
    # sum sin(2 pi 3 z/n) exp(+2 pi i 3 z/n) = i n/2 per (x, y) line
    assert abs(k[0, 0, 7] - n ** 3 / 2) < 1e-9 and abs(k[0, 0, 6]) < 1e-9


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref (reference Ooura build) absent")
@pytest.mark.parametrize("grid", GRIDS)
def test_inverse_of_a_non_hermitian_spectrum_matches_reference_ooura(grid):
    """A derivative's Nyquist modes make a spectrum that is not Hermitian on
    the kz = 0 and kz = NZ/2 planes; the reference's inverse reads only one
    member of each (k, -k) pair there.  The oracle must read it the same
    way, and a plain conjugate-symmetric average (numpy's irfftn) must not
    pass."""
    nx, ny, nz = grid
    o = Oracle(abi.make_params(nx, ny, nz))
    rng = np.random.default_rng(17)
    X = rng.standard_normal(o.shape)
    want = ref_A_k2a(X)[:, :, :nz]
    got = o.A_k2a(X)[:, :, :nz]
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    avg = np.fft.irfftn(X[:, :, 0::2] - 1j * X[:, :, 1::2], s=(nx, ny, nz), axes=(0, 1, 2))
    assert np.abs(avg - want).max() > 1e-3 * np.abs(want).max()


# Grids with factors 3 and 5 (the library's mixed-radix lengths; the
# reference's Ooura transform takes powers of two only, so numpy is the
# reference here).
MIXED = [(48, 48, 48), (80, 16, 96), (16, 96, 80), (8, 8, 192), (160, 8, 16), (8, 192, 8)]


@pytest.mark.parametrize("grid", MIXED)
def test_mixed_forward_matches_numpy(grid):
    o = Oracle(abi.make_params(*grid))
    a = real_field(o.shape, grid[2], seed=5)
    k = o.A2a_k(a)
    ref = numpy_spectrum(a, grid[2])
    assert np.abs(k - ref).max() / np.abs(ref).max() < 1e-13


@pytest.mark.parametrize("grid", MIXED)
def test_mixed_inverse_matches_numpy_and_round_trips(grid):
    nx, ny, nz = grid
    o = Oracle(abi.make_params(*grid))
    a = real_field(o.shape, nz, seed=6)
    spec = numpy_spectrum(a, nz)            # Hermitian: the spectrum of a real field
    b = o.A_k2a(spec.copy())
    assert np.abs(b[:, :, :nz] - a[:, :, :nz]).max() < 1e-13 * np.abs(a).max()
    assert np.all(b[:, :, nz:] == 0.0)


def test_mixed_brute_force_dft_48():
    nx, ny, nz = 48, 8, 80
    o = Oracle(abi.make_params(nx, ny, nz))
    a = real_field(o.shape, nz, seed=7)
    k = o.A2a_k(a)
    rng = np.random.default_rng(1)
    x, y, z = np.arange(nx), np.arange(ny), np.arange(nz)
    for _ in range(30):
        k1, k2, k3 = rng.integers(0, nx), rng.integers(0, ny), rng.integers(0, nz // 2 + 1)
        ph = np.exp(2j * np.pi * (k1 * x[:, None, None] / nx + k2 * y[None, :, None] / ny
                                  + k3 * z[None, None, :] / nz))
        X = (a[:, :, :nz] * ph).sum()
        assert abs(k[k1, k2, 2 * k3] - X.real) < 1e-11
        assert abs(k[k1, k2, 2 * k3 + 1] - X.imag) < 1e-11


@pytest.mark.parametrize("n", [12, 100, 7, 256 * 3, 2048])
def test_oracle_rejects_unsupported_lengths(n):
    with pytest.raises(Exception):
        Oracle(abi.make_params(n, 8, 8))
