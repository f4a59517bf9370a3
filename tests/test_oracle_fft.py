"""Pins the oracle's 3D real FFT (Grid_FFT::FourierTransformForward/Backward,
grid_fft.cc:207-270, unitary 1/sqrt(N) both ways, grid_fft.cc:69, 224, 257)
against numpy's pocketfft fixtures and exact identities."""
import numpy as np
import pytest

from oracle import oracle as O

GOLD = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "fft.npz"))


def _field(shape):
    n0, n1, n2 = shape
    a = np.zeros((n0, n1, n2 + 2))
    return a


def _r2c(x):
    n0, n1, n2 = x.shape
    a = _field(x.shape)
    a[:, :, :n2] = x
    O.lib().orc_fft3d_r2c(a, n0, n1, n2)
    return a.view(np.complex128).reshape(n0, n1, n2 // 2 + 1).copy()


def _c2r(k, n2):
    n0, n1, _ = k.shape
    a = np.zeros((n0, n1, n2 + 2))
    a.view(np.complex128).reshape(k.shape)[...] = k
    O.lib().orc_fft3d_c2r(a, n0, n1, n2)
    return a[:, :, :n2].copy()


@pytest.mark.parametrize("name", ["n8", "n12", "n16", "r6x10x12"])
def test_r2c_matches_pocketfft_fixture(name):
    x, k = GOLD[name + "_x"], GOLD[name + "_k"]
    got = _r2c(x)
    assert np.max(np.abs(got - k)) < 1e-13 * np.max(np.abs(k)) * 10


@pytest.mark.parametrize("name", ["n8", "n12", "n16", "r6x10x12"])
def test_c2r_matches_pocketfft_fixture(name):
    x, k = GOLD[name + "_x"], GOLD[name + "_k"]
    got = _c2r(k, x.shape[2])
    assert np.max(np.abs(got - x)) < 1e-13 * np.max(np.abs(x)) * 10


@pytest.mark.parametrize("n", [6, 8, 10, 12, 18, 24, 32])
def test_identities(n):
    rng = np.random.default_rng(n)
    # delta <-> constant 1/sqrt(N^3)
    d = np.zeros((n, n, n))
    d[0, 0, 0] = 1.0
    k = _r2c(d)
    assert np.allclose(k, 1.0 / np.sqrt(n ** 3), rtol=0, atol=1e-15)
    # single mode <-> cosine: x = cos(2 pi (a i + b j + c k)/n) has k-space
    # value sqrt(N^3)/2 at (a, b, c) (c < n/2)
    a, b, c = 1, n - 2, 2
    i, j, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    cosx = np.cos(2 * np.pi * (a * i + b * j + c * kk) / n)
    kc = _r2c(cosx)
    expect = np.zeros_like(kc)
    expect[a, b, c] = np.sqrt(n ** 3) / 2
    assert np.max(np.abs(kc - expect)) < 1e-12 * np.sqrt(n ** 3)
    # Parseval (unitary) and the round trip
    x = rng.standard_normal((n, n, n))
    k = _r2c(x)
    w = np.full(k.shape[2], 2.0)
    w[0] = 1.0
    if n % 2 == 0:
        w[-1] = 1.0
    assert abs(np.sum(w * np.abs(k) ** 2) - np.sum(x * x)) < 1e-11 * np.sum(x * x)
    assert np.max(np.abs(_c2r(k, n) - x)) < 1e-13 * 10
    # Hermitian symmetry of the kz = 0 plane
    p = k[:, :, 0]
    assert np.max(np.abs(p - np.conj(p[(-np.arange(n)) % n][:, (-np.arange(n)) % n]))) < 1e-13


def test_forward_sign_convention():
    # FFTW_FORWARD = exp(-i k x): a shift by +1 cell multiplies mode k by exp(-2 pi i k / n)
    n = 8
    rng = np.random.default_rng(3)
    x = rng.standard_normal((n, n, n))
    k0 = _r2c(x)
    k1 = _r2c(np.roll(x, 1, axis=2))
    ph = np.exp(-2j * np.pi * np.arange(n // 2 + 1) / n)
    assert np.max(np.abs(k1 - k0 * ph[None, None, :])) < 1e-12


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16, 18, 24, 36, 48, 64, 96, 384])
def test_fft1d_matches_numpy(n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got_f = O.fft1d(z.copy(), -1)
    got_b = O.fft1d(z.copy(), +1)
    assert np.max(np.abs(got_f - np.fft.fft(z))) < 1e-13 * n
    assert np.max(np.abs(got_b - np.fft.ifft(z) * n)) < 1e-13 * n
