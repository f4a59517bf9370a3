"""Pins the oracle's complex FFT and semiclassical PT fields
(ic_generator.cc:786-897):

* the c2c transform equals numpy.fft.fftn / ifftn(norm="ortho");
* the drift is unitary: the mean of delta = |psi|^2 - 1 vanishes;
* hbar follows the reference's neighbour-jump loop over the first n-1 x-planes;
* the drift conserves the momentum sum_x (1 + delta) v = vunit hbar
  sum_k k |psi0_k|^2 of the initial wave function (LPTorder 1);
* on large scales delta and v are the linear (Zel'dovich) density k^2 phi
  and velocity vfac grad phi of the growth-scaled potential, which a sign
  error in the phase or the drift turns into anti-correlation.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("shape", [(8, 12, 6), (16, 16, 16), (9, 18, 24)])
def test_c2c_matches_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert np.abs(O.fft_c2c(z, -1) - np.fft.fftn(z, norm="ortho")).max() < 1e-14 * np.abs(z).max() * 10
    assert np.abs(O.fft_c2c(z, +1) - np.fft.ifftn(z, norm="ortho")).max() < 1e-14 * np.abs(z).max() * 10


def _run(n, box, z, order, seed, sign=1.0):
    cfg = O.ic_cfg(n, box, order, seed)
    oc = O.baseline_cosmology(z)
    phi, phi2, _, _ = O.potentials(cfg, oc, O.ngenic_fill(seed, n))
    coef = O.lpt_coefs(cfg, oc)
    vu = 3.5
    (delta, vx, vy, vz), hbar = O.ppt_fields(cfg, coef, sign * phi, None if phi2 is None else sign * phi2, vu)
    return cfg, coef, phi, phi2, [O.rview(a) for a in (delta, vx, vy, vz)], hbar, vu


def _kgrid(n, box):
    m = np.arange(n)
    kk = np.where(m > n // 2, m - n, m) * (2 * math.pi / box)
    return np.meshgrid(kk, kk, kk, indexing="ij")


def test_unitarity_and_hbar():
    n, box = 32, 100.0
    cfg, coef, phi, _, (delta, *_), hbar, _ = _run(n, box, 49.0, 2, 12345)
    assert abs(delta.mean()) < 1e-12
    ph = np.fft.irfftn(O.kview(phi), s=(n, n, n), axes=(0, 1, 2), norm="ortho")
    jumps = [np.abs(np.roll(ph, -1, axis=a) - ph)[: n - 1] for a in range(3)]
    want = max(j.max() for j in jumps) / math.pi / coef.Dplus0 * 1.01
    assert abs(hbar / want - 1) < 1e-12


def test_drift_conserves_momentum():
    n, box = 32, 100.0
    cfg, coef, phi, _, (delta, vx, vy, vz), hbar, vu = _run(n, box, 49.0, 1, 777)
    D = coef.Dplus0
    ph = np.fft.irfftn(O.kview(phi), s=(n, n, n), axes=(0, 1, 2), norm="ortho")
    psi0 = np.exp(1j * ph / (hbar * D))
    pk = np.abs(np.fft.fftn(psi0, norm="ortho")) ** 2
    K = _kgrid(n, box)
    vunit = D * coef.vfac / box * vu
    for d, v in enumerate((vx, vy, vz)):
        got = np.sum((1.0 + delta) * v)
        want = vunit * hbar * np.sum(K[d] * pk)
        assert abs(got - want) <= 1e-9 * np.sum(np.abs((1.0 + delta) * v)), d


def _large_scale(f, n, kcut):
    k = np.fft.fftn(f)
    K = np.sqrt(sum(x ** 2 for x in np.meshgrid(*[np.fft.fftfreq(n, 1 / n)] * 3, indexing="ij")))
    k[K > kcut] = 0
    return np.real(np.fft.ifftn(k))


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_linear_limit(sign):
    """sign = -1 plants a phase-sign mistake: the fields flip against the
    Zel'dovich ones"""
    n, box = 32, 128.0
    cfg, coef, phi, _, (delta, vx, vy, vz), hbar, vu = _run(n, box, 99.0, 1, 4242, sign)
    K = _kgrid(n, box)
    kc = O.kview(phi)
    k2 = (K[0] ** 2 + K[1] ** 2 + K[2] ** 2)[:, :, : n // 2 + 1]
    dlin = np.fft.irfftn(k2 * kc, s=(n, n, n), axes=(0, 1, 2), norm="ortho")
    corr = lambda a, b: float(np.sum(a * b) / np.sqrt(np.sum(a * a) * np.sum(b * b)))  # noqa: E731
    a, b = _large_scale(delta, n, n / 8), _large_scale(dlin, n, n / 8)
    c = corr(a, b)
    ratio = np.std(a) / np.std(b)
    print(f"sign {sign}: delta corr {c:.4f}, amplitude ratio {ratio:.4f}")
    vcs = []
    for d, v in enumerate((vx, vy, vz)):
        g = np.fft.irfftn(1j * K[d][:, :, : n // 2 + 1] * kc, s=(n, n, n), axes=(0, 1, 2), norm="ortho") * coef.vfac / box * vu
        vcs.append(corr(_large_scale(v, n, n / 8), _large_scale(g, n, n / 8)))
    print(f"sign {sign}: velocity corr {vcs}")
    if sign > 0:
        assert c > 0.99 and 0.95 < ratio < 1.05 and min(vcs) > 0.99
    else:
        assert c < -0.9 and max(vcs) < -0.9
