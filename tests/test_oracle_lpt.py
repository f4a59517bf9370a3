"""Pins the oracle's LPT coefficients, assembly and lattice readout
(ic_generator.cc:258-273, 378-383, 509-516, 916-1060; particle_generator.hh:
499-544, 597-639; grid_fft.hh:832-842):

* growth coefficients g1 = -D, g2 = -3/7 D^2, g3 = D^3/3, g3c = D^3/7 and
  vfac_n = n vfac (ic_generator.cc:263-273), volfac = (L / N / 2 pi)^1.5;
* a single-mode white noise gives the analytic Zel'dovich displacement and
  velocity (phi2 = phi3 = A3 = 0 for a plane wave);
* simple-cubic readout is q + Psi in exact arithmetic; the bcc second
  sub-lattice equals the field's Fourier series evaluated at (ijk + 1/2) dx.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_coefficients():
    cfg = O.ic_cfg(64, 100.0, 3, 1)
    c = O.baseline_cosmology(49.0)
    k = O.lpt_coefs(cfg, c)
    D = O.lib().orc_cosmo_D(c, 1 / 50.0)
    assert k.Dplus0 == c.Dplus_start and abs(k.Dplus0 / D - 1) < 1e-15
    assert k.g1 == -D and k.g2 == -3.0 / 7.0 * D * D
    assert abs(k.g3 - D ** 3 / 3) < 1e-18 and abs(k.g3c - D ** 3 / 7) < 1e-18
    assert k.vfac == c.vfac_start and k.vfac2 == 2 * k.vfac and k.vfac3 == 3 * k.vfac
    assert abs(k.vfac / O.lib().orc_cosmo_vfact(c, 1 / 50.0) - 1) < 1e-14
    assert abs(k.volfac - (100.0 / 64 / 2 / math.pi) ** 1.5) < 1e-15


@pytest.mark.parametrize("order", [1, 2, 3])
def test_plane_wave_zeldovich(order):
    n, L = 16, 80.0
    cfg = O.ic_cfg(n, L, order, 7, lunit=L, vunit=2.0 * L)
    c = O.baseline_cosmology(24.0)
    co = O.lpt_coefs(cfg, c)
    wn = O.empty_field(n)
    q = (2, 3, 1)  # one mode with kz > 0 (no Hermitian partner stored)
    O.kview(wn)[q] = 0.8 - 0.35j
    pots = O.potentials(cfg, c, wn)
    if order > 1:
        assert np.max(np.abs(O.kview(pots[1]))) < 1e-12 * np.max(np.abs(O.kview(pots[0])))
    kf = 2 * math.pi / L
    kv = np.array(q, float) * kf
    k2 = kv @ kv
    A = O.lib().orc_amplitude(c, math.sqrt(k2))
    phi_q = -(O.kview(wn)[q] / co.volfac) * A / k2 * co.g1  # phi1 scaled by g1
    x = np.arange(n) * L / n
    X = np.stack(np.meshgrid(x, x, x, indexing="ij"))
    phase = np.exp(1j * np.tensordot(kv, X, axes=1))
    for comp in range(6):
        d = comp % 3
        kc = O.component_k(cfg, co, pots, comp)
        O.fft_c2r(kc)
        fac = (cfg.lunit if comp < 3 else cfg.vunit * co.vfac1) / L
        # unitary c2r of a single kz > 0 mode: 2 Re(v e^{ikx}) / sqrt(N^3)
        want = 2 * np.real(1j * kv[d] * phi_q * fac * phase) / n ** 1.5
        assert np.max(np.abs(O.rview(kc) - want)) < 1e-12 * np.max(np.abs(want))


def _band_limited(n, seed):
    rng = np.random.default_rng(seed)
    a = O.empty_field(n)
    O.rview(a)[...] = rng.standard_normal((n, n, n))
    O.fft_r2c(a)
    h = n // 2
    k = O.kview(a)
    k[h], k[:, h], k[:, :, h] = 0, 0, 0
    spec = k.copy()
    O.fft_c2r(a)
    return a, spec


def test_readout_sc_and_bcc():
    n, L, lunit = 8, 20.0, 3.0
    f, spec = _band_limited(n, 11)
    base = O.rview(f).copy()
    # simple cubic: pos = (ijk / n) lunit + Psi
    pos = np.zeros(3 * n ** 3)
    O.lib().orc_lattice_set_positions(n, L, 0, 1, lunit, f, pos)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    assert np.array_equal(pos[1::3], ((j / n) * lunit + base).reshape(-1))
    # bcc: second sub-lattice from the Fourier series at (ijk + 1/2) dx
    f2 = O.empty_field(n)
    O.rview(f2)[...] = base
    pos2 = np.zeros(3 * 2 * n ** 3)
    O.lib().orc_lattice_set_positions(n, L, 1, 0, lunit, f2, pos2)
    m = np.arange(n)
    sm = np.where(m > n // 2, m - n, m)
    kk = np.arange(n // 2 + 1)
    dx = L / n
    want = np.zeros((n, n, n))
    pts = (np.stack([i, j, k]).reshape(3, -1).T + 0.5) * dx
    kf = 2 * math.pi / L
    KX, KY, KZ = np.meshgrid(sm * kf, sm * kf, kk * kf, indexing="ij")
    w = np.where(KZ > 0, 2.0, 1.0)  # Hermitian partners of kz > 0
    for p, xyz in enumerate(pts):
        e = np.exp(1j * (KX * xyz[0] + KY * xyz[1] + KZ * xyz[2]))
        want.reshape(-1)[p] = np.real(np.sum(w * spec * e)) / n ** 1.5
    got = pos2[3 * n ** 3::3].reshape(n, n, n) - ((i + 0.5) / n) * lunit
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    assert np.array_equal(pos2[: 3 * n ** 3: 3].reshape(n, n, n), (i / n) * lunit + base)


def _series(spec, n, L, pts):
    """the band-limited field's Fourier series at points (npts, 3) [length]"""
    m = np.arange(n)
    sm = np.where(m > n // 2, m - n, m)
    kk = np.arange(n // 2 + 1)
    kf = 2 * math.pi / L
    KX, KY, KZ = np.meshgrid(sm * kf, sm * kf, kk * kf, indexing="ij")
    w = np.where(KZ > 0, 2.0, 1.0)  # Hermitian partners of kz > 0
    out = np.zeros(len(pts))
    for p, xyz in enumerate(pts):
        e = np.exp(1j * (KX * xyz[0] + KY * xyz[1] + KZ * xyz[2]))
        out[p] = np.real(np.sum(w * spec * e)) / n ** 1.5
    return out


SUBS = {0: [(0, 0, 0)], 1: [(0, 0, 0), (0.5, 0.5, 0.5)],
        2: [(0, 0, 0), (0, 0.5, 0.5), (0.5, 0, 0.5), (0.5, 0.5, 0)]}
SECOND = {0: (0.5, 0.5, 0.5), 1: (0.5, 0.5, 0.0), 2: (0.5, 0.5, 0.5)}


@pytest.mark.parametrize("lattice", [0, 1, 2])
@pytest.mark.parametrize("index", [0, 1])
def test_lattice_readout_series(lattice, index):
    """lattice index 0 and 1 (the baryon species, particle_generator.hh:
    47-63, 499-544): sub-lattice s reads the field's Fourier series at
    (ijk + shift_s + second) dx and places q there, shift_s the sc / bcc /
    fcc sub-lattice offsets, second = (1/2,1/2,1/2) for sc and fcc and
    (1/2,1/2,0) for bcc on index 1 (positions not wrapped); a planted offset
    (the sc shift on bcc) fails"""
    n, L, lunit = 8, 20.0, 3.0
    f, spec = _band_limited(n, 12 + lattice + 3 * index)
    base = O.rview(f).copy()
    second = np.array(SECOND[lattice]) if index else np.zeros(3)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ijk = np.stack([i, j, k]).reshape(3, -1).T
    dx = L / n
    for idim in range(3):
        g = O.empty_field(n)
        O.rview(g)[...] = base
        out = O.lattice_readout_index(g, L, lattice, index, idim, True, lunit)
        for s, sub in enumerate(SUBS[lattice]):
            off = second + np.array(sub)
            want = _series(spec, n, L, (ijk + off) * dx)
            got = out[s * n ** 3:(s + 1) * n ** 3, idim]
            q = (ijk[:, idim] + off[idim]) / n * lunit
            assert np.max(np.abs(got - q - want)) < 1e-12 * np.max(np.abs(want)), (idim, s)
            if lattice == 1 and index == 1 and s == 0 and idim == 2:
                wrong = _series(spec, n, L, (ijk + 0.5) * dx)
                assert np.max(np.abs(got - q - wrong)) > 1e-3 * np.max(np.abs(want))
