"""Tabulated transfer functions (the CAMB_file plugin, transfer_CAMB_file.cc):
the oracle's log-log natural spline against scipy's, its sigma_8 against the
Eisenstein & Hu one it was sampled from, the host side of the library against
the oracle, and the CAMB file reader."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from monofonic_amd import _lib as L, lpt, transfer
from oracle import oracle as O


def wiggly_table(n=600, kmin=1e-5, kmax=1e2):
    """E&H (baseline cosmology) times a smooth wiggle, so that the table
    differs from the fit it starts from"""
    c = O.baseline_cosmology(24.0)
    k = np.logspace(np.log10(kmin), np.log10(kmax), n)
    T = np.array([O.lib().orc_tk_eh(c, x) for x in k]) * (1.0 + 0.03 * np.sin(3.0 * np.log(k)))
    return k, T


BASE = dict(Omega_m=0.3, Omega_b=0.045, H0=70.0, sigma_8=0.81, n_s=0.961, zstart=24.0)


def test_oracle_table_spline_matches_scipy():
    k, T = wiggly_table(n=80)
    c = O.cosmology(transfer=(k, T), **BASE)
    sp = CubicSpline(np.log(k), np.log(T), bc_type="natural")
    kq = np.exp(np.random.default_rng(1).uniform(np.log(k[1]), np.log(k[-1]), 200))
    got = np.array([O.lib().orc_tk(c, x) for x in kq])
    assert np.allclose(got, np.exp(sp(np.log(kq))), rtol=1e-12, atol=0)
    # below the second row k is clamped to it (transfer_CAMB_file.cc:195-199)
    for x in (k[0], 0.3 * k[1], k[1]):
        assert O.lib().orc_tk(c, x) == pytest.approx(float(np.exp(sp(np.log(k[1])))), rel=1e-13)
    # the nodes themselves
    assert np.allclose([O.lib().orc_tk(c, x) for x in k[1:]], T[1:], rtol=1e-13, atol=0)


def test_oracle_table_sigma8_reproduces_eh():
    c_eh = O.baseline_cosmology(24.0)
    k = np.logspace(-5, 5, 4000)
    T = np.array([O.lib().orc_tk_eh(c_eh, x) for x in k])
    c_tab = O.cosmology(transfer=(k, T), **BASE)
    assert c_tab.sigma8_measured == pytest.approx(c_eh.sigma8_measured, rel=1e-7)
    assert c_tab.pnorm == pytest.approx(c_eh.pnorm, rel=2e-7)


def test_library_table_matches_oracle():
    k, T = wiggly_table()
    p = transfer.with_table(lpt.baseline_cosmology(24.0), k, T)
    c = O.cosmology(transfer=(k, T), **BASE)
    kq = np.array([3e-6, 1e-4, 3e-3, 0.05, 0.7, 5.0, 40.0])
    got = lpt.cosmo_eval(p, "T", kq)
    want = np.array([O.lib().orc_tk(c, x) for x in kq])
    assert np.allclose(got, want, rtol=1e-12, atol=0)
    s = lpt.cosmo_scalars(p)
    assert s.sigma8_measured == pytest.approx(c.sigma8_measured, rel=1e-10)
    assert s.pnorm == pytest.approx(c.pnorm, rel=1e-10)
    # the table is in use: the wiggle moves sigma_8 away from E&H
    s_eh = lpt.cosmo_scalars(lpt.baseline_cosmology(24.0))
    assert abs(s.sigma8_measured / s_eh.sigma8_measured - 1) > 1e-3


@pytest.mark.parametrize("bad", ["short", "decreasing", "negative"])
def test_library_rejects_bad_tables(bad):
    k, T = wiggly_table(n=40)
    if bad == "short":
        k, T = k[:5], T[:5]
    elif bad == "decreasing":
        k = k.copy()
        k[10], k[11] = k[11], k[10]
    else:
        T = T.copy()
        T[7] = -T[7]
    p = transfer.with_table(lpt.baseline_cosmology(24.0), k, T)
    with pytest.raises(L.MficError) as e:
        lpt.cosmo_scalars(p)
    assert e.value.code == -6


def test_camb_file_reader(tmp_path):
    k, T = wiggly_table(n=12)
    rows = []
    for i, (kk, tt) in enumerate(zip(k, T)):
        cols = [kk, 2.0 * tt, 0.5 * tt, 7.0, 8.0, 9.0, tt, 1.0, 1.0, 1.0, -tt, -2 * tt, tt]
        rows.append(" ".join(f"{x:.17e}" for x in cols))
    f = tmp_path / "camb_transfer.dat"
    f.write_text("# k cdm baryon photon nu mass_nu total no_nu total_de Weyl v_cdm v_b v_b-v_cdm\n" +
                 "\n".join(rows) + "\n")
    kr, Tr = transfer.read_camb_file(str(f))
    assert np.array_equal(kr, k) and np.array_equal(Tr, T)
    bad = tmp_path / "bad.dat"
    bad.write_text(rows[0] + "\n1.0 2.0 3.0\n")
    with pytest.raises(RuntimeError):
        transfer.read_camb_file(str(bad))
