"""Pins the oracle's cosmology (cosmology_calculator.hh:63-223, 334-432,
487-511; transfer_eisenstein.cc:58-201): growth D(a), f(a) against a scipy
DOP853 integration of the same ODE (fixture), the Einstein-de Sitter closed
form, the log-log natural spline against scipy's CubicSpline, sigma_8
against scipy quad, and the Eisenstein & Hu (1998) fit against an
evaluation written here from the paper's formulae."""
import json
import math
import os

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.interpolate import CubicSpline

from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cosmo.json")))

CASES = {
    "benchmark_zero_radiation": dict(Omega_m=0.3, Omega_b=0.045, H0=70.0, sigma_8=0.81, n_s=0.961,
                                     zero_radiation=1),
    "eds": dict(Omega_m=1.0, Omega_b=0.045, H0=70.0, sigma_8=0.81, n_s=0.961, zero_radiation=1,
                m_nu=(0.0, 0.0, 0.0), N_ur=0.0),
    "benchmark_radiation": dict(Omega_m=0.3, Omega_b=0.045, H0=70.0, sigma_8=0.81, n_s=0.961),
}


@pytest.mark.parametrize("case", list(CASES))
def test_growth_matches_ode_fixture(case):
    c = O.cosmology(**CASES[case])
    L = O.lib()
    for a, D, f in GOLD[case]["growth"]:
        assert abs(L.orc_cosmo_D(c, a) / D - 1) < 2e-7, (a, L.orc_cosmo_D(c, a), D)
        assert abs(L.orc_cosmo_f(c, a) / f - 1) < 2e-6, (a, L.orc_cosmo_f(c, a), f)
    assert abs(c.Omega_gamma / GOLD[case]["Omega_gamma"] - 1) < 1e-12
    assert abs(c.Omega_DE - GOLD[case]["Omega_DE"]) < 1e-12


def test_eds_closed_form():
    c = O.cosmology(**CASES["eds"])
    L = O.lib()
    for a in (0.01, 0.02, 0.1, 0.5, 1.0):
        assert abs(L.orc_cosmo_D(c, a) / a - 1) < 1e-8
        assert abs(L.orc_cosmo_f(c, a) - 1) < 1e-8
        # H(a) = H0 a^{-3/2}, vfact = f a H / h
        assert abs(L.orc_cosmo_H(c, a) / (70.0 * a ** -1.5) - 1) < 1e-14
        assert abs(L.orc_cosmo_vfact(c, a) / (a * 70.0 * a ** -1.5 / 0.7) - 1) < 1e-8


def test_spline_is_natural_cubic_in_loglog():
    c = O.baseline_cosmology(24.0)
    L = O.lib()
    cap = 100000
    a = np.zeros(cap)
    D = np.zeros(cap)
    f = np.zeros(cap)
    n = L.orc_cosmo_growth_table(c, a, D, f, cap)
    a, D, f = a[:n], D[:n], f[:n]
    sD = CubicSpline(np.log(a), np.log(D), bc_type="natural")
    sf = CubicSpline(np.log(a), np.log(f), bc_type="natural")
    xs = np.exp(0.5 * (np.log(a[:-1]) + np.log(a[1:])))[:: max(1, n // 200)]
    for x in xs:
        assert abs(L.orc_cosmo_D(c, x) * c.Dnow / math.exp(sD(math.log(x))) - 1) < 1e-11
        assert abs(L.orc_cosmo_f(c, x) / math.exp(sf(math.log(x))) - 1) < 1e-11


def test_sigma8_integral_and_normalisation():
    c = O.baseline_cosmology(24.0)
    L = O.lib()

    def integrand(k):
        x = 8.0 * k
        w = 1.0 - 0.1 * x * x if x < 1e-3 else 3.0 * (math.sin(x) - x * math.cos(x)) / x ** 3
        t = L.orc_tk_eh(c, k)
        return k * k * w * w * k ** 0.961 * t * t

    edges = np.logspace(-4, 4, 81)
    ref = sum(quad(integrand, lo, hi, epsabs=1e-30, epsrel=1e-12, limit=400)[0] for lo, hi in zip(edges[:-1], edges[1:]))
    got = L.orc_integrate_sigma8(c, 1e-4, 1e4, 1e-13)
    assert abs(got / ref - 1) < 1e-9
    assert abs(c.sigma8_measured / math.sqrt(4 * math.pi * ref) - 1) < 1e-9
    assert abs(c.pnorm * c.sigma8_measured ** 2 / 0.81 ** 2 - 1) < 1e-12
    # amplitude A(k) = k^{n_s/2} T(k) sqrt(pnorm)
    for k in (1e-3, 0.05, 1.0):
        assert abs(L.orc_amplitude(c, k) / (k ** 0.4805 * L.orc_tk_eh(c, k) * math.sqrt(c.pnorm)) - 1) < 1e-14


def eh98(k_hmpc, Om, Ob, h, Tcmb):
    """Eisenstein & Hu 1998 (ApJ 496, 605) baryon+CDM fit, eqs. 2-24; the
    2.718282 literal is the one of the authors' published code."""
    omhh = Om * h * h
    fb = Ob / Om
    obhh = omhh * fb
    th = Tcmb / 2.7
    zeq = 2.50e4 * omhh / th ** 4
    keq = 0.0746 * omhh / th ** 2
    b1 = 0.313 * omhh ** -0.419 * (1 + 0.607 * omhh ** 0.674)
    b2 = 0.238 * omhh ** 0.223
    zd = 1291 * omhh ** 0.251 / (1 + 0.659 * omhh ** 0.828) * (1 + b1 * obhh ** b2)
    Rd = 31.5 * obhh / th ** 4 * (1000 / (1 + zd))
    Req = 31.5 * obhh / th ** 4 * (1000 / zeq)
    s = 2 / (3 * keq) * math.sqrt(6 / Req) * math.log((math.sqrt(1 + Rd) + math.sqrt(Rd + Req)) / (1 + math.sqrt(Req)))
    ksilk = 1.6 * obhh ** 0.52 * omhh ** 0.73 * (1 + (10.4 * omhh) ** -0.95)
    a1 = (46.9 * omhh) ** 0.670 * (1 + (32.1 * omhh) ** -0.532)
    a2 = (12.0 * omhh) ** 0.424 * (1 + (45.0 * omhh) ** -0.582)
    alpha_c = a1 ** -fb * a2 ** (-fb ** 3)
    bb1 = 0.944 / (1 + (458 * omhh) ** -0.708)
    bb2 = (0.395 * omhh) ** -0.0266
    beta_c = 1 / (1 + bb1 * ((1 - fb) ** bb2 - 1))
    y = zeq / (1 + zd)
    G = y * (-6 * math.sqrt(1 + y) + (2 + 3 * y) * math.log((math.sqrt(1 + y) + 1) / (math.sqrt(1 + y) - 1)))
    alpha_b = 2.07 * keq * s * (1 + Rd) ** -0.75 * G
    beta_node = 8.41 * omhh ** 0.435
    beta_b = 0.5 + fb + (3 - 2 * fb) * math.sqrt((17.2 * omhh) ** 2 + 1)

    k = k_hmpc * h
    q = k / 13.41 / keq
    xx = k * s
    lnb = math.log(2.718282 + 1.8 * beta_c * q)
    lnnb = math.log(2.718282 + 1.8 * q)
    C_a = 14.2 / alpha_c + 386 / (1 + 69.9 * q ** 1.08)
    C_na = 14.2 + 386 / (1 + 69.9 * q ** 1.08)
    fcut = 1 / (1 + (xx / 5.4) ** 4)
    Tc = fcut * lnb / (lnb + C_na * q * q) + (1 - fcut) * lnb / (lnb + C_a * q * q)
    st = s / (1 + (beta_node / xx) ** 3) ** (1 / 3)
    xt = k * st
    T0 = lnnb / (lnnb + C_na * q * q)
    Tb = (T0 / (1 + (xx / 5.2) ** 2) + alpha_b / (1 + (beta_b / xx) ** 3) * math.exp(-(k / ksilk) ** 1.4)) \
        * math.sin(xt) / xt
    return fb * Tb + (1 - fb) * Tc


def test_eisenstein_hu_fit():
    c = O.baseline_cosmology(24.0)
    L = O.lib()
    for k in np.logspace(-4, 2, 49):
        ref = eh98(k, 0.3, 0.045, 0.7, 2.7255)
        assert abs(L.orc_tk_eh(c, k) - ref) < 1e-12 * max(abs(ref), 1e-6), k
    assert abs(L.orc_tk_eh(c, 1e-6) - 1) < 1e-6  # T(k -> 0) = 1


def test_derived_parameters():
    c = O.baseline_cosmology(24.0)
    # Omega_nu from m_nu = 0.06 eV (cosmology_parameters.hh:177-216)
    assert abs(c.Omega_nu_massive - 0.06 / (93.13861 * 0.49)) < 1e-4
    assert abs(c.Omega_c - (0.3 - 0.045 - c.Omega_nu_massive)) < 1e-15
    assert abs(c.Omega_DE - (1 - 0.3 - c.Omega_r)) < 1e-15
    assert c.Omega_r > 0
    assert abs(c.Dplus_start - O.lib().orc_cosmo_D(c, 1 / 25.0)) < 1e-15
