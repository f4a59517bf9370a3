"""Generate the golden fixtures that pin the CPU oracle.

Every value here comes from an implementation independent of the oracle:
  fft.npz        numpy.fft.rfftn(norm="ortho") (pocketfft) on random fields
                 of sizes 8^3, 12^3, 16^3, 6x10x12
  crmath.npz     mpmath (200-bit) log / sin / cos of RANLUX-like inputs and
                 edge cases, rounded to double
  cosmo.json     scipy solve_ivp (DOP853, rtol 1e-12) of the linear growth ODE
                 in conformal time (cosmology_calculator.hh:103-135), D(a)/D(1)
                 and f = dlnD/dlna at several a, for the benchmark cosmology
                 and an Einstein-de Sitter check; Omega_gamma from the
                 physical constants in physical_constants.hh
The reference monofonIC itself is not run (it cannot be built here: FFTW3
and GSL are absent).

Usage: python tests/golden/make_golden.py   (rewrites the files next to it)
"""
import json
import os

import mpmath
import numpy as np
from scipy.integrate import solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))


def make_fft():
    rng = np.random.default_rng(20261202)
    out = {}
    for name, shape in (("n8", (8, 8, 8)), ("n12", (12, 12, 12)), ("n16", (16, 16, 16)), ("r6x10x12", (6, 10, 12))):
        x = rng.standard_normal(shape)
        out[name + "_x"] = x
        out[name + "_k"] = np.fft.rfftn(x, norm="ortho")
    np.savez_compressed(os.path.join(HERE, "fft.npz"), **out)


def make_crmath():
    mpmath.mp.prec = 200
    rng = np.random.default_rng(7)
    # RANLUX-like: multiples of 2^-48 in (0, 1)
    u = rng.integers(1, 2 ** 48, size=3000, dtype=np.int64).astype(np.float64) * 2.0 ** -48
    u = np.concatenate([u, [2.0 ** -48, 1 - 2.0 ** -48, 0.5, 0.25, 0.75, 1e-10, 0.999999]])
    ph = np.concatenate([u * 2 * np.pi, [0.0, np.pi / 2, np.pi, 1.5 * np.pi, 2 * np.pi - 1e-12]])
    lg = np.array([float(mpmath.log(mpmath.mpf(float(x)))) for x in u])
    sn = np.array([float(mpmath.sin(mpmath.mpf(float(x)))) for x in ph])
    cs = np.array([float(mpmath.cos(mpmath.mpf(float(x)))) for x in ph])
    np.savez_compressed(os.path.join(HERE, "crmath.npz"), u=u, log=lg, phase=ph, sin=sn, cos=cs)


# physical constants (physical_constants.hh:28-72)
PI = 3.141592653589793115997963468544185161590576171875
MPC, C_SI, G_SI = 3.0857e22, 2.99792458e8, 6.6740800e-11
KB, HBAR = 1.38064852e-23, 1.054571800e-34
SIGMA_SB = (PI * PI) * KB ** 4 / 60.0 / HBAR ** 3 / C_SI ** 2
RHOC = 3 * 1e10 / (8 * PI * G_SI) / MPC / MPC


def growth(Om, Ob, H0, mnu, Nur, Tcmb, zero_rad, a_eval):
    h = H0 / 100.0
    Onu = mnu / (93.13861 * h * h)
    Og = 4 * SIGMA_SB / C_SI ** 3 * Tcmb ** 4 / RHOC / h ** 2
    Or = 0.0 if zero_rad else Og + Nur * Og * 7 / 8 * (4 / 11) ** (4 / 3)
    Ode = 1.0 - Om - Or

    def H(a):
        return H0 * np.sqrt(Or / a ** 4 + Om / a ** 3 + Ode)

    # D'' + a H D' - 3/2 Om H0^2 D / a = 0, da/dtau = a^2 H; integrate in ln a
    # with y = (D, dD/dlna): dD/dlna = D'/(aH)
    def rhs(lna, y):
        a = np.exp(lna)
        Hv = H(a)
        # d ln H / d ln a
        dlnH = -0.5 * (4 * Or / a ** 4 + 3 * Om / a ** 3) / (Or / a ** 4 + Om / a ** 3 + Ode)
        D, Dp = y
        return [Dp, -(2 + dlnH) * Dp + 1.5 * Om * H0 ** 2 / (a ** 3 * Hv ** 2) * D]

    # the reference's initial condition (cosmology_calculator.hh:99-103):
    # a0 = 1e-10, D0 = a0, dD/dtau = 2 D0 H(a0) / c^2  ->  dD/dlna = 2 / c^2
    a0 = 1e-10
    y0 = [a0, 2.0 / C_SI ** 2]
    sol = solve_ivp(rhs, (np.log(a0), np.log(2.0)), y0, method="DOP853", rtol=1e-12, atol=1e-40,
                    dense_output=True)
    D1 = sol.sol(0.0)[0]
    res = []
    for a in a_eval:
        D, Dp = sol.sol(np.log(a))
        res.append((a, D / D1, Dp / D))
    return res, dict(Omega_gamma=Og, Omega_r=Or, Omega_DE=Ode, Omega_nu=Onu)


def make_cosmo():
    a_eval = [1 / 50.0, 1 / 25.0, 0.1, 0.3, 0.5, 1.0]
    out = {}
    res, extra = growth(0.3, 0.045, 70.0, 0.06, 2.046, 2.7255, True, a_eval)
    out["benchmark_zero_radiation"] = dict(growth=res, **extra)
    res, extra = growth(1.0, 0.045, 70.0, 0.0, 0.0, 2.7255, True, a_eval)
    out["eds"] = dict(growth=res, **extra)
    res, extra = growth(0.3, 0.045, 70.0, 0.06, 2.046, 2.7255, False, a_eval)
    out["benchmark_radiation"] = dict(growth=res, **extra)
    with open(os.path.join(HERE, "cosmo.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    make_fft()
    make_crmath()
    make_cosmo()
    print("golden fixtures written to", HERE)
