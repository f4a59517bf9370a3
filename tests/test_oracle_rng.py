"""Pins the oracle's white-noise generator: RANLUX ranlxd (gsl_rng_ranlxd1,
random_ngenic.cc:44-45) against GSL's published known answers, the
correctly rounded log/sin/cos against 200-bit mpmath values, and the NGenIC
seed table / per-column draw order (random_ngenic.cc:49-154) against a
restatement written here from the reference text plus structural invariants.

The GSL known answers are the values its own test suite checks after 10000
gsl_rng_get calls from seed 1: ranlxd1 -> 1998227290, ranlxd2 -> 3949287736.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

CR = np.load(os.path.join(os.path.dirname(__file__), "golden", "crmath.npz"))


@pytest.mark.parametrize("luxury,expect", [(202, 1998227290), (397, 3949287736)])
def test_ranlxd_gsl_known_answers(luxury, expect):
    r = O.Ranlux(1, luxury)
    for _ in range(9999):
        r.get()
    assert r.get() == expect


def test_ranlxd_uniform_is_48bit_grid():
    r = O.Ranlux(12345)
    u = np.array([r.uniform() for _ in range(5000)])
    assert np.all((u >= 0) & (u < 1))
    assert np.all(u * 2.0 ** 48 == np.floor(u * 2.0 ** 48))  # multiples of 2^-48
    assert abs(u.mean() - 0.5) < 0.02


def test_correctly_rounded_log_sin_cos():
    u = CR["u"]
    lg = np.array([O.cr_log(float(x)) for x in u])
    assert np.array_equal(lg, CR["log"])
    sc = np.array([O.cr_sincos(float(x)) for x in CR["phase"]])
    assert np.array_equal(sc[:, 0], CR["sin"])
    assert np.array_equal(sc[:, 1], CR["cos"])


def _seedtable_restated(seed, n):
    """NGenIC's shell order (random_ngenic.cc:49-74), restated."""
    r = O.Ranlux(seed)
    t = np.zeros((n, n), dtype=np.int64)

    def put(i, j):
        t[i, j] = int(0x7fffffff * r.uniform())

    for i in range(n // 2):
        for j in range(i):
            put(i, j)
        for j in range(i + 1):
            put(j, i)
        for j in range(i):
            put(n - 1 - i, j)
        for j in range(i + 1):
            put(n - 1 - j, i)
        for j in range(i):
            put(i, n - 1 - j)
        for j in range(i + 1):
            put(j, n - 1 - i)
        for j in range(i):
            put(n - 1 - i, n - 1 - j)
        for j in range(i + 1):
            put(n - 1 - j, n - 1 - i)
    return t.reshape(-1)


@pytest.mark.parametrize("n", [4, 8, 16])
def test_seedtable_shell_order(n):
    assert np.array_equal(O.seedtable(12345, n).astype(np.int64), _seedtable_restated(12345, n))


def test_seedtable_covers_every_column_once():
    n = 16
    t = O.seedtable(777, n)
    assert np.all(t > 0) and len(np.unique(t)) == n * n


def _column_restated(seed_word, n):
    """one (i, j) column of Fill_Grid (random_ngenic.cc:106-121): phase first,
    then the amplitude rejection loop, for k = 0..n/2"""
    r = O.Ranlux(int(seed_word))
    out = []
    for _ in range(n // 2 + 1):
        phase = r.uniform() * 2 * math.pi
        ampl = 0.0
        while True:
            ampl = r.uniform()
            if ampl != 0.0 and ampl != 1.0:
                break
        a = math.sqrt(-math.log(ampl))
        # glibc (math.*, what the reference's std:: calls reach) is correctly
        # rounded for this draw when its log, cos and sin equal the oracle's
        s, c = O.cr_sincos(phase)
        agree = math.log(ampl) == O.cr_log(ampl) and math.cos(phase) == c and math.sin(phase) == s
        out.append((complex(a * math.cos(phase), a * math.sin(phase)), agree))
    return out


def test_ngenic_columns_and_structure():
    n = 8
    seed = 12345
    t = O.seedtable(seed, n)
    k = O.kview(O.ngenic_fill(seed, n))
    h = n // 2
    for (i, j) in [(1, 2), (3, 5), (6, 1), (2, 7), (5, 6)]:
        col = _column_restated(t[i * n + j], n)
        for kk in range(1, h):  # k > 0, off the Nyquist planes
            if i == h or j == h:
                continue
            z, agree = col[kk]
            if agree:
                assert k[i, j, kk] == z  # bit for bit where glibc rounds correctly
            else:
                assert abs(k[i, j, kk] - z) <= 4e-16 * abs(z)
    # Nyquist planes and DC are zero
    assert np.all(k[h] == 0) and np.all(k[:, h] == 0) and np.all(k[:, :, h] == 0)
    assert k[0, 0, 0] == 0
    # the kz = 0 plane is Hermitian
    p = k[:, :, 0]
    idx = (-np.arange(n)) % n
    assert np.array_equal(p, np.conj(p[idx][:, idx]))


def test_ngenic_statistics():
    n = 32
    k = O.kview(O.ngenic_fill(4242, n))
    h = n // 2
    m = np.ones(k.shape, bool)
    m[h], m[:, h], m[:, :, h] = False, False, False
    m[0, 0, 0] = False
    z = k[m]
    p2 = np.abs(z) ** 2
    assert abs(p2.mean() - 1.0) < 0.03  # <|z|^2> = 1 (exponential)
    ph = np.angle(z)
    hist, _ = np.histogram(ph, bins=8, range=(-np.pi, np.pi))
    assert hist.min() > 0.9 * hist.mean()


def test_gaussian_draws_differ_from_glibc_only_where_glibc_misrounds():
    """C1's 64^3 seed-12345 draws: the stored values differ from the host
    libm evaluation (the reference's std::log/cos/sin on glibc) only in draws
    where one of glibc's log, cos, sin is not correctly rounded, by one
    rounding; the counts are those measured for glibc 2.35 (64^3: 386 of
    125,023 draws; 512^3: 189,028 of 66,716,415, both 0.3%)"""
    c = O.ngenic_libm_compare(12345, 64)
    assert c["draws"] == 125023
    assert c["value"] <= c["log"] + c["cos"] + c["sin"]
    assert c["value"] < 0.005 * c["draws"]
