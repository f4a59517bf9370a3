"""Pins the oracle's Orszag 3/2 convolver (convolution.hh:412-651) and the
2LPT / 3LPT sources (ic_generator.cc:523-629) by brute force and closed
forms, independently of any FFT:

* For band-limited operands (zero on the i, j or k == N/2 planes) the 3/2
  padded product followed by the unpad truncation equals the direct
  (non-periodic) convolution of the spectra, N^-3/2 sum_p F(p) G(k - p),
  restricted to |k_a| < N/2.  That sum is evaluated here directly (O(N^6)).
* A single plane wave has a vanishing second-order source; two orthogonal
  cosines give mu2 = A B kx^2 ky^2 cos(kx x) cos(ky y) exactly.
"""
import numpy as np
import pytest
from scipy.signal import convolve

from oracle import oracle as O

TWO_PI = 2 * np.pi


def signed(n):
    """signed mode numbers of the N-grid without the Nyquist index"""
    return np.arange(-(n // 2) + 1, n // 2)


def half_to_signed(kh, n):
    """half spectrum (n, n, n/2+1) -> full signed spectrum on (-n/2, n/2)^3"""
    s = signed(n)
    full = np.zeros((n, n, n), complex)
    full[:, :, : n // 2 + 1] = kh
    for k in range(n // 2 + 1, n):
        full[:, :, k] = np.conj(kh[(-np.arange(n)) % n][:, (-np.arange(n)) % n, n - k])
    idx = s % n
    return full[np.ix_(idx, idx, idx)]


def signed_to_half(fs, n):
    s = signed(n)
    out = np.zeros((n, n, n // 2 + 1), complex)
    kk = s[s >= 0]
    out[np.ix_(s % n, s % n, kk)] = fs[:, :, s >= 0]
    return out


def conv(fs, gs, n):
    h = n // 2
    full = convolve(fs, gs, mode="full", method="direct")
    sl = slice(h - 1, 3 * h - 2)
    return full[sl, sl, sl] / n ** 1.5


def kvec(n, L):
    s = signed(n) * TWO_PI / L
    return np.meshgrid(s, s, s, indexing="ij")


def hess(fs, a, b, n, L):
    k = kvec(n, L)
    return -(k[a] * k[b]) * fs


def band_limited(n, seed):
    rng = np.random.default_rng(seed)
    a = O.empty_field(n)
    O.rview(a)[...] = rng.standard_normal((n, n, n))
    O.fft_r2c(a)
    kh = O.kview(a)
    h = n // 2
    kh[h], kh[:, h], kh[:, :, h] = 0, 0, 0
    return a


def max_rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("n", [8, 12])
def test_convolve2_brute_force(n):
    L = 37.0
    A, B = band_limited(n, 1), band_limited(n, 2)
    As, Bs = half_to_signed(O.kview(A), n), half_to_signed(O.kview(B), n)
    k = kvec(n, L)
    cases = [
        ((O.HESS, (0, 1, 0, 0)), (O.HESS, (2, 2, 0, 0)), hess(As, 0, 1, n, L), hess(Bs, 2, 2, n, L)),
        ((O.SUM_HESS, (0, 0, 1, 1)), (O.DIFF_HESS, (1, 2, 0, 2)),
         hess(As, 0, 0, n, L) + hess(As, 1, 1, n, L), hess(Bs, 1, 2, n, L) - hess(Bs, 0, 2, n, L)),
        ((O.GRAD, (1, 0, 0, 0)), (O.IDENT, (0, 0, 0, 0)), 1j * k[1] * As, Bs),
    ]
    for (ka, da), (kb, db), fa, fb in cases:
        expect = signed_to_half(conv(fa, fb, n), n)
        out = O.empty_field(n)
        O.convolve2(n, L, O.operand(A, ka, da), O.operand(B, kb, db), out, O.ASSIGN)
        assert max_rel(O.kview(out), expect) < 1e-12
        # output operators (operators.hh:25-35) on a pre-filled output
        base = O.kview(band_limited(n, 9)).copy()
        for op, coef, want in ((O.ADD, 0.0, base + expect), (O.SUB, 0.0, base - expect),
                               (O.MADD, 0.75, base + 0.75 * expect)):
            o2 = O.empty_field(n)
            O.kview(o2)[...] = base
            O.convolve2(n, L, O.operand(A, ka, da), O.operand(B, kb, db), o2, op, coef)
            assert max_rel(O.kview(o2), want) < 1e-12


def test_convolve3_is_two_stage_truncation():
    n, L = 8, 11.0
    A, B, Cf = band_limited(n, 3), band_limited(n, 4), band_limited(n, 5)
    As, Bs, Cs = (half_to_signed(O.kview(x), n) for x in (A, B, Cf))
    fa, fb, fc = hess(As, 0, 0, n, L), hess(Bs, 1, 1, n, L), hess(Cs, 0, 2, n, L)
    expect = signed_to_half(conv(conv(fa, fb, n), fc, n), n)
    out = O.empty_field(n)
    O.convolve3(n, L, O.operand(A, O.HESS, (0, 0)), O.operand(B, O.HESS, (1, 1)), O.operand(Cf, O.HESS, (0, 2)),
                out, O.ASSIGN)
    assert max_rel(O.kview(out), expect) < 1e-12
    # a single-stage triple product differs (trunc(trunc(ab)c) is not trunc(abc))
    h = n // 2
    big = convolve(convolve(fa, fb, mode="full", method="direct"), fc, mode="full", method="direct")
    sl = slice(2 * (h - 1), 2 * (h - 1) + n - 1)
    one_stage = signed_to_half(big[sl, sl, sl] / n ** 3, n)
    assert max_rel(one_stage, expect) > 1e-3


def test_plane_wave_and_two_cosines_closed_forms():
    n, L = 16, 50.0
    x = np.arange(n) * L / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    kx, ky = 2 * TWO_PI / L, 3 * TWO_PI / L
    # two orthogonal cosines: mu2 term phi_xx phi_yy
    Aamp, Bamp = 0.7, -1.3
    phi = O.empty_field(n)
    O.rview(phi)[...] = Aamp * np.cos(kx * X) + Bamp * np.cos(ky * Y)
    O.fft_r2c(phi)
    out = O.empty_field(n)
    O.convolve2(n, L, O.operand(phi, O.HESS, (0, 0)), O.operand(phi, O.HESS, (1, 1)), out, O.ASSIGN)
    O.fft_c2r(out)
    want = Aamp * Bamp * kx ** 2 * ky ** 2 * np.cos(kx * X) * np.cos(ky * Y)
    assert np.max(np.abs(O.rview(out) - want)) < 1e-12 * np.max(np.abs(want))
    # a single oblique plane wave: mu2 = sum_{a<b} (phi_aa phi_bb - phi_ab^2) = 0
    kv = np.array([1, 2, 3]) * TWO_PI / L
    pw = O.empty_field(n)
    O.rview(pw)[...] = np.cos(kv[0] * X + kv[1] * Y + kv[2] * Z)
    O.fft_r2c(pw)
    mu2 = O.empty_field(n)
    for (a, b, c, d, s) in ((0, 0, 1, 1, O.ADD), (0, 0, 2, 2, O.ADD), (1, 1, 2, 2, O.ADD),
                            (0, 1, 0, 1, O.SUB), (0, 2, 0, 2, O.SUB), (1, 2, 1, 2, O.SUB)):
        O.convolve2(n, L, O.operand(pw, O.HESS, (a, b)), O.operand(pw, O.HESS, (c, d)), mu2, s)
    scale = np.max(np.abs(kv)) ** 4
    assert np.max(np.abs(O.kview(mu2))) < 1e-13 * scale * n ** 1.5


def inv_lap(fs, n, L):
    k = kvec(n, L)
    k2 = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    out = np.zeros_like(fs)
    m = k2 > 0
    out[m] = -fs[m] / k2[m]
    return out


def test_lpt_potentials_brute_force():
    """phi2, phi3 and A3 of the oracle's pipeline (ic_generator.cc:523-658)
    against the direct-sum evaluation of the reference's formulae"""
    n, L = 8, 64.0
    cfg = O.ic_cfg(n, L, 3, 4321)
    cosmo = O.baseline_cosmology(24.0)
    co = O.lpt_coefs(cfg, cosmo)
    wn = O.ngenic_fill(4321, n)
    phi, phi2, phi3, A3 = O.potentials(cfg, cosmo, wn)
    p1 = half_to_signed(O.kview(phi), n) / co.g1
    H1 = {(a, b): hess(p1, a, b, n, L) for a in range(3) for b in range(3)}
    src2 = (conv(H1[0, 0], H1[1, 1] + H1[2, 2], n) + conv(H1[1, 1], H1[2, 2], n)
            - conv(H1[0, 1], H1[0, 1], n) - conv(H1[0, 2], H1[0, 2], n) - conv(H1[1, 2], H1[1, 2], n))
    p2 = inv_lap(src2, n, L)
    assert max_rel(O.kview(phi2), signed_to_half(co.g2 * p2, n)) < 1e-11
    H2 = {(a, b): hess(p2, a, b, n, L) for a in range(3) for b in range(3)}

    def c3(a, b, c):
        return conv(conv(H1[a], H1[b], n), H1[c], n)

    s3a = (c3((0, 0), (1, 1), (2, 2)) + 2 * c3((0, 1), (0, 2), (1, 2)) - c3((1, 2), (1, 2), (0, 0))
           - c3((0, 2), (0, 2), (1, 1)) - c3((0, 1), (0, 1), (2, 2)))
    s3b = sum(-5 / 7 * conv(H1[i, i], H2[j, j] + H2[k, k], n) for i, j, k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)))
    s3b = s3b + sum(10 / 7 * conv(H1[a, b], H2[a, b], n) for a, b in ((0, 1), (0, 2), (1, 2)))
    assert max_rel(O.kview(phi3), signed_to_half(co.g3 * inv_lap(s3a + s3b, n, L), n)) < 1e-11
    for d in range(3):
        d1, d2 = (d + 1) % 3, (d + 2) % 3
        s = (conv(H2[d, d1], H1[d, d2], n) - conv(H2[d, d2], H1[d, d1], n)
             + conv(H1[d1, d2], H2[d1, d1] - H2[d2, d2], n) - conv(H2[d1, d2], H1[d1, d1] - H1[d2, d2], n))
        assert max_rel(O.kview(A3[d]), signed_to_half(co.g3c * inv_lap(s, n, L), n)) < 1e-11
