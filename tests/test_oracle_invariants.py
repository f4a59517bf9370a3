"""Mathematical pins of the oracle's 3LPT transverse term and PLT tables.

Cauchy invariant (the reference's own diagnostic, testing.cc:130-270): with
x = q + Psi(q) and v(q) the Lagrangian velocity, a potential flow keeps
  I_ij = dv_i/dq_j - dv_j/dq_i + sum_k (dPsi_k/dq_i dv_k/dq_j - dPsi_k/dq_j dv_k/dq_i)
at zero.  In 3LPT the first and second order cancel by symmetry and the
third order cancels only through the transverse displacement curl A3
(ic_generator.cc:616-627, 957, 1013), so I is of fourth order in the growth
factor D.  A swapped d'/d'' in the A3 source or in the assembly curl (both
turn A3 into -A3), or A3 taken for the wrong component, leaves a third-order
residual.  The tests measure the order of I in D from two starting redshifts
and plant those mistakes in the oracle's inputs.  The velocity factor vfac
(f a H / h) is divided out, so the order in D is 4 with the correct A3 and 3
with a mistake.

PLT tables (particle_plt.hh:461-499): the corrected gradient direction
v(k) = e_r + D_theta e_theta + D_phi e_phi is the longitudinal eigenvector of
the lattice dynamical matrix scaled to unit radial part; for sc and bcc the
matrix is invariant under the 48 cubic point-group operations R, so
v(Rk) = R v(k) at every table site away from Nyquist indices, whatever the
first-Brillouin-zone folding and neighbour averaging did.  Two exceptions are
left out: the six bcc sites next to k = 0, which average in the k = 0
eigenvector (an exactly isotropic matrix there, so its eigenvector is the
solver's choice), and the Ewald truncation itself: the reference centres the
7^3 image block on each site wrapped into [0,1)^3 (vec3.hh:111-114,
particle_plt.hh:234-248), not on the origin, so its truncation error (the
1e-3 of the Kohn sum rule) is not cubic symmetric.  A wrong fold or a
misplaced table entry moves a direction by O(0.1-1), far above that.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O


def _grad(n, L, axis):
    """Grid_FFT::gradient factor along axis (0 at the Nyquist index), broadcast
    to the (n, n, n//2+1) k-space shape"""
    m = np.arange(n if axis < 2 else n // 2 + 1, dtype=float)
    g = np.where(m > n // 2, m - n, m) * (2 * math.pi / L)
    g[m == n // 2] = 0.0
    shape = [1, 1, 1]
    shape[axis] = -1
    return g.reshape(shape)


def components_k(n, L, z, seed, a3_map=None):
    """k-space displacement (Mpc/h: lunit = L) and velocity components from
    the oracle's potentials; a3_map(A3) replaces the A3 triple (planted
    mistakes)"""
    cfg = O.ic_cfg(n, L, 3, seed, lunit=L, vunit=1.0)
    oc = O.baseline_cosmology(z)
    phi, phi2, phi3, A3 = O.potentials(cfg, oc, O.ngenic_fill(seed, n))
    if a3_map is not None:
        A3 = a3_map(A3)
    coef = O.lpt_coefs(cfg, oc)
    return [O.component_k(cfg, coef, (phi, phi2, phi3, A3), c) for c in range(6)], coef


def cauchy_invariant(n, L, comps):
    """standard deviations of I_xy, I_yz, I_zx (testing.cc:221-268): the
    linear curl pointwise in k-space, the products as Orszag convolutions of
    gradient operands"""
    psi, vel = comps[:3], comps[3:]
    out = []
    for i in range(3):
        j = (i + 1) % 3
        inv = O.empty_field(n)
        O.kview(inv)[:] = 1j * _grad(n, L, j) * O.kview(vel[i]) - 1j * _grad(n, L, i) * O.kview(vel[j])
        for k in range(3):
            O.convolve2(n, L, O.operand(psi[k], O.GRAD, (i,)), O.operand(vel[k], O.GRAD, (j,)), inv, O.ADD)
        for k in range(3):
            O.convolve2(n, L, O.operand(psi[k], O.GRAD, (j,)), O.operand(vel[k], O.GRAD, (i,)), inv, O.SUB)
        O.fft_c2r(inv)
        out.append(float(np.std(O.rview(inv))))
    return np.array(out)


def _negate(A3):
    return tuple(-a for a in A3)


def _cycle(A3):
    return (A3[1], A3[2], A3[0])


N, BOX, SEED = 32, 64.0, 2024
Z_HI, Z_LO = 99.0, 24.0


@pytest.fixture(scope="module")
def invariant_orders():
    res = {}
    for name, fn in (("correct", None), ("a3_negated", _negate), ("a3_cycled", _cycle), ("no_a3", "zero")):
        vals = []
        for z in (Z_HI, Z_LO):
            mp = (lambda A3: tuple(np.zeros_like(a) for a in A3)) if fn == "zero" else fn
            comps, coef = components_k(N, BOX, z, SEED, mp)
            # v carries vfac = f a H (about D^-1/2 in matter domination); I is
            # linear in v, so I / vfac is the quantity of order D^4
            vals.append((coef.Dplus0, cauchy_invariant(N, BOX, comps) / coef.vfac))
        (D1, I1), (D2, I2) = vals
        res[name] = (np.log(I2 / I1) / math.log(D2 / D1), I1, I2)
    return res


def test_cauchy_invariant_is_fourth_order(invariant_orders):
    p, I1, I2 = invariant_orders["correct"]
    print("order of I in D (correct A3):", p, I1, I2)
    assert np.all(p > 3.8), p


@pytest.mark.parametrize("mistake", ["a3_negated", "a3_cycled", "no_a3"])
def test_cauchy_invariant_catches_planted_a3_mistakes(invariant_orders, mistake):
    p, I1, _ = invariant_orders[mistake]
    _, I1c, _ = invariant_orders["correct"]
    print(mistake, "order", p, "std ratio vs correct at z=99", I1 / I1c)
    assert np.all(p < 3.3), p
    assert np.all(I1 / I1c > 10.0)


# ----------------------------------------------------------- PLT symmetry --
def _cubic_group():
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3), dtype=int)
            for r, (c, s) in enumerate(zip(perm, signs)):
                R[r, c] = s
            yield R


def plt_direction_field(tab, ng):
    """v(k) per table site: e_r + D_theta e_theta + D_phi e_phi with the
    spherical basis of particle_plt.hh:477-479, k at the table's integer
    wave numbers"""
    idx = np.arange(ng)
    s = np.where(idx > ng // 2, idx - ng, idx).astype(float)
    kz = np.arange(ng // 2 + 1, dtype=float)
    KX, KY, KZ = np.meshgrid(s, s, kz, indexing="ij")
    kr = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        ph = np.where(kr > 0, np.arctan2(KY, KX), 0.0)
        th = np.where(kr > 0, np.arccos(np.where(kr > 0, KZ / kr, 1.0)), 0.0)
    st, ct, sp, cp = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
    er = np.stack([st * cp, st * sp, ct], -1)
    et = np.stack([ct * cp, ct * sp, -st], -1)
    ep = np.stack([-sp, cp, np.zeros_like(sp)], -1)
    v = er + tab["Dtheta"][..., None] * et + tab["Dphi"][..., None] * ep
    return v, np.stack([KX, KY, KZ], -1)


def check_cubic_symmetry(tab, ng, lattice):
    """max |v(Rk) - R v(k)| over sites with every |index| < ng/2, k != 0;
    v(-k) = -v(k) maps images with negative kz into the stored half"""
    v, K = plt_direction_field(tab, ng)
    h = ng // 2
    ints = np.rint(K).astype(int)
    ok = np.all(np.abs(ints) < h, axis=-1) & np.any(ints != 0, axis=-1)
    if lattice >= 1:
        ok &= np.abs(ints).sum(axis=-1) > 1
    sites = np.argwhere(ok)
    worst = 0.0
    for R in _cubic_group():
        k = ints[ok] @ R.T
        sgn = np.where(k[:, 2] < 0, -1, 1)
        k = k * sgn[:, None]
        ii, jj, kk = k[:, 0] % ng, k[:, 1] % ng, k[:, 2]
        vi = v[ii, jj, kk] * sgn[:, None]
        vk = v[sites[:, 0], sites[:, 1], sites[:, 2]] @ R.T
        worst = max(worst, float(np.max(np.abs(vi - vk))))
    vf = tab["vfac"]
    for R in _cubic_group():
        k = ints[ok] @ R.T
        k = k * np.where(k[:, 2] < 0, -1, 1)[:, None]
        worst = max(worst, float(np.max(np.abs(vf[k[:, 0] % ng, k[:, 1] % ng, k[:, 2]] -
                                                vf[sites[:, 0], sites[:, 1], sites[:, 2]]))))
    return worst


@pytest.mark.parametrize("lattice", [0, 1, 2])
def test_plt_tables_cubic_symmetry(lattice):
    ng = 16
    p = O.plt_tables(lattice, ng)
    tab = O.plt_arrays(p)
    O.lib().orc_plt_free(p)
    err = check_cubic_symmetry(tab, ng, lattice)
    print("lattice", lattice, "max cubic-symmetry violation", err)
    assert err < 3e-3


def test_plt_symmetry_check_catches_a_broken_fold():
    """a table whose D_theta is perturbed on one site and not on its images
    (as a fold that sends one octant to the wrong zone would) fails"""
    ng = 16
    p = O.plt_tables(1, ng)
    tab = O.plt_arrays(p)
    O.lib().orc_plt_free(p)
    tab["Dtheta"][3, 2, 1] += 0.05
    assert check_cubic_symmetry(tab, ng, 1) > 0.03
