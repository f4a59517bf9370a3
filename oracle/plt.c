/*
 * Oracle: PLT lattice corrections (TEST INFRASTRUCTURE ONLY).
 *
 * Restates particle::lattice_gradient (particle_plt.hh:50-587):
 *   init_D  (:68-515)  Ewald-split lattice dynamical matrix on an ng^3 site
 *                      mesh, unitary r2c of its 6 components, per-mode
 *                      symmetric eigen-decomposition (GSL gsl_eigen_symmv +
 *                      ascending sort in mat3.hh:169-182; here cyclic Jacobi),
 *                      folding into the first Brillouin zone, neighbour
 *                      averaging of off-lattice modes, projection on the
 *                      spherical basis of k;
 *   gradient(:551-573) / vfac_corr (:575-585) per N-grid mode via k-space CIC
 *                      (Grid_FFT::get_cic_kspace, grid_fft.hh:347-373).
 * The reference inserts FBZ keys into a std::map from an OpenMP loop under a
 * critical section (particle_plt.hh:381-382), so for a key reached from two
 * modes the stored mode depends on thread timing; the oracle uses the serial
 * loop order (first insertion wins), i.e. the one-thread result.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static const double PI_ = 3.14159265358979323846;

typedef struct { double m[9]; } m3;
#define M(a, i, j) ((a).m[3 * (i) + (j)])

static void mv(const double A[9], const double v[3], double out[3]) {
    for (int mu = 0; mu < 3; ++mu) {
        out[mu] = 0.0;
        for (int nu = 0; nu < 3; ++nu) out[mu] += A[3 * mu + nu] * v[nu];
    }
}

/* cyclic Jacobi for a symmetric 3x3 matrix; eigenvalues ascending,
 * eigenvectors as columns of V */
static void eigen3(const double Ain[9], double ev[3], double V[9]) {
    double A[9];
    memcpy(A, Ain, sizeof(A));
    for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = fabs(A[1]) + fabs(A[2]) + fabs(A[5]);
        if (off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double apq = A[3 * p + q];
                if (apq == 0.0) continue;
                double app = A[3 * p + p], aqq = A[3 * q + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    double akp = A[3 * k + p], akq = A[3 * k + q];
                    A[3 * k + p] = c * akp - s * akq;
                    A[3 * k + q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    double apk = A[3 * p + k], aqk = A[3 * q + k];
                    A[3 * p + k] = c * apk - s * aqk;
                    A[3 * q + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = V[3 * k + p], vkq = V[3 * k + q];
                    V[3 * k + p] = c * vkp - s * vkq;
                    V[3 * k + q] = s * vkp + c * vkq;
                }
            }
    }
    double e[3] = {A[0], A[4], A[8]};
    int idx[3] = {0, 1, 2};
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (e[idx[b]] < e[idx[a]]) { int t = idx[a]; idx[a] = idx[b]; idx[b] = t; }
    double W[9];
    for (int c = 0; c < 3; ++c) {
        ev[c] = e[idx[c]];
        for (int r = 0; r < 3; ++r) W[3 * r + c] = V[3 * r + idx[c]];
    }
    memcpy(V, W, sizeof(W));
}

static const double BRAV[3][9] = {
    {1, 0, 0, 0, 1, 0, 0, 0, 1},
    {1, 0, 0.5, 0, 1, 0.5, 0, 0, 0.5},
    {0, 0.5, 0, 0.5, 0, 1, 0.5, 0.5, 0}};
static const int INVREC[3][9] = {
    {2, 0, 0, 0, 2, 0, 0, 0, 2},
    {2, 0, 0, 0, 2, 0, 1, 1, 1},
    {0, 1, 1, 1, 0, 1, 0, 2, 0}};
static const int CHARGE_FAC[3] = {1, 2, 4};

static void reciprocal(int lat, double R[9]) {
    const double tp = 2.0 * PI_, fp = 4.0 * PI_;
    const double sc[9] = {tp, 0, 0, 0, tp, 0, 0, 0, tp};
    const double bcc[9] = {tp, 0, 0, 0, tp, 0, -tp, -tp, fp};
    const double fcc[9] = {-fp, fp, tp, 0, 0, tp, fp, 0, -tp};
    memcpy(R, lat == 2 ? fcc : lat == 1 ? bcc : sc, sizeof(double) * 9);
}

static int normals(int lat, double nv[][3]) {
    const double p = PI_, tp = 2.0 * PI_;
    int k = 0;
    if (lat == 0) {
        double t[6][3] = {{p, 0, 0}, {-p, 0, 0}, {0, p, 0}, {0, -p, 0}, {0, 0, p}, {0, 0, -p}};
        memcpy(nv, t, sizeof(t)); k = 6;
    } else if (lat == 1) {
        double t[12][3] = {{0, p, p}, {0, -p, p}, {0, p, -p}, {0, -p, -p},
                           {p, 0, p}, {-p, 0, p}, {p, 0, -p}, {-p, 0, -p},
                           {p, p, 0}, {-p, p, 0}, {p, -p, 0}, {-p, -p, 0}};
        memcpy(nv, t, sizeof(t)); k = 12;
    } else {
        double t[14][3] = {{tp, 0, 0}, {-tp, 0, 0}, {0, tp, 0}, {0, -tp, 0}, {0, 0, tp}, {0, 0, -tp},
                           {p, p, p}, {p, p, -p}, {p, -p, p}, {p, -p, -p},
                           {-p, p, p}, {-p, p, -p}, {-p, -p, p}, {-p, -p, -p}};
        memcpy(nv, t, sizeof(t)); k = 14;
    }
    return k;
}

/* signed index along one axis of an ng mesh as get_k: idx > ng/2 -> idx-ng */
static inline double sidx(int ng, int i) { return (double)i - (double)(i > ng / 2) * ng; }

#define KEYR 140
static inline int key_of(int x, int y, int z) {
    return ((x + KEYR) * (2 * KEYR + 1) + (y + KEYR)) * (2 * KEYR + 1) + (z + KEYR);
}

/* the six real-space dynamical-matrix components on the site mesh, unitary
 * r2c transformed in place (particle_plt.hh:160-315) */
static void plt_dynmat(int lat, int ng, double *D[6]) {
    const double dx = 1.0 / ng;
    const double eta = 4.0, alpha = 1.0 / sqrt(2) / eta, alpha2 = alpha * alpha, alpha3 = alpha2 * alpha;
    const double charge = 1.0 / pow((double)ng, 3) / CHARGE_FAC[lat];
    const double fft_norm = 1.0 / pow((double)ng, 3.0);
    const double sqrtpi = sqrt(PI_), pi32 = pow(PI_, 1.5);
    double Rm[9];
    reciprocal(lat, Rm);
    const double *B = BRAV[lat];

#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < ng; ++i)
        for (int j = 0; j < ng; ++j)
            for (int k = 0; k < ng; ++k) {
                double x[3] = {dx * i, dx * j, dx * k}, ar[3];
                mv(B, x, ar);
                for (int a = 0; a < 3; ++a) ar[a] = fmod(2 * 1.0 + ar[a], 1.0);
                double Dm[9] = {0};
                for (int ix = -3; ix <= 3; ix++)
                    for (int iy = -3; iy <= 3; iy++)
                        for (int iz = -3; iz <= 3; iz++) {
                            double nn[3] = {ix, iy, iz}, bn[3], d[3];
                            mv(B, nn, bn);
                            for (int a = 0; a < 3; ++a) d[a] = ar[a] - bn[a];
                            double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                            if (r < 1e-14) continue;
                            double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
                            double K1 = -alpha3 / pi32 * exp(-alpha2 * r2) / r2;
                            double K2 = (erfc(alpha * r) + 2.0 * alpha / sqrtpi * exp(-alpha2 * r2) * r) / (4.0 * PI_);
                            for (int mu = 0; mu < 3; ++mu)
                                for (int nu = mu; nu < 3; ++nu) {
                                    double dd = d[mu] * d[nu] * K1 + ((mu == nu ? 1.0 : 0.0) / r3 - 3.0 * (d[mu] * d[nu]) / r5) * K2;
                                    Dm[3 * mu + nu] += dd;
                                    if (mu != nu) Dm[3 * nu + mu] += dd;
                                }
                        }
                for (int ix = -3; ix <= 3; ix++)
                    for (int iy = -3; iy <= 3; iy++)
                        for (int iz = -3; iz <= 3; iz++) {
                            if (abs(ix) + abs(iy) + abs(iz) == 0) continue;
                            double kn[3] = {(double)ix / ng, (double)iy / ng, (double)iz / ng}, ak[3];
                            mv(Rm, kn, ak);
                            double kmod2 = ak[0] * ak[0] + ak[1] * ak[1] + ak[2] * ak[2];
                            double term = exp(-kmod2 / (4 * alpha2)) * cos(ak[0] * ar[0] + ak[1] * ar[1] + ak[2] * ar[2]) / kmod2 * fft_norm;
                            for (int mu = 0; mu < 3; ++mu)
                                for (int nu = mu; nu < 3; ++nu) {
                                    double dd = ak[mu] * ak[nu] * term;
                                    Dm[3 * mu + nu] += dd;
                                    if (mu != nu) Dm[3 * nu + mu] += dd;
                                }
                        }
                size_t id = ((size_t)i * ng + j) * (ng + 2) + k;
                D[0][id] = Dm[0] * charge; D[1][id] = Dm[1] * charge; D[2][id] = Dm[2] * charge;
                D[3][id] = Dm[4] * charge; D[4][id] = Dm[5] * charge; D[5][id] = Dm[8] * charge;
            }
    D[0][0] = 1.0 / 3.0; D[1][0] = 0.0; D[2][0] = 0.0;
    D[3][0] = 1.0 / 3.0; D[4][0] = 0.0; D[5][0] = 1.0 / 3.0;
    for (int c = 0; c < 6; ++c) orc_fft3d_r2c(D[c], ng, ng, ng);
}

void orc_plt_dynmat(int lat, int ng, double *out6) {
    const int nc = ng / 2 + 1;
    const size_t nr = (size_t)ng * ng * (ng + 2), nk = (size_t)ng * ng * nc;
    const double fft_norm12 = 1.0 / pow((double)ng, 1.5);
    double *D[6];
    for (int c = 0; c < 6; ++c) D[c] = (double *)calloc(nr, sizeof(double));
    plt_dynmat(lat, ng, D);
    for (int c = 0; c < 6; ++c) {
        for (size_t q = 0; q < nk; ++q) out6[c * nk + q] = D[c][2 * q] / fft_norm12;
        free(D[c]);
    }
}

int orc_plt_init(orc_plt *P, int lat, int ng) {
    const int nc = ng / 2 + 1;
    const size_t nr = (size_t)ng * ng * (ng + 2), nk = (size_t)ng * ng * nc;
    const double fft_norm12 = 1.0 / pow((double)ng, 1.5);
    double Rm[9];
    reciprocal(lat, Rm);
    double *D[6];
    for (int c = 0; c < 6; ++c) D[c] = (double *)calloc(nr, sizeof(double));
    plt_dynmat(lat, ng, D);

    /* per mode: eigen data, default = packed D values (never read unless set) */
    double *t6 = (double *)malloc(sizeof(double) * 6 * nk);   /* l0 l1 l2 v0 v1 v2 */
    int *set = (int *)calloc(nk, sizeof(int));
    const int KS = 2 * KEYR + 1;
    int *map = (int *)malloc(sizeof(int) * (size_t)KS * KS * KS);
    for (size_t q = 0; q < (size_t)KS * KS * KS; ++q) map[q] = -1;
    double nv[14][3];
    int nn = normals(lat, nv);
    for (int i = 0; i < ng; i++)
        for (int j = 0; j < ng; j++)
            for (int k = 0; k < nc; k++) {
                size_t q = ((size_t)i * ng + j) * nc + k;
                double kv[3] = {sidx(ng, i) * 2.0 * PI_, sidx(ng, j) * 2.0 * PI_, sidx(ng, k) * 2.0 * PI_};
                double A[9];
                A[0] = D[0][2 * q] / fft_norm12;
                A[1] = A[3] = D[1][2 * q] / fft_norm12;
                A[2] = A[6] = D[2][2 * q] / fft_norm12;
                A[4] = D[3][2 * q] / fft_norm12;
                A[5] = A[7] = D[4][2 * q] / fft_norm12;
                A[8] = D[5][2 * q] / fft_norm12;
                double ev[3], V[9];
                if (q == 0) {
                    /* k = 0: D is isotropic in exact arithmetic, so every
                     * vector is an eigenvector and a solver's choice follows
                     * rounding noise; take the Cartesian basis, which a
                     * symmetric eigensolver returns for an exactly isotropic
                     * matrix (largest-eigenvalue vector = e_z) */
                    for (int c = 0; c < 9; ++c) V[c] = (c % 4 == 0) ? 1.0 : 0.0;
                    ev[0] = ev[1] = ev[2] = (A[0] + A[4] + A[8]) / 3.0;
                } else {
                    eigen3(A, ev, V);
                }
                double e3[3] = {V[2] / (2.0 * PI_ * ng), V[5] / (2.0 * PI_ * ng), V[8] / (2.0 * PI_ * ng)};
                double arv[3] = {kv[0] / (2.0 * PI_ * ng), kv[1] / (2.0 * PI_ * ng), kv[2] / (2.0 * PI_ * ng)}, a[3];
                mv(Rm, arv, a);
                for (int l1 = -1; l1 <= 1; ++l1)
                    for (int l2 = -1; l2 <= 1; ++l2)
                        for (int l3 = -1; l3 <= 1; ++l3)
                            for (int is = 0; is <= 1; ++is) {
                                double sg = 2.0 * is - 1.0;
                                double ls[3] = {l1, l2, l3}, rl[3], vk[3];
                                mv(Rm, ls, rl);
                                for (int c = 0; c < 3; ++c) vk[c] = sg * a[c] + rl[c];
                                int inside = 1;
                                for (int m = 0; m < nn; ++m) {
                                    double nd = nv[m][0] * vk[0] + nv[m][1] * vk[1] + nv[m][2] * vk[2];
                                    double n2 = nv[m][0] * nv[m][0] + nv[m][1] * nv[m][1] + nv[m][2] * nv[m][2];
                                    if (nd > 1.0001 * n2) { inside = 0; break; }
                                }
                                if (!inside) continue;
                                int ix = (int)round(vk[0] * ng / (2.0 * PI_));
                                int iy = (int)round(vk[1] * ng / (2.0 * PI_));
                                int iz = (int)round(vk[2] * ng / (2.0 * PI_));
                                if (abs(ix) <= KEYR && abs(iy) <= KEYR && abs(iz) <= KEYR) {
                                    int key = key_of(ix, iy, iz);
                                    if (map[key] < 0) map[key] = (int)q;
                                }
                                t6[6 * q + 0] = ev[2]; t6[6 * q + 1] = ev[1]; t6[6 * q + 2] = ev[0];
                                t6[6 * q + 3] = e3[0]; t6[6 * q + 4] = e3[1]; t6[6 * q + 5] = e3[2];
                                set[q] = 1;
                            }
            }
    for (int c = 0; c < 6; ++c) free(D[c]);
    P->ng = ng;
    P->Dr = (double *)malloc(sizeof(double) * nk);
    P->Dtheta = (double *)malloc(sizeof(double) * nk);
    P->Dphi = (double *)malloc(sizeof(double) * nk);
    P->vfac = (double *)malloc(sizeof(double) * nk);
    const int *IR = INVREC[lat];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ng; i++)
        for (int j = 0; j < ng; j++)
            for (int k = 0; k < nc; k++) {
                size_t q = ((size_t)i * ng + j) * nc + k;
                int ii = (i > ng / 2) ? i - ng : i, jj = (j > ng / 2) ? j - ng : j, kk = (k > ng / 2) ? k - ng : k;
                double kvi[3] = {ii, jj, kk};
                double l[3] = {0, 0, 0}, v[3] = {0, 0, 0};
                int vv0 = IR[0] * i + IR[1] * j + IR[2] * k, vv1 = IR[3] * i + IR[4] * j + IR[5] * k,
                    vv2 = IR[6] * i + IR[7] * j + IR[8] * k;
                int on_lattice = (vv0 % 2 == 0) && (vv1 % 2 == 0) && (vv2 % 2 == 0);
                int nb[6][3] = {{ii - 1, jj, kk}, {ii + 1, jj, kk}, {ii, jj - 1, kk}, {ii, jj + 1, kk}, {ii, jj, kk - 1}, {ii, jj, kk + 1}};
                int cnt = 0, nlook = on_lattice ? 1 : 6;
                for (int m = 0; m < nlook; ++m) {
                    int kx = on_lattice ? ii : nb[m][0], ky = on_lattice ? jj : nb[m][1], kz = on_lattice ? kk : nb[m][2];
                    if (abs(kx) > KEYR || abs(ky) > KEYR || abs(kz) > KEYR) continue;
                    int r = map[key_of(kx, ky, kz)];
                    if (r < 0) continue;
                    ++cnt;
                    double vt[3] = {t6[6 * r + 3], t6[6 * r + 4], t6[6 * r + 5]};
                    double dt = vt[0] * kvi[0] + vt[1] * kvi[1] + vt[2] * kvi[2];
                    double sg = (dt < 0.0) ? -1.0 : 1.0;
                    for (int c = 0; c < 3; ++c) { l[c] += t6[6 * r + c]; v[c] += vt[c] * sg; }
                }
                if (!on_lattice) {
                    for (int c = 0; c < 3; ++c) { l[c] /= (double)cnt; v[c] /= (double)cnt; }
                }
                /* projection on the spherical basis (PRODUCTION branch) */
                double kv[3] = {sidx(ng, i) * 2.0 * PI_, sidx(ng, j) * 2.0 * PI_, sidx(ng, k) * 2.0 * PI_};
                double mu1 = l[0];
                double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                double e1[3] = {v[0] / vn, v[1] / vn, v[2] / vn};
                double kr = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
                double kphi = kr > 0.0 ? atan2(kv[1], kv[0]) : 0.0;
                double kth = kr > 0.0 ? acos(kv[2] / kr) : 0.0;
                double st = sin(kth), ct = cos(kth), sp = sin(kphi), cp = cos(kphi);
                double er[3] = {st * cp, st * sp, ct}, et[3] = {ct * cp, ct * sp, -st}, ep[3] = {-sp, cp, 0.0};
                double renorm = e1[0] * er[0] + e1[1] * er[1] + e1[2] * er[2];
                if (renorm < 0.01) renorm = 1.0;
                P->Dr[q] = 1.0;
                P->Dtheta[q] = (e1[0] * et[0] + e1[1] * et[1] + e1[2] * et[2]) / renorm;
                P->Dphi[q] = (e1[0] * ep[0] + e1[1] * ep[1] + e1[2] * ep[2]) / renorm;
                P->vfac[q] = 1.0 / (0.25 * (sqrt(1. + 24 * mu1) - 1.));
            }
    P->vfac[0] = 1.0; P->Dr[0] = 1.0; P->Dtheta[0] = 0.0; P->Dphi[0] = 0.0;
    (void)set;
    free(t6); free(set); free(map);
    return 0;
}

void orc_plt_free(orc_plt *P) {
    free(P->Dr); free(P->Dtheta); free(P->Dphi); free(P->vfac);
    P->Dr = P->Dtheta = P->Dphi = P->vfac = NULL;
}

/* Grid_FFT::get_cic_kspace on an ng x ng x (ng/2+1) real table */
static double cic_k(const double *T, int ng, double x, double y, double z) {
    const int nc = ng / 2 + 1;
    int ix = (int)floor(x), iy = (int)floor(y), iz = (int)floor(z);
    double dx = x - ix, tx = 1.0 - dx, dy = y - iy, ty = 1.0 - dy, dz = z - iz, tz = 1.0 - dz;
    int ix1 = (ix + 1) % ng, iy1 = (iy + 1) % ng;
    int iz1 = (iz + 1) < (nc - 1) ? (iz + 1) : (nc - 1);
#define TV(a, b, c) T[((size_t)(a) * ng + (b)) * nc + (c)]
    double v = 0.0;
    v += TV(ix, iy, iz) * tx * ty * tz;
    v += TV(ix, iy, iz1) * tx * ty * dz;
    v += TV(ix, iy1, iz) * tx * dy * tz;
    v += TV(ix, iy1, iz1) * tx * dy * dz;
    v += TV(ix1, iy, iz) * dx * ty * tz;
    v += TV(ix1, iy, iz1) * dx * ty * dz;
    v += TV(ix1, iy1, iz) * dx * dy * tz;
    v += TV(ix1, iy1, iz1) * dx * dy * dz;
#undef TV
    return v;
}

double orc_plt_gradient(const orc_plt *P, int idim, int n, double L, int i, int j, int k) {
    const int ng = P->ng;
    const double mapratio = (double)ng / (double)n;
    double ix = i * mapratio, iy = j * mapratio, iz = k * mapratio;
    const double kf = 2.0 * PI_;   /* box length 1 */
    double kv[3] = {(ix - (double)(ix > (double)(ng / 2)) * ng) * kf,
                    (iy - (double)(iy > (double)(ng / 2)) * ng) * kf,
                    (iz - (double)(iz > (double)(ng / 2)) * ng) * kf};
    double kr = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
    double kmod = kr / mapratio / L;
    double Dr = cic_k(P->Dr, ng, ix, iy, iz);
    double Dt = cic_k(P->Dtheta, ng, ix, iy, iz);
    double Dp = cic_k(P->Dphi, ng, ix, iy, iz);
    double kphi = kr > 0.0 ? atan2(kv[1], kv[0]) : 0.0;
    double kth = kr > 0.0 ? acos(kv[2] / kr) : 0.0;
    double st = sin(kth), ct = cos(kth), sp = sin(kphi), cp = cos(kphi);
    if (idim == 0) return kmod * (Dr * st * cp + Dt * ct * cp - Dp * sp);
    if (idim == 1) return kmod * (Dr * st * sp + Dt * ct * sp + Dp * cp);
    return kmod * (Dr * ct - Dt * st);
}

double orc_plt_vfac_corr(const orc_plt *P, int n, int i, int j, int k) {
    const int ng = P->ng;
    const double mapratio = (double)ng / (double)n;
    double a = 1.0 / cic_k(P->vfac, ng, i * mapratio, j * mapratio, k * mapratio);
    return 1.0 / a;
}
