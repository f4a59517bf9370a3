/*
 * Oracle: Orszag 3/2 convolver, LPT potentials, assembly and lattice readout
 * in the reference's own schedule (TEST INFRASTRUCTURE ONLY).
 *
 *   operands      convolution.hh:74-275 (Hessian via Grid_FFT::gradient,
 *                 Nyquist index -> 0 along the differentiated axis,
 *                 grid_fft.hh:376-385)
 *   convolve2     convolution.hh:495-518, pad_insert :552-575, unpad :600-636
 *   convolve3     convolution.hh:525-532 (two-stage truncation)
 *   output ops    operators.hh:25-35
 *   phi1          ic_generator.cc:246-247, 378-383, 509-516
 *   phi2          ic_generator.cc:523-555
 *   phi3, A3      ic_generator.cc:571-629
 *   growth        ic_generator.cc:258-273, 644-658
 *   assembly      ic_generator.cc:931-970 (positions), 987-1047 (velocities)
 *   readout       particle_generator.hh:499-544, 597-639; shift_field
 *                 grid_fft.hh:832-842
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

static const double PI_ = 3.14159265358979323846;

static double wtime(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

static size_t nreal(int n) { return (size_t)n * n * (n + 2); }

/* Grid_FFT::gradient imaginary part: (idx - [idx>n/2] n) * 2pi/L, 0 at n/2 */
static inline double grad1(int n, double kfac, int idx) {
    if (idx == n / 2) return 0.0;
    return ((double)idx - (double)(idx > n / 2) * n) * kfac;
}
/* Grid_FFT::get_k component */
static inline double kcomp(int n, double kfac, int idx) {
    return ((double)idx - (double)(idx > n / 2) * n) * kfac;
}

/* operand value at N-grid mode (i,j,k): returns re/im */
static void operand_eval(const orc_operand *o, int n, double kfac, int i, int j, int k,
                         double *re, double *im) {
    const int nc = n / 2 + 1;
    const double *e = o->f + 2 * (((size_t)i * n + j) * nc + k);
    int ijk[3] = {i, j, k};
    double fr = e[0], fi = e[1];
    double s;
    switch (o->kind) {
    case ORC_KIND_HESS: {
        double g0 = grad1(n, kfac, ijk[o->d[0]]), g1 = grad1(n, kfac, ijk[o->d[1]]);
        s = -(g0 * g1);      /* (i g0)(i g1) */
        *re = s * fr; *im = s * fi; return;
    }
    case ORC_KIND_SUM_HESS: {
        double g0 = grad1(n, kfac, ijk[o->d[0]]), g1 = grad1(n, kfac, ijk[o->d[1]]);
        double g2 = grad1(n, kfac, ijk[o->d[2]]), g3 = grad1(n, kfac, ijk[o->d[3]]);
        s = -(g0 * g1) + -(g2 * g3);
        *re = s * fr; *im = s * fi; return;
    }
    case ORC_KIND_DIFF_HESS: {
        double g0 = grad1(n, kfac, ijk[o->d[0]]), g1 = grad1(n, kfac, ijk[o->d[1]]);
        double g2 = grad1(n, kfac, ijk[o->d[2]]), g3 = grad1(n, kfac, ijk[o->d[3]]);
        s = -(g0 * g1) - -(g2 * g3);
        *re = s * fr; *im = s * fi; return;
    }
    case ORC_KIND_GRAD: {
        double g0 = grad1(n, kfac, ijk[o->d[0]]);
        *re = -g0 * fi; *im = g0 * fr; return;
    }
    default:
        *re = fr; *im = fi; return;
    }
}

static void pad_insert(const orc_operand *o, int n, double L, double *fp) {
    const int np = 3 * n / 2, ncp = np / 2 + 1, nh = n / 2;
    const double kfac = 2.0 * PI_ / L;
    const double rfac = pow(1.5, 1.5);
    memset(fp, 0, nreal(np) * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        int ip = (i > nh) ? i + nh : i;
        for (int j = 0; j < n; ++j) {
            int jp = (j > nh) ? j + nh : j;
            for (int k = 0; k < nh + 1; ++k) {
                double re, im;
                operand_eval(o, n, kfac, i, j, k, &re, &im);
                double *e = fp + 2 * (((size_t)ip * np + jp) * ncp + k);
                e[0] = re * rfac; e[1] = im * rfac;
            }
        }
    }
}

static void apply_op(double *out, size_t idx, double v, int op, double coef) {
    switch (op) {
    case ORC_OP_ASSIGN: out[idx] = v; break;
    case ORC_OP_ADD: out[idx] += v; break;
    case ORC_OP_SUB: out[idx] -= v; break;
    default: out[idx] += v * coef; break;
    }
}

static void unpad(const double *fp, int n, double *out, int op, double coef) {
    const int np = 3 * n / 2, ncp = np / 2 + 1, nc = n / 2 + 1, nh = n / 2;
    const double rfac = sqrt((double)((size_t)np * np * np)) / sqrt((double)((size_t)n * n * n));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        int ip = (i > nh) ? i + nh : i;
        for (int j = 0; j < n; ++j) {
            int jp = (j > nh) ? j + nh : j;
            for (int k = 0; k < nc; ++k) {
                const double *e = fp + 2 * (((size_t)ip * np + jp) * ncp + k);
                double re = e[0] / rfac, im = e[1] / rfac;
                if (i == nh || j == nh || k == nh) { re = 0.0; im = 0.0; }
                size_t idx = 2 * (((size_t)i * n + j) * nc + k);
                apply_op(out, idx, re, op, coef);
                apply_op(out, idx + 1, im, op, coef);
            }
        }
    }
}

void orc_convolve2(int n, double L, const orc_operand *a, const orc_operand *b,
                   double *out, int op, double coef) {
    const int np = 3 * n / 2;
    double *f1 = (double *)malloc(nreal(np) * sizeof(double));
    double *f2 = (double *)malloc(nreal(np) * sizeof(double));
    pad_insert(a, n, L, f1);
    pad_insert(b, n, L, f2);
    orc_fft3d_c2r(f1, np, np, np);
    orc_fft3d_c2r(f2, np, np, np);
    size_t tot = nreal(np);
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < tot; ++r) f2[r] *= f1[r];
    orc_fft3d_r2c(f2, np, np, np);
    unpad(f2, n, out, op, coef);
    free(f1); free(f2);
}

void orc_convolve3(int n, double L, const orc_operand *a, const orc_operand *b,
                   const orc_operand *c, double *out, int op, double coef) {
    double *fbuf = (double *)calloc(nreal(n), sizeof(double));
    orc_convolve2(n, L, a, b, fbuf, ORC_OP_ASSIGN, 0.0);
    orc_operand id = {fbuf, ORC_KIND_IDENT, {0, 0, 0, 0}};
    orc_convolve2(n, L, &id, c, out, op, coef);
    free(fbuf);
}

/* ------------------------------------------------------ k-space ops ---- */
static void inverse_laplacian(int n, double L, double *f) {
    const int nc = n / 2 + 1;
    const double kfac = 2.0 * PI_ / L;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < nc; ++k) {
                double kx = kcomp(n, kfac, i), ky = kcomp(n, kfac, j), kz = kcomp(n, kfac, k);
                double k2 = kx * kx + ky * ky + kz * kz;
                double *e = f + 2 * (((size_t)i * n + j) * nc + k);
                e[0] = -e[0] / k2; e[1] = -e[1] / k2;
            }
    f[0] = 0.0; f[1] = 0.0;
}

static void scale(int n, double *f, double s) {
    size_t tot = nreal(n);
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < tot; ++r) f[r] *= s;
}

void orc_lpt_coefs(const orc_ic_cfg *cfg, const orc_cosmo *c, orc_lpt_coef *k) {
    k->volfac = pow(cfg->L / cfg->n / 2.0 / PI_, 1.5);
    k->Dplus0 = c->Dplus_start;
    k->vfac = c->vfac_start;
    k->g1 = -k->Dplus0;
    k->g2 = (cfg->lpt_order > 1) ? -3.0 / 7.0 * k->Dplus0 * k->Dplus0 : 0.0;
    k->g3 = (cfg->lpt_order > 2) ? 1.0 / 3.0 * k->Dplus0 * k->Dplus0 * k->Dplus0 : 0.0;
    k->g3c = (cfg->lpt_order > 2) ? 1.0 / 7.0 * k->Dplus0 * k->Dplus0 * k->Dplus0 : 0.0;
    k->vfac1 = k->vfac;
    k->vfac2 = 2 * k->vfac;
    k->vfac3 = 3 * k->vfac;
}

#define H(f, a, b) {f, ORC_KIND_HESS, {a, b, 0, 0}}
#define SH(f, a, b, c, d) {f, ORC_KIND_SUM_HESS, {a, b, c, d}}
#define DH(f, a, b, c, d) {f, ORC_KIND_DIFF_HESS, {a, b, c, d}}

static void conv2(int n, double L, orc_operand a, orc_operand b, double *out, int op, double cf) {
    orc_convolve2(n, L, &a, &b, out, op, cf);
}
static void conv3(int n, double L, orc_operand a, orc_operand b, orc_operand c, double *out, int op, double cf) {
    orc_convolve3(n, L, &a, &b, &c, out, op, cf);
}

void orc_lpt_potentials(const orc_ic_cfg *cfg, const orc_cosmo *c, const double *wn_in,
                        double *phi, double *phi2, double *phi3, double *A3x,
                        double *A3y, double *A3z) {
    const int n = cfg->n, nc = n / 2 + 1;
    const double L = cfg->L, kfac = 2.0 * PI_ / L;
    orc_lpt_coef co;
    orc_lpt_coefs(cfg, c, &co);
    /* noise normalisation (ic_generator.cc:378-383) and phi1 (509-516) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < nc; ++k) {
                size_t id = 2 * (((size_t)i * n + j) * nc + k);
                double wr = wn_in[id], wi = wn_in[id + 1];
                if (cfg->do_fixing) {
                    double a = hypot(wr, wi);
                    if (a != 0.0) { wr /= a; wi /= a; }
                }
                double sg = cfg->do_inversion ? -1.0 : 1.0;
                wr = sg * wr / co.volfac; wi = sg * wi / co.volfac;
                double kx = kcomp(n, kfac, i), ky = kcomp(n, kfac, j), kz = kcomp(n, kfac, k);
                double kmod = sqrt(kx * kx + ky * ky + kz * kz);
                double A = orc_amplitude(c, kmod);
                double dr = wr * A, di = wi * A;
                phi[id] = -dr / (kmod * kmod);
                phi[id + 1] = -di / (kmod * kmod);
            }
    phi[0] = 0.0; phi[1] = 0.0;

    if (cfg->lpt_order > 1) {
        conv2(n, L, (orc_operand)H(phi, 0, 0), (orc_operand)SH(phi, 1, 1, 2, 2), phi2, ORC_OP_ASSIGN, 0);
        conv2(n, L, (orc_operand)H(phi, 1, 1), (orc_operand)H(phi, 2, 2), phi2, ORC_OP_ADD, 0);
        conv2(n, L, (orc_operand)H(phi, 0, 1), (orc_operand)H(phi, 0, 1), phi2, ORC_OP_SUB, 0);
        conv2(n, L, (orc_operand)H(phi, 0, 2), (orc_operand)H(phi, 0, 2), phi2, ORC_OP_SUB, 0);
        conv2(n, L, (orc_operand)H(phi, 1, 2), (orc_operand)H(phi, 1, 2), phi2, ORC_OP_SUB, 0);
        inverse_laplacian(n, L, phi2);
    }
    if (cfg->lpt_order > 2) {
        conv3(n, L, (orc_operand)H(phi, 0, 0), (orc_operand)H(phi, 1, 1), (orc_operand)H(phi, 2, 2), phi3, ORC_OP_ASSIGN, 0);
        conv3(n, L, (orc_operand)H(phi, 0, 1), (orc_operand)H(phi, 0, 2), (orc_operand)H(phi, 1, 2), phi3, ORC_OP_MADD, 2.0);
        conv3(n, L, (orc_operand)H(phi, 1, 2), (orc_operand)H(phi, 1, 2), (orc_operand)H(phi, 0, 0), phi3, ORC_OP_SUB, 0);
        conv3(n, L, (orc_operand)H(phi, 0, 2), (orc_operand)H(phi, 0, 2), (orc_operand)H(phi, 1, 1), phi3, ORC_OP_SUB, 0);
        conv3(n, L, (orc_operand)H(phi, 0, 1), (orc_operand)H(phi, 0, 1), (orc_operand)H(phi, 2, 2), phi3, ORC_OP_SUB, 0);
        conv2(n, L, (orc_operand)H(phi, 0, 0), (orc_operand)SH(phi2, 1, 1, 2, 2), phi3, ORC_OP_MADD, -5.0 / 7.0);
        conv2(n, L, (orc_operand)H(phi, 1, 1), (orc_operand)SH(phi2, 2, 2, 0, 0), phi3, ORC_OP_MADD, -5.0 / 7.0);
        conv2(n, L, (orc_operand)H(phi, 2, 2), (orc_operand)SH(phi2, 0, 0, 1, 1), phi3, ORC_OP_MADD, -5.0 / 7.0);
        conv2(n, L, (orc_operand)H(phi, 0, 1), (orc_operand)H(phi2, 0, 1), phi3, ORC_OP_MADD, +10.0 / 7.0);
        conv2(n, L, (orc_operand)H(phi, 0, 2), (orc_operand)H(phi2, 0, 2), phi3, ORC_OP_MADD, +10.0 / 7.0);
        conv2(n, L, (orc_operand)H(phi, 1, 2), (orc_operand)H(phi2, 1, 2), phi3, ORC_OP_MADD, +10.0 / 7.0);
        inverse_laplacian(n, L, phi3);
        double *A3[3] = {A3x, A3y, A3z};
        for (int d = 0; d < 3; ++d) {
            int dp = (d + 1) % 3, dpp = (d + 2) % 3;
            conv2(n, L, (orc_operand)H(phi2, d, dp), (orc_operand)H(phi, d, dpp), A3[d], ORC_OP_ASSIGN, 0);
            conv2(n, L, (orc_operand)H(phi2, d, dpp), (orc_operand)H(phi, d, dp), A3[d], ORC_OP_SUB, 0);
            conv2(n, L, (orc_operand)H(phi, dp, dpp), (orc_operand)DH(phi2, dp, dp, dpp, dpp), A3[d], ORC_OP_ADD, 0);
            conv2(n, L, (orc_operand)H(phi2, dp, dpp), (orc_operand)DH(phi, dp, dp, dpp, dpp), A3[d], ORC_OP_SUB, 0);
            inverse_laplacian(n, L, A3[d]);
        }
    }
    scale(n, phi, co.g1);
    if (cfg->lpt_order > 1) scale(n, phi2, co.g2);
    if (cfg->lpt_order > 2) {
        scale(n, phi3, co.g3);
        scale(n, A3x, co.g3c); scale(n, A3y, co.g3c); scale(n, A3z, co.g3c);
    }
}

/* lattice gradient: Fourier (operators.hh:38-61) or PLT */
static inline double lg_grad(const orc_plt *plt, int n, double L, int idim, int i, int j, int k) {
    if (plt) return orc_plt_gradient(plt, idim, n, L, i, j, k);
    int ijk[3] = {i, j, k};
    int nh = n / 2;
    double rg = (ijk[idim] != nh) ? ((double)ijk[idim] - (double)(ijk[idim] > nh) * n) : 0.0;
    return rg * (2.0 * PI_ / L);
}

void orc_lpt_component_k(const orc_ic_cfg *cfg, const orc_lpt_coef *co, const orc_plt *plt,
                         const double *phi, const double *phi2, const double *phi3,
                         const double *A3x, const double *A3y, const double *A3z,
                         int comp, double *out) {
    const int n = cfg->n, nc = n / 2 + 1, ord = cfg->lpt_order;
    const double L = cfg->L;
    const int vel = comp >= 3, d = comp % 3, dp = (d + 1) % 3, dpp = (d + 2) % 3;
    const double *A3[3] = {A3x, A3y, A3z};
    const double unit = vel ? cfg->vunit / L : cfg->lunit / L;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < nc; ++k) {
                size_t id = 2 * (((size_t)i * n + j) * nc + k);
                double pr, pi;
                if (!vel) {
                    pr = phi[id]; pi = phi[id + 1];
                    if (ord > 1) { pr += phi2[id]; pi += phi2[id + 1]; }
                    if (ord > 2) { pr += phi3[id]; pi += phi3[id + 1]; }
                } else {
                    pr = co->vfac1 * phi[id]; pi = co->vfac1 * phi[id + 1];
                    if (ord > 1) { pr += co->vfac2 * phi2[id]; pi += co->vfac2 * phi2[id + 1]; }
                    if (ord > 2) { pr += co->vfac3 * phi3[id]; pi += co->vfac3 * phi3[id + 1]; }
                }
                double g = lg_grad(plt, n, L, d, i, j, k);
                /* (i g) * (pr + i pi) */
                double vr = -g * pi, vi = g * pr;
                if (ord > 2) {
                    double gp = lg_grad(plt, n, L, dp, i, j, k), gpp = lg_grad(plt, n, L, dpp, i, j, k);
                    const double *a1 = A3[dpp] + id, *a2 = A3[dp] + id;
                    /* (i gp) a1 - (i gpp) a2 */
                    double cr = (-gp * a1[1]) - (-gpp * a2[1]);
                    double ci = (gp * a1[0]) - (gpp * a2[0]);
                    if (vel) { cr *= co->vfac3; ci *= co->vfac3; }
                    vr += cr; vi += ci;
                }
                if (vel) {
                    double vc = plt ? orc_plt_vfac_corr(plt, n, i, j, k) : 1.0;
                    vr *= vc; vi *= vc;
                }
                out[id] = vr * unit; out[id + 1] = vi * unit;
            }
    out[0] = 0.0; out[1] = 0.0;
}

/* ------------------------------------------------------- readout ------- */
void orc_shift_field(int n, double L, const double s[3], double *f) {
    const int nc = n / 2 + 1;
    const double kfac = 2.0 * PI_ / L, dx = L / n;
    orc_fft3d_r2c(f, n, n, n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < nc; ++k) {
                double sh = s[0] * kcomp(n, kfac, i) * dx + s[1] * kcomp(n, kfac, j) * dx +
                            s[2] * kcomp(n, kfac, k) * dx;
                double c = cos(sh), sn = sin(sh);
                double *e = f + 2 * (((size_t)i * n + j) * nc + k);
                double re = e[0] * c - e[1] * sn, im = e[0] * sn + e[1] * c;
                e[0] = re; e[1] = im;
            }
    orc_fft3d_c2r(f, n, n, n);
}

/* lattice_shifts (particle_generator.hh:47-53): sc, bcc, fcc */
static const double SHIFTS[3][4][3] = {{{0, 0, 0}},
                                       {{0, 0, 0}, {0.5, 0.5, 0.5}},
                                       {{0, 0, 0}, {0, 0.5, 0.5}, {0.5, 0, 0.5}, {0.5, 0.5, 0}}};
/* second_lattice_shift (particle_generator.hh:58-63): sc, bcc, fcc */
static const double SECOND[3][3] = {{0.5, 0.5, 0.5}, {0.5, 0.5, 0.0}, {0.5, 0.5, 0.5}};

/* set_positions / set_velocities (particle_generator.hh:499-544, 597-639)
 * for lattice index 0 or 1: index 1 shifts the field by the secondary
 * lattice's offset before the first sub-lattice and offsets q by it */
static void readout(int n, double L, int lattice, int lattice_index, int idim, double lunit, int is_pos,
                    double *f, double *outv) {
    const size_t np1 = (size_t)n * n * n;
    const int nshift = 1 << lattice;
    for (int is = 0; is < nshift; ++is) {
        if (is == 0 && lattice_index == 1) orc_shift_field(n, L, SECOND[lattice], f);
        if (is > 0) {
            double s[3] = {SHIFTS[lattice][is][0] - SHIFTS[lattice][is - 1][0],
                           SHIFTS[lattice][is][1] - SHIFTS[lattice][is - 1][1],
                           SHIFTS[lattice][is][2] - SHIFTS[lattice][is - 1][2]};
            orc_shift_field(n, L, s, f);
        }
        double sh[3];
        for (int a = 0; a < 3; ++a) sh[a] = SHIFTS[lattice][is][a] + (lattice_index == 1 ? SECOND[lattice][a] : 0.0);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                for (int k = 0; k < n; ++k) {
                    size_t ip = is * np1 + ((size_t)i * n + j) * n + k;
                    double v = f[((size_t)i * n + j) * (n + 2) + k];
                    if (is_pos) {
                        int ijk[3] = {i, j, k};
                        double q = ((double)ijk[idim] + sh[idim]) / (double)n;
                        v = q * lunit + v;
                    }
                    outv[3 * ip + idim] = v;
                }
    }
}

void orc_lattice_set_positions(int n, double L, int lattice, int idim, double lunit,
                               double *field, double *pos) {
    readout(n, L, lattice, 0, idim, lunit, 1, field, pos);
}
void orc_lattice_set_velocities(int n, double L, int lattice, int idim,
                                double *field, double *vel) {
    readout(n, L, lattice, 0, idim, 0.0, 0, field, vel);
}
void orc_lattice_readout_index(int n, double L, int lattice, int lattice_index, int idim, double lunit,
                               int is_pos, double *field, double *out) {
    readout(n, L, lattice, lattice_index, idim, lunit, is_pos, field, out);
}

/* ------------------------------------------------------ full run ------- */
int orc_ic_run(const orc_ic_cfg *cfg, orc_cosmo *c, double *fields6, double *pos,
               double *vel, double st[8]) {
    const int n = cfg->n;
    const size_t nr = nreal(n);
    double t0 = wtime();
    for (int i = 0; i < 8; ++i) if (st) st[i] = 0.0;
    uint32_t *tab = (uint32_t *)malloc(sizeof(uint32_t) * n * n);
    double *wn = (double *)malloc(nr * sizeof(double));
    orc_ngenic_seedtable(cfg->seed, n, tab);
    orc_ngenic_fill(tab, n, wn);
    free(tab);
    double t1 = wtime();
    double *phi = (double *)calloc(nr, sizeof(double));
    double *phi2 = cfg->lpt_order > 1 ? (double *)calloc(nr, sizeof(double)) : NULL;
    double *phi3 = NULL, *A3x = NULL, *A3y = NULL, *A3z = NULL;
    if (cfg->lpt_order > 2) {
        phi3 = (double *)calloc(nr, sizeof(double));
        A3x = (double *)calloc(nr, sizeof(double));
        A3y = (double *)calloc(nr, sizeof(double));
        A3z = (double *)calloc(nr, sizeof(double));
    }
    orc_lpt_potentials(cfg, c, wn, phi, phi2, phi3, A3x, A3y, A3z);
    double t2 = wtime();
    orc_plt plt_s, *plt = NULL;
    if (cfg->plt) {
        if (orc_plt_init(&plt_s, cfg->lattice, 64) != 0) return -1;
        plt = &plt_s;
    }
    double t3 = wtime();
    orc_lpt_coef co;
    orc_lpt_coefs(cfg, c, &co);
    double *tmp = (double *)malloc(nr * sizeof(double));
    for (int comp = 0; comp < 6; ++comp) {
        orc_lpt_component_k(cfg, &co, plt, phi, phi2, phi3, A3x, A3y, A3z, comp, tmp);
        orc_fft3d_c2r(tmp, n, n, n);
        if (fields6) memcpy(fields6 + comp * nr, tmp, nr * sizeof(double));
        if (pos && comp < 3) orc_lattice_set_positions(n, cfg->L, cfg->lattice, comp, cfg->lunit, tmp, pos);
        if (vel && comp >= 3) orc_lattice_set_velocities(n, cfg->L, cfg->lattice, comp - 3, tmp, vel);
    }
    double t4 = wtime();
    if (st) { st[0] = t1 - t0; st[1] = t2 - t1; st[2] = t3 - t2; st[3] = t4 - t3; st[4] = t4 - t0; }
    free(tmp); free(wn); free(phi); free(phi2); free(phi3); free(A3x); free(A3y); free(A3z);
    if (plt) orc_plt_free(plt);
    return 0;
}

/* ---------------------------------------------- exported k-space helpers */
/* apply_InverseLaplacian (grid_fft.hh:440-449): -f/|k|^2, then DC := 0 */
void orc_inverse_laplacian(int n, double L, double *f) { inverse_laplacian(n, L, f); }
/* zero_DC_mode, k-space branch (grid_fft.hh:844-850) */
void orc_zero_dc(double *f) { f[0] = 0.0; f[1] = 0.0; }
/* operator*= (grid_fft.hh:400-412): every stored real times s */
void orc_scale(int n, double *f, double s) { scale(n, f, s); }
