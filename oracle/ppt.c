/*
 * Oracle: semiclassical perturbation theory fields of the field_eulerian
 * output (ic_generator.cc:786-897), TEST INFRASTRUCTURE ONLY.
 *
 *   rho  = sqrt(1 + C_species delta_bc) = 1 (no baryons: C_species = 0)
 *   hbar = max |phi(neighbour) - phi| / pi / D+ * 1.01 over i < n-1 (the
 *          reference's loop, ic_generator.cc:815-841)
 *   psi  = rho exp(i (phi + phi2) / (hbar D+))   (phi alone for LPTorder 1)
 *   psi  = c2c^-1[ c2c[psi] exp(-i hbar |k|^2 D+ / 2) ]
 *   psi *= exp(i phi2 / (hbar D+))               (LPTorder >= 2)
 *   delta = |psi|^2 - 1
 *   v_d  = vunit Re[(psi* g - psi g*) / (2 i / hbar)] / (1 + delta),
 *          g = c2c^-1[ i k_d c2c[psi] ], vunit = D+ vfac / L * velocity_unit
 * phi, phi2 are the growth-scaled k-space potentials (orc_lpt_potentials);
 * the outputs are real n x n x (n+2) buffers.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static const double PI_ = 3.14159265358979323846;

static double kget(int n, double kf, int i) { return ((double)i - (double)(i > n / 2) * n) * kf; }

int orc_ppt_fields(const orc_ic_cfg *cfg, const orc_lpt_coef *co, const double *phi_k, const double *phi2_k,
                   double velocity_unit, double *delta, double *vx, double *vy, double *vz, double *hbar_out) {
    const int n = cfg->n;
    const size_t nr = (size_t)n * n * (n + 2), n3 = (size_t)n * n * n;
    const double L = cfg->L, kf = 2.0 * PI_ / L, D = co->Dplus0;
    const int two = cfg->lpt_order >= 2;
    double *phi = (double *)malloc(nr * sizeof(double));
    double *phi2 = two ? (double *)malloc(nr * sizeof(double)) : NULL;
    double *psi = (double *)malloc(2 * n3 * sizeof(double));
    double *g = (double *)malloc(2 * n3 * sizeof(double));
    if (!phi || (two && !phi2) || !psi || !g) return -1;
    memcpy(phi, phi_k, nr * sizeof(double));
    orc_fft3d_c2r(phi, n, n, n);
    if (two) {
        memcpy(phi2, phi2_k, nr * sizeof(double));
        orc_fft3d_c2r(phi2, n, n, n);
    }
#define R(f, i, j, k) (f)[((size_t)(i) * n + (j)) * (n + 2) + (k)]
    double maxdphi = -1.0;
#pragma omp parallel for reduction(max : maxdphi) schedule(static)
    for (int i = 0; i < n - 1; ++i) {
        const int ir = (i + 1) % n;
        for (int j = 0; j < n; ++j) {
            const int jr = (j + 1) % n;
            for (int k = 0; k < n; ++k) {
                const int kr = (k + 1) % n;
                const double c = R(phi, i, j, k);
                const double dx = fabs(R(phi, ir, j, k) - c), dy = fabs(R(phi, i, jr, k) - c),
                             dz = fabs(R(phi, i, j, kr) - c);
                maxdphi = fmax(maxdphi, fmax(dx, fmax(dy, dz)));
            }
        }
    }
    const double hbar = maxdphi / PI_ / D * 1.01;
    if (hbar_out) *hbar_out = hbar;
    /* psi = exp(i (phi + phi2) / hbar / D+) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                const double a = (R(phi, i, j, k) + (two ? R(phi2, i, j, k) : 0.0)) / hbar / D;
                const size_t c = ((size_t)i * n + j) * n + k;
                psi[2 * c] = cos(a);
                psi[2 * c + 1] = sin(a);
            }
    orc_fft3d_c2c(psi, n, n, n, -1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                const double kx = kget(n, kf, i), ky = kget(n, kf, j), kz = kget(n, kf, k);
                const double th = -0.5 * hbar * (kx * kx + ky * ky + kz * kz) * D;
                const size_t c = ((size_t)i * n + j) * n + k;
                const double re = psi[2 * c], im = psi[2 * c + 1], cs = cos(th), sn = sin(th);
                psi[2 * c] = re * cs - im * sn;
                psi[2 * c + 1] = re * sn + im * cs;
            }
    orc_fft3d_c2c(psi, n, n, n, +1);
    if (two) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                for (int k = 0; k < n; ++k) {
                    const double a = R(phi2, i, j, k) / hbar / D;
                    const size_t c = ((size_t)i * n + j) * n + k;
                    const double re = psi[2 * c], im = psi[2 * c + 1], cs = cos(a), sn = sin(a);
                    psi[2 * c] = re * cs - im * sn;
                    psi[2 * c + 1] = re * sn + im * cs;
                }
    }
    memset(delta, 0, nr * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                const size_t c = ((size_t)i * n + j) * n + k;
                R(delta, i, j, k) = psi[2 * c] * psi[2 * c] + psi[2 * c + 1] * psi[2 * c + 1] - 1.0;
            }
    const double vunit = D * co->vfac / L * velocity_unit;
    double *v[3] = {vx, vy, vz};
    for (int d = 0; d < 3; ++d) {
        memcpy(g, psi, 2 * n3 * sizeof(double));
        orc_fft3d_c2c(g, n, n, n, -1);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                for (int k = 0; k < n; ++k) {
                    const int ijk[3] = {i, j, k};
                    const double kd = kget(n, kf, ijk[d]);
                    const size_t c = ((size_t)i * n + j) * n + k;
                    const double re = g[2 * c], im = g[2 * c + 1];
                    g[2 * c] = -kd * im;  /* x * (0 + i k_d) */
                    g[2 * c + 1] = kd * re;
                }
        orc_fft3d_c2c(g, n, n, n, +1);
        memset(v[d], 0, nr * sizeof(double));
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                for (int k = 0; k < n; ++k) {
                    const size_t c = ((size_t)i * n + j) * n + k;
                    const double pr = psi[2 * c], pi = psi[2 * c + 1], gr = g[2 * c], gi = g[2 * c + 1];
                    /* psi* g - psi g* = 2 i Im(psi* g); divided by 2 i / hbar */
                    const double im = pr * gi - pi * gr;
                    R(v[d], i, j, k) = vunit * (hbar * im) / (1.0 + R(delta, i, j, k));
                }
    }
#undef R
    free(phi); free(phi2); free(psi); free(g);
    return 0;
}
