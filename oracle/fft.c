/*
 * Oracle: FP64 FFTs in the Grid_FFT conventions (TEST INFRASTRUCTURE ONLY).
 *
 * 1D: mixed-radix Stockham autosort (radices 4, 2, 3, 5, and a direct DFT for
 * any other prime factor) with twiddles evaluated in long double.
 * 3D: in-place r2c / c2r on the padded layout double[n0][n1][n2+2]
 * (grid_fft.cc:45-54), unitary scaling 1/sqrt(n0 n1 n2) in both directions
 * (grid_fft.cc:69, 199-205), FFTW sign convention (forward e^{-ikx}).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static void twiddle(long num, long den, int sign, double *re, double *im) {
    /* exp(sign * 2 pi i num / den), evaluated in long double */
    long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)(num % den) / (long double)den;
    *re = (double)cosl(a);
    *im = (double)(sign * sinl(a));
}

/* plan: factorisation + per-stage twiddles + radix DFT matrices; tables hold
 * exp(+2 pi i ...), the transform direction multiplies the imaginary parts */
typedef struct {
    int n, nf, fac[64];
    double *tw[64];     /* stage f: Ns*R entries (re,im) of exp(2 pi i r k/(Ns R)) */
    double *dm[64];     /* stage f: R*R entries of exp(2 pi i r q / R)             */
} plan_t;

static plan_t *g_plans[64];
static int g_nplans = 0;

static plan_t *make_plan(int n) {
    plan_t *p = (plan_t *)calloc(1, sizeof(plan_t));
    p->n = n;
    int t = n;
    while (t % 4 == 0) { p->fac[p->nf++] = 4; t /= 4; }
    while (t % 2 == 0) { p->fac[p->nf++] = 2; t /= 2; }
    for (int q = 3; t > 1; q += 2)
        while (t % q == 0) { p->fac[p->nf++] = q; t /= q; }
    int Ns = 1;
    for (int f = 0; f < p->nf; ++f) {
        int R = p->fac[f];
        p->tw[f] = (double *)malloc(sizeof(double) * 2 * Ns * R);
        for (int k = 0; k < Ns; ++k)
            for (int r = 0; r < R; ++r)
                twiddle((long)r * k, (long)Ns * R, +1, &p->tw[f][2 * (k * R + r)], &p->tw[f][2 * (k * R + r) + 1]);
        p->dm[f] = (double *)malloc(sizeof(double) * 2 * R * R);
        for (int q = 0; q < R; ++q)
            for (int r = 0; r < R; ++r)
                twiddle((long)r * q, R, +1, &p->dm[f][2 * (q * R + r)], &p->dm[f][2 * (q * R + r) + 1]);
        Ns *= R;
    }
    return p;
}

static plan_t *get_plan(int n) {
    plan_t *res = NULL;
#pragma omp critical(orc_fft_plans)
    {
        for (int i = 0; i < g_nplans; ++i)
            if (g_plans[i]->n == n) { res = g_plans[i]; break; }
        if (!res) {
            res = make_plan(n);
            if (g_nplans < 64) g_plans[g_nplans++] = res;
        }
    }
    return res;
}

/* one Stockham stage: input x, output y, radix R, current span Ns */
static void stage(const plan_t *p, int f, const double *x, double *y, int Ns, int sign) {
    const int n = p->n, R = p->fac[f], m = n / R;
    const double *tw = p->tw[f], *dm = p->dm[f];
    double vr[4096], vi[4096];
    for (int j = 0; j < m; ++j) {
        int k = j % Ns;
        for (int r = 0; r < R; ++r) {
            double ar = x[2 * (j + r * m)], ai = x[2 * (j + r * m) + 1];
            double tr = tw[2 * (k * R + r)], ti = sign * tw[2 * (k * R + r) + 1];
            vr[r] = ar * tr - ai * ti;
            vi[r] = ar * ti + ai * tr;
        }
        int base = (j / Ns) * Ns * R + k;
        for (int q = 0; q < R; ++q) {
            double sr = 0, si = 0;
            for (int r = 0; r < R; ++r) {
                double tr = dm[2 * (q * R + r)], ti = sign * dm[2 * (q * R + r) + 1];
                sr += vr[r] * tr - vi[r] * ti;
                si += vr[r] * ti + vi[r] * tr;
            }
            y[2 * (base + q * Ns)] = sr;
            y[2 * (base + q * Ns) + 1] = si;
        }
    }
}

void orc_fft1d(double *z, int n, int sign) {
    if (n <= 1) return;
    const plan_t *p = get_plan(n);
    if (p->fac[p->nf - 1] > 4096) { /* direct DFT for a huge prime */
        double *o = (double *)malloc(2 * sizeof(double) * n);
        for (int q = 0; q < n; ++q) {
            long double sr = 0, si = 0;
            for (int r = 0; r < n; ++r) {
                double tr, ti;
                twiddle((long)r * q, n, sign, &tr, &ti);
                sr += (long double)z[2 * r] * tr - (long double)z[2 * r + 1] * ti;
                si += (long double)z[2 * r] * ti + (long double)z[2 * r + 1] * tr;
            }
            o[2 * q] = (double)sr; o[2 * q + 1] = (double)si;
        }
        memcpy(z, o, 2 * sizeof(double) * n);
        free(o);
        return;
    }
    double *buf = (double *)malloc(2 * sizeof(double) * n);
    double *a = z, *b = buf;
    int Ns = 1;
    for (int f = 0; f < p->nf; ++f) {
        stage(p, f, a, b, Ns, sign);
        Ns *= p->fac[f];
        double *tmp = a; a = b; b = tmp;
    }
    if (a != z) memcpy(z, a, 2 * sizeof(double) * n);
    free(buf);
}

/* transform all pencils along axis `ax` (0 or 1) of the complex half-spectrum */
static void pencils(double *d, int n0, int n1, int nc, int ax, int sign) {
    int len = ax == 0 ? n0 : n1;
    get_plan(len);
    long stride = ax == 0 ? (long)n1 * nc : nc;
    long outer = ax == 0 ? n1 : n0;
    long ostride = ax == 0 ? nc : (long)n1 * nc;
#pragma omp parallel
    {
        double *z = (double *)malloc(2 * sizeof(double) * len);
#pragma omp for collapse(2) schedule(static)
        for (long o = 0; o < outer; ++o) {
            for (long k = 0; k < nc; ++k) {
                double *base = d + 2 * (o * ostride + k);
                for (int p = 0; p < len; ++p) {
                    z[2 * p] = base[2 * p * stride];
                    z[2 * p + 1] = base[2 * p * stride + 1];
                }
                orc_fft1d(z, len, sign);
                for (int p = 0; p < len; ++p) {
                    base[2 * p * stride] = z[2 * p];
                    base[2 * p * stride + 1] = z[2 * p + 1];
                }
            }
        }
        free(z);
    }
}

void orc_fft3d_r2c(double *d, int n0, int n1, int n2) {
    const int nc = n2 / 2 + 1;
    const double norm = 1.0 / sqrt((double)n0 * (double)n1 * (double)n2);
    get_plan(n2);
#pragma omp parallel
    {
        double *z = (double *)malloc(2 * sizeof(double) * n2);
#pragma omp for schedule(static)
        for (long row = 0; row < (long)n0 * n1; ++row) {
            double *r = d + row * (n2 + 2);
            for (int p = 0; p < n2; ++p) { z[2 * p] = r[p]; z[2 * p + 1] = 0.0; }
            orc_fft1d(z, n2, -1);
            for (int p = 0; p < 2 * nc; ++p) r[p] = z[p];
        }
        free(z);
    }
    pencils(d, n0, n1, nc, 1, -1);
    pencils(d, n0, n1, nc, 0, -1);
    long tot = (long)n0 * n1 * 2 * nc;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < tot; ++i) d[i] *= norm;
}

void orc_fft3d_c2r(double *d, int n0, int n1, int n2) {
    const int nc = n2 / 2 + 1;
    const double norm = 1.0 / sqrt((double)n0 * (double)n1 * (double)n2);
    pencils(d, n0, n1, nc, 0, +1);
    pencils(d, n0, n1, nc, 1, +1);
    get_plan(n2);
#pragma omp parallel
    {
        double *z = (double *)malloc(2 * sizeof(double) * n2);
#pragma omp for schedule(static)
        for (long row = 0; row < (long)n0 * n1; ++row) {
            double *r = d + row * (n2 + 2);
            /* Hermitian completion; imaginary parts of DC and Nyquist dropped */
            for (int k = 0; k < nc; ++k) { z[2 * k] = r[2 * k]; z[2 * k + 1] = r[2 * k + 1]; }
            z[1] = 0.0;
            if (n2 % 2 == 0) z[2 * (n2 / 2) + 1] = 0.0;
            for (int k = nc; k < n2; ++k) { z[2 * k] = r[2 * (n2 - k)]; z[2 * k + 1] = -r[2 * (n2 - k) + 1]; }
            orc_fft1d(z, n2, +1);
            for (int p = 0; p < n2; ++p) r[p] = z[2 * p] * norm;
            r[n2] = 0.0; r[n2 + 1] = 0.0;
        }
        free(z);
    }
}

/* Grid_FFT<ccomplex_t> transform (grid_fft.cc:56-63, 207-270) of a complex
 * [n0][n1][n2] array, in place, unitary, sign -1 forward / +1 backward.
 * pinned: numpy.fft.fftn / ifftn(norm="ortho"). */
void orc_fft3d_c2c(double *d, int n0, int n1, int n2, int sign) {
    get_plan(n2);
    const long rows = (long)n0 * n1;
#pragma omp parallel for schedule(static)
    for (long r = 0; r < rows; ++r) orc_fft1d(d + 2 * r * n2, n2, sign);
    pencils(d, n0, n1, n2, 1, sign);
    pencils(d, n0, n1, n2, 0, sign);
    const double norm = 1.0 / sqrt((double)n0 * (double)n1 * (double)n2);
    const long tot = 2L * n0 * n1 * n2;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < tot; ++i) d[i] *= norm;
}
