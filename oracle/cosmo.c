/*
 * Oracle: cosmology host scalars (TEST INFRASTRUCTURE ONLY).
 *
 *   parameters      cosmology_parameters.hh:80-241, cosmology_parameters.cc:110-169
 *   H(a)            cosmology_calculator.hh:334-342
 *   growth D, f     cosmology_calculator.hh:92-152 with the adaptive Cash-Karp
 *                   step of math/ode_integrate.hh:39-111 (including its retry
 *                   behaviour: a rejected trial step is retried from the state
 *                   the rejected step produced, ode_integrate.hh:79-91)
 *   D(a), f(a)      natural cubic spline in (log a, log y), interpolate.hh:30-103
 *   vfac            cosmology_calculator.hh:373-376
 *   sigma_8, pnorm  cosmology_calculator.hh:392-404, 487-511; GSL QAG-61 at
 *                   rel. 1e-10 is replaced by adaptive Gauss-Kronrod 7/15 at
 *                   rel. 1e-13 (same integral, tighter tolerance)
 *   E&H T(k)        transfer_eisenstein.cc:58-201 (Eisenstein & Hu 1998 fit)
 * pinned: scipy solve_ivp / CubicSpline / quad goldens, EdS closed form,
 * E&H goldens from an independent Python evaluation (tests/golden).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* physical constants, physical_constants.hh:28-72 */
static const double PI_ = 3.141592653589793115997963468544185161590576171875;
static const double Mpc_SI = 3.0857e22;
static const double c_SI = 2.99792458e8;
static const double G_SI = 6.6740800e-11;
static const double kB_SI = 1.38064852e-23;
static const double hbar_SI = 1.054571800e-34;

static double sigma_SI(void) {
    return (PI_ * PI_) * (kB_SI * kB_SI * kB_SI * kB_SI) / 60. /
           (hbar_SI * hbar_SI * hbar_SI) / (c_SI * c_SI);
}
static double rhocrit_h2_SI(void) { return 3 * 1e10 / (8 * PI_ * G_SI) / Mpc_SI / Mpc_SI; }

void orc_cosmo_defaults(orc_cosmo *c) {
    memset(c, 0, sizeof(*c));
    c->H0 = 67.742; c->Omega_m = 0.3099; c->Omega_b = 0.048891054;
    c->Omega_k = 0.0; c->w_0 = -1.0; c->w_a = 0.0; c->n_s = 0.96822;
    c->A_s = 2.1064e-9; c->alpha_s = 0.0; c->beta_s = 0.0; c->k_p = 0.05;
    c->N_ur = 2.046; c->m_nu[0] = 0.06; c->m_nu[1] = 0.0; c->m_nu[2] = 0.0;
    c->deg_nu[0] = c->deg_nu[1] = c->deg_nu[2] = 1.0;
    c->Tcmb = 2.7255; c->O_nu_norm = 93.13861;
    c->sigma_8 = -1.0; c->zstart = 100.0; c->ztarget = 0.0;
}

/* ------------------------------------------------------------ E&H ------ */
enum { EH_omhh, EH_obhh, EH_theta, EH_zeq, EH_keq, EH_zd, EH_Rd, EH_Req, EH_s,
       EH_ksilk, EH_ac, EH_bc, EH_ab, EH_bb, EH_bnode, EH_kpeak, EH_sfit, EH_ag,
       EH_h, EH_fb, EH_fc };
#define SQR(x) ((x) * (x))
#define CUBE(x) ((x) * SQR(x))
#define POW4(x) SQR(SQR(x))

static void eh_set(orc_cosmo *c) {
    double *e = c->eh;
    double omhh = c->Omega_m * c->H0 * c->H0 * (0.01 * 0.01);
    double f_baryon = c->Omega_b / c->Omega_m;
    double Tcmb = c->Tcmb;
    if (Tcmb <= 0.0) Tcmb = 2.728;
    double obhh = omhh * f_baryon;
    double theta = Tcmb / 2.7;
    double zeq = 2.50e4 * omhh / POW4(theta);
    double keq = 0.0746 * omhh / SQR(theta);
    double zdb1 = 0.313 * pow(omhh, -0.419) * (1 + 0.607 * pow(omhh, 0.674));
    double zdb2 = 0.238 * pow(omhh, 0.223);
    double zd = 1291 * pow(omhh, 0.251) / (1 + 0.659 * pow(omhh, 0.828)) * (1 + zdb1 * pow(obhh, zdb2));
    double Rd = 31.5 * obhh / POW4(theta) * (1000 / (1 + zd));
    double Req = 31.5 * obhh / POW4(theta) * (1000 / zeq);
    double s = 2. / 3. / keq * sqrt(6. / Req) * log((sqrt(1 + Rd) + sqrt(Rd + Req)) / (1 + sqrt(Req)));
    double ksilk = 1.6 * pow(obhh, 0.52) * pow(omhh, 0.73) * (1 + pow(10.4 * omhh, -0.95));
    double aca1 = pow(46.9 * omhh, 0.670) * (1 + pow(32.1 * omhh, -0.532));
    double aca2 = pow(12.0 * omhh, 0.424) * (1 + pow(45.0 * omhh, -0.582));
    double ac = pow(aca1, -f_baryon) * pow(aca2, -CUBE(f_baryon));
    double bcb1 = 0.944 / (1 + pow(458 * omhh, -0.708));
    double bcb2 = pow(0.395 * omhh, -0.0266);
    double bc = 1.0 / (1 + bcb1 * (pow(1 - f_baryon, bcb2) - 1));
    double y = zeq / (1 + zd);
    double abG = y * (-6. * sqrt(1 + y) + (2. + 3. * y) * log((sqrt(1 + y) + 1) / (sqrt(1 + y) - 1)));
    double ab = 2.07 * keq * s * pow(1 + Rd, -0.75) * abG;
    double bnode = 8.41 * pow(omhh, 0.435);
    double bb = 0.5 + f_baryon + (3. - 2. * f_baryon) * sqrt(pow(17.2 * omhh, 2.0) + 1);
    e[EH_omhh] = omhh; e[EH_obhh] = obhh; e[EH_theta] = theta; e[EH_zeq] = zeq;
    e[EH_keq] = keq; e[EH_zd] = zd; e[EH_Rd] = Rd; e[EH_Req] = Req; e[EH_s] = s;
    e[EH_ksilk] = ksilk; e[EH_ac] = ac; e[EH_bc] = bc; e[EH_ab] = ab; e[EH_bb] = bb;
    e[EH_bnode] = bnode;
    e[EH_h] = c->H0 * 0.01;
    e[EH_fb] = c->Omega_b / c->Omega_m;
    e[EH_fc] = (c->Omega_m - c->Omega_b) / c->Omega_m;
}

/* TFfit_onek baryon and CDM pieces, k in 1/Mpc */
static void eh_onek(const double *e, double k, double *tb, double *tc) {
    k = fabs(k);
    if (k == 0.0) { *tb = 1.0; *tc = 1.0; return; }
    double q = k / 13.41 / e[EH_keq];
    double xx = k * e[EH_s];
    double lnb = log(2.718282 + 1.8 * e[EH_bc] * q);
    double lnnb = log(2.718282 + 1.8 * q);
    double Ca = 14.2 / e[EH_ac] + 386.0 / (1 + 69.9 * pow(q, 1.08));
    double Cna = 14.2 + 386.0 / (1 + 69.9 * pow(q, 1.08));
    double f = 1.0 / (1.0 + POW4(xx / 5.4));
    double T_c = f * lnb / (lnb + Cna * SQR(q)) + (1 - f) * lnb / (lnb + Ca * SQR(q));
    double st = e[EH_s] * pow(1. + CUBE(e[EH_bnode] / xx), -1. / 3.);
    double xt = k * st;
    double T0 = lnnb / (lnnb + Cna * SQR(q));
    double T_b = sin(xt) / (xt) * (T0 / (1. + SQR(xx / 5.2)) +
                                   e[EH_ab] / (1. + CUBE(e[EH_bb] / xx)) * exp(-pow(k / e[EH_ksilk], 1.4)));
    *tb = T_b; *tc = T_c;
}

static double running_factor(const orc_cosmo *c, double k) {
    if (c->alpha_s == 0.0 && c->beta_s == 0.0) return 1.0;
    double lnk = log(k / c->k_p);
    double run = 0.5 * (c->alpha_s * lnk + c->beta_s * lnk * lnk / 3.0);
    return pow(k / c->k_p, 0.5 * run);
}

double orc_tk_eh(const orc_cosmo *c, double k) {
    double tb, tc;
    eh_onek(c->eh, k * c->eh[EH_h], &tb, &tc);
    return (c->eh[EH_fb] * tb + c->eh[EH_fc] * tc) * running_factor(c, k);
}

static double spline_eval(int n, const double *x, const double *y, const double *m, double xv);

/* transfer_CAMB_file.cc:195-199, 244-245: k clamped to the second row, the
 * running factor taken at k h */
double orc_tk(const orc_cosmo *c, double k) {
    if (c->tk_n <= 0) return orc_tk_eh(c, k);
    double kc = k > c->tk_k[1] ? k : c->tk_k[1];
    return exp(spline_eval(c->tk_n, c->tk_lx, c->tk_ly, c->tk_m, log(kc))) * running_factor(c, kc * c->h);
}

double orc_amplitude(const orc_cosmo *c, double k) {
    return pow(k, 0.5 * c->n_s) * orc_tk(c, k) * c->sqrtpnorm;
}

/* ------------------------------------------------------------ H(a) ----- */
double orc_cosmo_H(const orc_cosmo *c, double a) {
    double HH2 = 0.0;
    HH2 += c->Omega_r / (a * a * a * a);
    HH2 += c->Omega_m / (a * a * a);
    HH2 += c->Omega_k / (a * a);
    HH2 += c->Omega_DE * pow(a, -3. * (1. + c->w_0 + c->w_a)) * exp(-3. * (1.0 - a) * c->w_a);
    return c->H0 * sqrt(HH2);
}

/* --------------------------------------------------------- growth ODE -- */
static void rhs(const orc_cosmo *c, const double y[3], double dy[3]) {
    double a = y[0], D = y[1], Dp = y[2];
    double H = orc_cosmo_H(c, a);
    dy[0] = a * a * H;
    dy[1] = Dp;
    dy[2] = -a * H * Dp + 3.0 / 2.0 * c->Omega_m * pow(c->H0, 2) * D / a;
}

/* Cash-Karp 5(4) step; advances y in place, returns the error estimate */
static void ckrk5(const orc_cosmo *c, double h, double y[3], double yerr[3]) {
    static const double
        b21 = 0.20, b31 = 3.0 / 40.0, b32 = 9.0 / 40.0,
        b41 = 0.30, b42 = -0.90, b43 = 1.20,
        b51 = -11.0 / 54.0, b52 = 2.50, b53 = -70.0 / 27.0, b54 = 35.0 / 27.0,
        b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0, b63 = 575.0 / 13824.0,
        b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0,
        c1 = 37.0 / 378.0, c3 = 250.0 / 621.0, c4 = 125.0 / 594.0, c6 = 512.0 / 1771.0;
    const double dc1 = c1 - 2825.0 / 27648.0, dc3 = c3 - 18575.0 / 48384.0,
                 dc4 = c4 - 13525.0 / 55296.0, dc5 = -277.0 / 14336.0, dc6 = c6 - 0.250;
    double k1[3], k2[3], k3[3], k4[3], k5[3], k6[3], t[3];
    int i;
    rhs(c, y, k1); for (i = 0; i < 3; ++i) k1[i] *= h;
    for (i = 0; i < 3; ++i) t[i] = y[i] + b21 * k1[i];
    rhs(c, t, k2); for (i = 0; i < 3; ++i) k2[i] *= h;
    for (i = 0; i < 3; ++i) t[i] = y[i] + b31 * k1[i] + b32 * k2[i];
    rhs(c, t, k3); for (i = 0; i < 3; ++i) k3[i] *= h;
    for (i = 0; i < 3; ++i) t[i] = y[i] + b41 * k1[i] + b42 * k2[i] + b43 * k3[i];
    rhs(c, t, k4); for (i = 0; i < 3; ++i) k4[i] *= h;
    for (i = 0; i < 3; ++i) t[i] = y[i] + b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i];
    rhs(c, t, k5); for (i = 0; i < 3; ++i) k5[i] *= h;
    for (i = 0; i < 3; ++i) t[i] = y[i] + b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i];
    rhs(c, t, k6); for (i = 0; i < 3; ++i) k6[i] *= h;
    for (i = 0; i < 3; ++i) {
        y[i] += c1 * k1[i] + c3 * k3[i] + c4 * k4[i] + c6 * k6[i];
        yerr[i] = dc1 * k1[i] + dc3 * k3[i] + dc4 * k4[i] + dc5 * k5[i] + dc6 * k6[i];
    }
}

/* natural cubic spline second derivatives (tridiagonal solve) */
static void spline_init(int n, const double *x, const double *y, double *m) {
    double *u = (double *)malloc(sizeof(double) * n);
    m[0] = 0.0; u[0] = 0.0;
    for (int i = 1; i < n - 1; ++i) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * m[i - 1] + 2.0;
        m[i] = (sig - 1.0) / p;
        u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    m[n - 1] = 0.0;
    for (int k = n - 2; k >= 0; --k) m[k] = m[k] * m[k + 1] + u[k];
    free(u);
}
static double spline_eval(int n, const double *x, const double *y, const double *m, double xv) {
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (hi + lo) >> 1;
        if (x[mid] > xv) hi = mid; else lo = mid;
    }
    double h = x[hi] - x[lo];
    double A = (x[hi] - xv) / h, B = (xv - x[lo]) / h;
    return A * y[lo] + B * y[hi] + ((A * A * A - A) * m[lo] + (B * B * B - B) * m[hi]) * (h * h) / 6.0;
}

static int compute_growth(orc_cosmo *c) {
    const double a0 = 1e-10, D0 = a0;
    const double Dprime0 = 2.0 * D0 * orc_cosmo_H(c, a0) / pow(c_SI, 2);
    double t = 1.0 / (a0 * orc_cosmo_H(c, a0));
    double y[3] = {a0, D0, Dprime0};
    double dt = 1e-9;
    const double amax = 2.0, eps = 1e-10;
    int cap = 4096, n = 0;
    double *ta = (double *)malloc(sizeof(double) * cap);
    double *tD = (double *)malloc(sizeof(double) * cap);
    double *tf = (double *)malloc(sizeof(double) * cap);
    while (y[0] < amax) {
        double d0[3], ysc[3];
        rhs(c, y, d0);
        for (int i = 0; i < 3; ++i) ysc[i] = fabs(y[i]) + dt * fabs(d0[i]);
        /* rk_step_qs: the trial state persists across rejected trials */
        double h = dt, ytemp[3] = {y[0], y[1], y[2]}, yerr[3], errmax;
        for (;;) {
            ckrk5(c, h, ytemp, yerr);
            errmax = 0.0;
            for (int i = 0; i < 3; ++i) errmax = fmax(errmax, fabs(yerr[i] / ysc[i]));
            errmax = errmax / eps;
            if (errmax > 1.0) {
                h *= fmax(0.1, 0.9 * pow(errmax, -0.25));
                if (t + h == t) { free(ta); free(tD); free(tf); return -1; }
                continue;
            }
            dt = (errmax > 1.89e-4) ? h * 0.9 * pow(errmax, -0.2) : 5 * h;
            t += h;
            y[0] = ytemp[0]; y[1] = ytemp[1]; y[2] = ytemp[2];
            break;
        }
        if (n == cap) {
            cap *= 2;
            ta = (double *)realloc(ta, sizeof(double) * cap);
            tD = (double *)realloc(tD, sizeof(double) * cap);
            tf = (double *)realloc(tf, sizeof(double) * cap);
        }
        ta[n] = y[0]; tD[n] = y[1]; tf[n] = y[2];
        ++n;
    }
    for (int i = 0; i < n; ++i) tf[i] = tf[i] / (ta[i] * orc_cosmo_H(c, ta[i]) * tD[i]);
    c->ntab = n;
    c->tab_loga = ta; c->tab_logD = tD; c->tab_logf = tf;
    for (int i = 0; i < n; ++i) { ta[i] = log(ta[i]); tD[i] = log(tD[i]); tf[i] = log(tf[i]); }
    c->spl_D = (double *)malloc(sizeof(double) * n);
    c->spl_f = (double *)malloc(sizeof(double) * n);
    spline_init(n, ta, tD, c->spl_D);
    spline_init(n, ta, tf, c->spl_f);
    return 0;
}

static double D_raw(const orc_cosmo *c, double a) {
    return exp(spline_eval(c->ntab, c->tab_loga, c->tab_logD, c->spl_D, log(a)));
}
double orc_cosmo_D(const orc_cosmo *c, double a) { return D_raw(c, a) / c->Dnow; }
double orc_cosmo_f(const orc_cosmo *c, double a) {
    return exp(spline_eval(c->ntab, c->tab_loga, c->tab_logf, c->spl_f, log(a)));
}
double orc_cosmo_vfact(const orc_cosmo *c, double a) {
    return orc_cosmo_f(c, a) * a * orc_cosmo_H(c, a) / c->h;
}
int orc_cosmo_growth_table(const orc_cosmo *c, double *a, double *D, double *f, int max) {
    int n = c->ntab < max ? c->ntab : max;
    for (int i = 0; i < n; ++i) {
        a[i] = exp(c->tab_loga[i]); D[i] = exp(c->tab_logD[i]); f[i] = exp(c->tab_logf[i]);
    }
    return c->ntab;
}

/* ------------------------------------------------ sigma_8 quadrature --- */
static const double XGK[8] = {
    0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
    0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
    0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double WGK[8] = {
    0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
    0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
    0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
    0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double WG[4] = {
    0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
    0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

static double dsigma8(const orc_cosmo *c, double k) {
    const double x = k * 8.0;
    const double w = (x < 0.001) ? 1.0 - 0.1 * x * x : 3.0 * (sin(x) - x * cos(x)) / (x * x * x);
    double tf = orc_tk(c, k);
    return k * k * w * w * pow(k, c->n_s) * tf * tf;
}

static void gk15(const orc_cosmo *c, double a, double b, double *res, double *err) {
    double cen = 0.5 * (a + b), hl = 0.5 * (b - a);
    double fc = dsigma8(c, cen);
    double rk = fc * WGK[7], rg = fc * WG[3];
    for (int j = 0; j < 7; ++j) {
        double dx = hl * XGK[j];
        double f1 = dsigma8(c, cen - dx), f2 = dsigma8(c, cen + dx);
        rk += WGK[j] * (f1 + f2);
        if (j & 1) rg += WG[j / 2] * (f1 + f2);
    }
    *res = rk * hl;
    *err = fabs((rk - rg) * hl);
}

double orc_integrate_sigma8(const orc_cosmo *c, double a, double b, double rel) {
    int cap = 20000, n = 1;
    double *lo = (double *)malloc(sizeof(double) * cap), *hi = (double *)malloc(sizeof(double) * cap);
    double *rs = (double *)malloc(sizeof(double) * cap), *er = (double *)malloc(sizeof(double) * cap);
    lo[0] = a; hi[0] = b;
    gk15(c, a, b, &rs[0], &er[0]);
    double tot = rs[0], terr = er[0];
    while (terr > rel * fabs(tot) && n < cap - 1) {
        int w = 0;
        for (int i = 1; i < n; ++i) if (er[i] > er[w]) w = i;
        double m = 0.5 * (lo[w] + hi[w]);
        double r1, e1, r2, e2;
        gk15(c, lo[w], m, &r1, &e1);
        gk15(c, m, hi[w], &r2, &e2);
        lo[n] = m; hi[n] = hi[w]; rs[n] = r2; er[n] = e2;
        hi[w] = m; rs[w] = r1; er[w] = e1;
        ++n;
        tot = 0; terr = 0;
        for (int i = 0; i < n; ++i) { tot += rs[i]; terr += er[i]; }
    }
    free(lo); free(hi); free(rs); free(er);
    return tot;
}

int orc_cosmo_init(orc_cosmo *c) {
    c->h = c->H0 / 100.0;
    double h = c->h;
    double sum_m_nu = c->m_nu[0] * c->deg_nu[0] + c->m_nu[1] * c->deg_nu[1] + c->m_nu[2] * c->deg_nu[2];
    c->Omega_nu_massive = sum_m_nu / (c->O_nu_norm * h * h);
    c->Omega_gamma = 4 * sigma_SI() / pow(c_SI, 3) * pow(c->Tcmb, 4.0) / rhocrit_h2_SI() / (h * h);
    double Omega_nu_massless = c->N_ur * c->Omega_gamma * 7. / 8. * pow(4. / 11., 4. / 3.);
    c->Omega_r = c->Omega_gamma + Omega_nu_massless;
    c->Omega_c = c->Omega_m - c->Omega_b - c->Omega_nu_massive;
    if (c->zero_radiation) c->Omega_r = 0.0;
    c->Omega_DE = 1.0 - c->Omega_m - c->Omega_r - c->Omega_k;

    if (compute_growth(c) != 0) return -1;
    c->Dnow = D_raw(c, 1.0);
    double astart = 1.0 / (1.0 + c->zstart), atarget = 1.0 / (1.0 + c->ztarget);
    c->Dplus_start = D_raw(c, astart) / c->Dnow;
    c->Dplus_target = D_raw(c, atarget) / c->Dnow;

    eh_set(c);
    double kmin = 1e-4, kmax = 1e4;   /* transfer_eisenstein.cc:245-253 */
    if (c->tk_n > 0) {
        if (c->tk_n < 6) return -1;
        int n = c->tk_n;
        c->tk_lx = (double *)malloc(sizeof(double) * n);
        c->tk_ly = (double *)malloc(sizeof(double) * n);
        c->tk_m = (double *)malloc(sizeof(double) * n);
        for (int i = 0; i < n; ++i) {
            if (!(c->tk_k[i] > 0.0) || !(c->tk_T[i] > 0.0)) return -1;
            c->tk_lx[i] = log(c->tk_k[i]);
            c->tk_ly[i] = log(c->tk_T[i]);
        }
        spline_init(n, c->tk_lx, c->tk_ly, c->tk_m);
        kmin = c->tk_k[1];                /* transfer_CAMB_file.cc:156-158 */
        kmax = c->tk_k[n - 2];
    }
    c->sigma8_measured = sqrt(4.0 * PI_ * orc_integrate_sigma8(c, kmin, kmax, 1e-13));
    if (!(c->A_s > -1.0))
        c->pnorm = c->sigma_8 * c->sigma_8 / (c->sigma8_measured * c->sigma8_measured);
    else
        c->pnorm = c->A_s * pow(c->k_p / h, 1.0 - c->n_s) * (h * h * h * h) /
                   (4 * PI_ * c->Dplus_target * c->Dplus_target);
    c->sqrtpnorm = sqrt(c->pnorm);
    c->f_start = orc_cosmo_f(c, astart);
    c->vfac_start = orc_cosmo_vfact(c, astart);
    return 0;
}

void orc_cosmo_free(orc_cosmo *c) {
    free(c->tab_loga); free(c->tab_logD); free(c->tab_logf); free(c->spl_D); free(c->spl_f);
    c->tab_loga = c->tab_logD = c->tab_logf = c->spl_D = c->spl_f = NULL;
    free(c->tk_lx); free(c->tk_ly); free(c->tk_m);
    c->tk_lx = c->tk_ly = c->tk_m = NULL;
}
