/*
 * Oracle: correctly rounded log and sin/cos (TEST INFRASTRUCTURE ONLY).
 *
 * random_ngenic.cc:110-121 turns RANLUX uniforms into Gaussian modes with
 * std::log, std::cos and std::sin.  To make host and device draws bit-exact,
 * both sides evaluate these three functions the same way: in double-double
 * arithmetic (Dekker / Knuth error-free transforms, fma based), rounded once
 * to double.  The result is the correctly rounded value except when the exact
 * result lies within ~2^-100 (relative) of a rounding boundary.
 *
 *   log(x):  x = m 2^e, m in [sqrt(1/2), sqrt(2));  log m = 2 atanh(s),
 *            s = (m-1)/(m+1), |s| <= 0.1716, series in s^2 (Horner, dd).
 *   sin/cos: reduce by pi/2 with a triple-double pi/2 (Cody-Waite style,
 *            exact products via fma), Taylor series on |r| <= pi/4 in dd.
 *
 * Compiled with -ffp-contract=off: every fused multiply-add is explicit.
 */
#include <math.h>
#include "oracle.h"

typedef struct { double hi, lo; } dd_t;

static inline dd_t two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    dd_t r = {s, e};
    return r;
}
static inline dd_t fast_two_sum(double a, double b) {
    double s = a + b;
    double e = b - (s - a);
    dd_t r = {s, e};
    return r;
}
static inline dd_t two_prod(double a, double b) {
    double p = a * b;
    double e = fma(a, b, -p);
    dd_t r = {p, e};
    return r;
}
static inline dd_t dd_add(dd_t x, dd_t y) {
    dd_t s = two_sum(x.hi, y.hi);
    dd_t t = two_sum(x.lo, y.lo);
    double c = s.lo + t.hi;
    dd_t v = fast_two_sum(s.hi, c);
    c = t.lo + v.lo;
    return fast_two_sum(v.hi, c);
}
static inline dd_t dd_mul(dd_t x, dd_t y) {
    dd_t p = two_prod(x.hi, y.hi);
    double e = fma(x.hi, y.lo, x.lo * y.hi) + p.lo;
    return fast_two_sum(p.hi, e);
}
static inline dd_t dd_mul_d(dd_t x, double d) {
    dd_t p = two_prod(x.hi, d);
    double e = fma(x.lo, d, p.lo);
    return fast_two_sum(p.hi, e);
}
/* x / y in double-double (one Newton correction on the quotient) */
static inline dd_t dd_div(dd_t x, dd_t y) {
    double q1 = x.hi / y.hi;
    dd_t p = dd_mul_d(y, q1);
    dd_t r = dd_add(x, (dd_t){-p.hi, -p.lo});
    double q2 = r.hi / y.hi;
    p = dd_mul_d(y, q2);
    r = dd_add(r, (dd_t){-p.hi, -p.lo});
    double q3 = r.hi / y.hi;
    dd_t q = fast_two_sum(q1, q2);
    return dd_add(q, (dd_t){q3, 0.0});
}
/* 1/d as a double-double for a small exact integer d */
static inline dd_t dd_recip_d(double d) {
    double hi = 1.0 / d;
    double r = fma(-hi, d, 1.0);          /* exact residual */
    dd_t out = {hi, r / d};
    return out;
}

/* ln 2 = LN2_HI + LN2_LO */
static const double LN2_HI = 6.93147180559945286227e-01;
static const double LN2_LO = 2.31904681384629955842e-17;
/* pi/2 = PIO2_1 + PIO2_2 + PIO2_3 (each the rounding of the remainder) */
static const double PIO2_1 = 1.57079632679489655800e+00;
static const double PIO2_2 = 6.12323399573676603587e-17;
static const double PIO2_3 = -1.49738490485916983634e-33;

double orc_cr_log(double x) {
    if (!(x > 0.0)) return (x == 0.0) ? -INFINITY : NAN;
    if (isinf(x)) return x;
    int e;
    double m = frexp(x, &e);                  /* x = m 2^e, m in [0.5,1) */
    if (m < 0.70710678118654752440) { m *= 2.0; e -= 1; }
    /* s = (m-1)/(m+1); m-1 is exact (Sterbenz), m+1 kept exact as dd */
    dd_t num = {m - 1.0, 0.0};
    dd_t den = two_sum(m, 1.0);
    dd_t s = dd_div(num, den);
    dd_t t = dd_mul(s, s);
    /* P(t) = sum_{k>=0} t^k / (2k+1), Horner from k = 26 down */
    dd_t p = dd_recip_d(53.0);
    for (int k = 25; k >= 0; --k) {
        p = dd_mul(p, t);
        p = dd_add(p, dd_recip_d((double)(2 * k + 1)));
    }
    dd_t lm = dd_mul(s, p);
    lm.hi *= 2.0; lm.lo *= 2.0;               /* log m = 2 s P(t) */
    dd_t el = dd_mul_d((dd_t){LN2_HI, LN2_LO}, (double)e);
    dd_t r = dd_add(el, lm);
    return r.hi + r.lo;
}

/* sin(r), cos(r) for |r| <= pi/4 + tiny, r given as double-double */
static void dd_sincos_small(dd_t r, dd_t *s, dd_t *c) {
    dd_t t = dd_mul(r, r);
    /* sin r / r = sum (-t)^k / (2k+1)!,  cos r = sum (-t)^k / (2k)! */
    const int K = 15;
    dd_t cs[16], cc[16];
    cs[0] = (dd_t){1.0, 0.0};
    cc[0] = (dd_t){1.0, 0.0};
    for (int k = 1; k <= K; ++k) {
        /* cs[k] = cs[k-1] / ((2k)(2k+1)), cc[k] = cc[k-1] / ((2k-1)(2k)) */
        cs[k] = dd_div(cs[k - 1], (dd_t){(double)((2 * k) * (2 * k + 1)), 0.0});
        cc[k] = dd_div(cc[k - 1], (dd_t){(double)((2 * k - 1) * (2 * k)), 0.0});
    }
    dd_t ps = cs[K], pc = cc[K];
    if (K & 1) { ps.hi = -ps.hi; ps.lo = -ps.lo; pc.hi = -pc.hi; pc.lo = -pc.lo; }
    for (int k = K - 1; k >= 0; --k) {
        dd_t a = cs[k], b = cc[k];
        if (k & 1) { a.hi = -a.hi; a.lo = -a.lo; b.hi = -b.hi; b.lo = -b.lo; }
        ps = dd_add(dd_mul(ps, t), a);
        pc = dd_add(dd_mul(pc, t), b);
    }
    *s = dd_mul(ps, r);
    *c = pc;
}

void orc_cr_sincos(double x, double *s, double *c) {
    if (isnan(x) || isinf(x)) { *s = NAN; *c = NAN; return; }
    double ax = fabs(x);
    if (ax > 1.0e6) {       /* not needed by the path; plain libm */
        *s = sin(x); *c = cos(x); return;
    }
    /* q = nearest integer to x / (pi/2) */
    double qd = nearbyint(x * 0.63661977236758134308);
    dd_t r;
    if (qd == 0.0) {
        r = (dd_t){x, 0.0};
    } else {
        dd_t p1 = two_prod(qd, PIO2_1);
        dd_t p2 = two_prod(qd, PIO2_2);
        double p3 = qd * PIO2_3;
        r = two_sum(x, -p1.hi);
        r = dd_add(r, (dd_t){-p1.lo, 0.0});
        r = dd_add(r, (dd_t){-p2.hi, -p2.lo});
        r = dd_add(r, (dd_t){-p3, 0.0});
    }
    dd_t ss, cc;
    dd_sincos_small(r, &ss, &cc);
    double sv = ss.hi + ss.lo, cv = cc.hi + cc.lo;
    long q = (long)qd;
    switch (((q % 4) + 4) % 4) {
    case 0: *s = sv;  *c = cv;  break;
    case 1: *s = cv;  *c = -sv; break;
    case 2: *s = -sv; *c = -cv; break;
    default: *s = -cv; *c = sv; break;
    }
}
