/*
 * Oracle: RANLUX (ranlxd) and the NGenIC white-noise generator
 * (TEST INFRASTRUCTURE ONLY).
 *
 * RANLUX, double-precision variant (M. Lüscher, Comput. Phys. Commun. 79
 * (1994) 100; ranlxd v1 1997), as exposed by GSL as gsl_rng_ranlxd1/2:
 *   state: 12 words of 48 bits, a borrow bit;
 *   step:  x_n = x_{n-5} - x_{n-12} - c_{n-1}  (mod 2^48, c = borrow);
 *   a cycle runs `luxury` (202 / 397) steps and the last 12 words produced
 *   are returned in the order they were produced, scaled by 2^-48;
 *   seeding (gsl_rng_set): a 31-bit shift register initialised from the
 *   seed bits (taps 0 and 18) fills each word with 48 complemented bits.
 * GSL keeps the words as doubles k*2^-48; every operation on them is exact,
 * so 48-bit integers give identical values.
 *
 * NGenIC: random_ngenic.cc:39-75 (seed table) and :84-154 (Fill_Grid,
 * non-distributed branch).
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

#define MASK48 ((((uint64_t)1) << 48) - 1)

void orc_ranlxd_set(orc_ranlxd *st, unsigned long seed, int luxury) {
    if (seed == 0) seed = 1;
    /* the seed is held in a C int (two's complement, truncating division) */
    int i = (int)(uint32_t)(seed & 0xFFFFFFFFUL);
    int xbit[31];
    for (int k = 0; k < 31; ++k) {
        xbit[k] = i % 2;
        i /= 2;
    }
    int ibit = 0, jbit = 18;
    for (int k = 0; k < 12; ++k) {
        uint64_t x = 0;
        for (int l = 1; l <= 48; ++l) {
            int y = (xbit[ibit] + 1) % 2;
            x = 2 * x + (uint64_t)(int64_t)y;
            xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
            ibit = (ibit + 1) % 31;
            jbit = (jbit + 1) % 31;
        }
        st->x[k] = x & MASK48;
    }
    st->carry = 0;
    st->ir = 11;
    st->jr = 7;
    st->ir_old = 0;
    st->pr = luxury;
}

/* one cycle of `pr` subtract-with-borrow steps starting at position ir */
static void ranlxd_cycle(orc_ranlxd *st) {
    int ir = st->ir, jr = st->jr, c = st->carry;
    for (int k = 0; k < st->pr; ++k) {
        int64_t d = (int64_t)st->x[jr] - (int64_t)st->x[ir] - c;
        if (d < 0) { d += ((int64_t)1 << 48); c = 1; } else { c = 0; }
        st->x[ir] = (uint64_t)d;
        ir = (ir + 1) % 12;
        jr = (jr + 1) % 12;
    }
    st->ir = ir;
    st->ir_old = ir;
    st->jr = jr;
    st->carry = c;
}

uint64_t orc_ranlxd_next48(orc_ranlxd *st) {
    st->ir = (st->ir + 1) % 12;
    if (st->ir == st->ir_old) ranlxd_cycle(st);
    return st->x[st->ir];
}

double orc_ranlxd_uniform(orc_ranlxd *st) {
    return (double)orc_ranlxd_next48(st) * (1.0 / 281474976710656.0);
}

unsigned long orc_ranlxd_get(orc_ranlxd *st) {
    /* gsl_rng_get: uniform * 2^32 truncated */
    return (unsigned long)(orc_ranlxd_uniform(st) * 4294967296.0);
}

/* ----------------------------------------------------------- NGenIC --- */

static uint32_t seed_draw(orc_ranlxd *r) {
    return (uint32_t)(2147483647.0 * orc_ranlxd_uniform(r));
}

void orc_ngenic_seedtable(long seed, size_t n, uint32_t *T) {
    orc_ranlxd r;
    orc_ranlxd_set(&r, (unsigned long)seed, 202);
    memset(T, 0, n * n * sizeof(uint32_t));
    /* NGenIC shell order: eight edges of the growing square per step i */
    for (size_t i = 0; i < n / 2; ++i) {
        for (size_t j = 0; j < i; j++)      T[i * n + j] = seed_draw(&r);
        for (size_t j = 0; j < i + 1; j++)  T[j * n + i] = seed_draw(&r);
        for (size_t j = 0; j < i; j++)      T[(n - 1 - i) * n + j] = seed_draw(&r);
        for (size_t j = 0; j < i + 1; j++)  T[(n - 1 - j) * n + i] = seed_draw(&r);
        for (size_t j = 0; j < i; j++)      T[i * n + (n - 1 - j)] = seed_draw(&r);
        for (size_t j = 0; j < i + 1; j++)  T[j * n + (n - 1 - i)] = seed_draw(&r);
        for (size_t j = 0; j < i; j++)      T[(n - 1 - i) * n + (n - 1 - j)] = seed_draw(&r);
        for (size_t j = 0; j < i + 1; j++)  T[(n - 1 - j) * n + (n - 1 - i)] = seed_draw(&r);
    }
}

void orc_ngenic_fill(const uint32_t *T, size_t n, double *kd) {
    const size_t nc = n / 2 + 1;
    memset(kd, 0, n * n * (n + 2) * sizeof(double));
    /* every (i,j) column is an independent stream; columns write disjoint
     * elements (k>0: their own; k=0: the Hermitian partner pairs) */
#pragma omp parallel for schedule(dynamic, 4)
    for (long il = 0; il < (long)n; ++il) {
        size_t i = (size_t)il;
        size_t ii = (i > 0) ? n - i : 0;
        orc_ranlxd r;
        for (size_t j = 0; j < n; ++j) {
            size_t jj = (j > 0) ? n - j : 0;
            orc_ranlxd_set(&r, T[i * n + j], 202);
            for (size_t k = 0; k < nc; ++k) {
                double phase = orc_ranlxd_uniform(&r) * 2 * 3.14159265358979323846;
                double ampl;
                do {
                    ampl = orc_ranlxd_uniform(&r);
                } while (ampl == 0 || ampl == 1);
                if (i == n / 2 || j == n / 2 || k == n / 2) continue;
                if (i == 0 && j == 0 && k == 0) continue;
                ampl = sqrt(-orc_cr_log(ampl));
                double sn, cs;
                orc_cr_sincos(phase, &sn, &cs);
                double zr = ampl * cs, zi = ampl * sn;
                double *e;
                if (k > 0) {
                    e = kd + 2 * ((i * n + j) * nc + k);
                    e[0] = zr; e[1] = zi;
                } else if (i == 0) {
                    if (j < n / 2) {
                        e = kd + 2 * ((i * n + j) * nc);
                        e[0] = zr; e[1] = zi;
                        e = kd + 2 * ((i * n + jj) * nc);
                        e[0] = zr; e[1] = -zi;
                    }
                } else if (i < n / 2) {
                    e = kd + 2 * ((i * n + j) * nc);
                    e[0] = zr; e[1] = zi;
                    e = kd + 2 * ((ii * n + jj) * nc);
                    e[0] = zr; e[1] = -zi;
                }
            }
        }
    }
}

/* The reference's Gaussian transform evaluates std::log, std::cos and
 * std::sin of the host libm (random_ngenic.cc:119-120); this library uses
 * correctly rounded versions so that host and device agree bit for bit.
 * Walks every draw of the non-distributed fill (the same column streams and
 * skips as orc_ngenic_fill) and counts where the two differ:
 *   out[0] draws stored, out[1] log differs, out[2] cos differs,
 *   out[3] sin differs, out[4] the stored value (re or im) differs.
 * The libm calls go through volatile pointers so the compiler neither
 * folds nor fuses them. */
void orc_ngenic_libm_compare(const uint32_t *T, size_t n, uint64_t out[5]) {
    double (*volatile p_log)(double) = log;
    double (*volatile p_cos)(double) = cos;
    double (*volatile p_sin)(double) = sin;
    const size_t nc = n / 2 + 1;
    uint64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : c0, c1, c2, c3, c4)
    for (long il = 0; il < (long)n; ++il) {
        size_t i = (size_t)il;
        orc_ranlxd r;
        for (size_t j = 0; j < n; ++j) {
            orc_ranlxd_set(&r, T[i * n + j], 202);
            for (size_t k = 0; k < nc; ++k) {
                double phase = orc_ranlxd_uniform(&r) * 2 * 3.14159265358979323846;
                double ampl;
                do {
                    ampl = orc_ranlxd_uniform(&r);
                } while (ampl == 0 || ampl == 1);
                if (i == n / 2 || j == n / 2 || k == n / 2) continue;
                if (i == 0 && j == 0 && k == 0) continue;
                /* only draws the fill stores (k > 0, or half of the k = 0 plane) */
                if (k == 0 && !((i == 0 && j < n / 2) || (i > 0 && i < n / 2))) continue;
                const double lc = orc_cr_log(ampl), lg = p_log(ampl);
                double sn, cs;
                orc_cr_sincos(phase, &sn, &cs);
                const double cg = p_cos(phase), sg = p_sin(phase);
                const double a1 = sqrt(-lc), a2 = sqrt(-lg);
                c0 += 1;
                c1 += lc != lg;
                c2 += cs != cg;
                c3 += sn != sg;
                c4 += (a1 * cs != a2 * cg) || (a1 * sn != a2 * sg);
            }
        }
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3; out[4] = c4;
}
