// NGenIC white noise on the device (random_ngenic.cc:39-154).
//
// Each (i, j) column of the k-space grid is an independent RANLUX stream
// reseeded from the seed table, so one thread generates one column.  The
// RANLUX state (12 words of 48 bits) lives in registers; every 202-step cycle
// is unrolled with static indices and the register names are rotated by 10
// afterwards, so a cycle always starts at word 0 and the 12 outputs of the
// cycle are words 0..11 in order.  The Gaussian transform evaluates log and
// sin/cos in double-double and rounds once (same operation sequence as the
// host seed-table code and the CPU oracle), giving bit-identical draws.
//
// This file is compiled with -ffp-contract=off: every fma is explicit.
#include <cmath>
#include <cstdint>
#include <vector>

#include "fft_core.hpp"
#include "mfic_internal.hpp"
#include "ngenic.hpp"

namespace mfic {

// ----------------------------------------------------- double-double ----
struct dd { double hi, lo; };

__host__ __device__ inline dd two_sum(double a, double b) {
    double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ inline dd fast_two_sum(double a, double b) {
    double s = a + b;
    return {s, b - (s - a)};
}
__host__ __device__ inline dd two_prod(double a, double b) {
    double p = a * b;
    return {p, fma(a, b, -p)};
}
__host__ __device__ inline dd dd_add(dd x, dd y) {
    dd s = two_sum(x.hi, y.hi), t = two_sum(x.lo, y.lo);
    double c = s.lo + t.hi;
    dd v = fast_two_sum(s.hi, c);
    c = t.lo + v.lo;
    return fast_two_sum(v.hi, c);
}
__host__ __device__ inline dd dd_mul(dd x, dd y) {
    dd p = two_prod(x.hi, y.hi);
    double e = fma(x.hi, y.lo, x.lo * y.hi) + p.lo;
    return fast_two_sum(p.hi, e);
}
__host__ __device__ inline dd dd_mul_d(dd x, double d) {
    dd p = two_prod(x.hi, d);
    double e = fma(x.lo, d, p.lo);
    return fast_two_sum(p.hi, e);
}
__host__ __device__ inline dd dd_div(dd x, dd y) {
    double q1 = x.hi / y.hi;
    dd p = dd_mul_d(y, q1);
    dd r = dd_add(x, dd{-p.hi, -p.lo});
    double q2 = r.hi / y.hi;
    p = dd_mul_d(y, q2);
    r = dd_add(r, dd{-p.hi, -p.lo});
    double q3 = r.hi / y.hi;
    dd q = fast_two_sum(q1, q2);
    return dd_add(q, dd{q3, 0.0});
}
__host__ __device__ inline dd dd_recip_d(double d) {
    double hi = 1.0 / d;
    double r = fma(-hi, d, 1.0);
    return {hi, r / d};
}

// series coefficients, computed once on the host with the same operations as
// the CPU oracle: log: 1/(2k+1), k = 0..26;  sin / cos Taylor terms
struct CrTables {
    double logc[27][2];
    double sinc[16][2];
    double cosc[16][2];
};

static CrTables make_tables() {
    CrTables t;
    for (int k = 0; k <= 26; ++k) {
        dd r = dd_recip_d((double)(2 * k + 1));
        t.logc[k][0] = r.hi; t.logc[k][1] = r.lo;
    }
    dd cs = {1.0, 0.0}, cc = {1.0, 0.0};
    t.sinc[0][0] = 1.0; t.sinc[0][1] = 0.0; t.cosc[0][0] = 1.0; t.cosc[0][1] = 0.0;
    for (int k = 1; k <= 15; ++k) {
        cs = dd_div(cs, dd{(double)((2 * k) * (2 * k + 1)), 0.0});
        cc = dd_div(cc, dd{(double)((2 * k - 1) * (2 * k)), 0.0});
        t.sinc[k][0] = cs.hi; t.sinc[k][1] = cs.lo;
        t.cosc[k][0] = cc.hi; t.cosc[k][1] = cc.lo;
    }
    return t;
}

__device__ inline double cr_log(double x, const CrTables& T) {
    // x in (0, 1) here: positive, finite, normal
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) { m *= 2.0; e -= 1; }
    dd num = {m - 1.0, 0.0};
    dd den = two_sum(m, 1.0);
    dd s = dd_div(num, den);
    dd t = dd_mul(s, s);
    dd p = {T.logc[26][0], T.logc[26][1]};
    for (int k = 25; k >= 0; --k) {
        p = dd_mul(p, t);
        p = dd_add(p, dd{T.logc[k][0], T.logc[k][1]});
    }
    dd lm = dd_mul(s, p);
    lm.hi *= 2.0; lm.lo *= 2.0;
    dd el = dd_mul_d(dd{6.93147180559945286227e-01, 2.31904681384629955842e-17}, (double)e);
    dd r = dd_add(el, lm);
    return r.hi + r.lo;
}

__device__ inline void cr_sincos(double x, double* sn, double* cs, const CrTables& T) {
    const double PIO2_1 = 1.57079632679489655800e+00, PIO2_2 = 6.12323399573676603587e-17,
                 PIO2_3 = -1.49738490485916983634e-33;
    double qd = rint(x * 0.63661977236758134308);
    dd r;
    if (qd == 0.0) {
        r = dd{x, 0.0};
    } else {
        dd p1 = two_prod(qd, PIO2_1), p2 = two_prod(qd, PIO2_2);
        double p3 = qd * PIO2_3;
        r = two_sum(x, -p1.hi);
        r = dd_add(r, dd{-p1.lo, 0.0});
        r = dd_add(r, dd{-p2.hi, -p2.lo});
        r = dd_add(r, dd{-p3, 0.0});
    }
    dd t = dd_mul(r, r);
    const int K = 15;
    dd ps = {-T.sinc[K][0], -T.sinc[K][1]}, pc = {-T.cosc[K][0], -T.cosc[K][1]};  // K odd
    for (int k = K - 1; k >= 0; --k) {
        dd a = {T.sinc[k][0], T.sinc[k][1]}, b = {T.cosc[k][0], T.cosc[k][1]};
        if (k & 1) { a.hi = -a.hi; a.lo = -a.lo; b.hi = -b.hi; b.lo = -b.lo; }
        ps = dd_add(dd_mul(ps, t), a);
        pc = dd_add(dd_mul(pc, t), b);
    }
    dd ss = dd_mul(ps, r);
    double sv = ss.hi + ss.lo, cv = pc.hi + pc.lo;
    long q = (long)qd;
    switch (((q % 4) + 4) % 4) {
    case 0: *sn = sv; *cs = cv; break;
    case 1: *sn = cv; *cs = -sv; break;
    case 2: *sn = -sv; *cs = -cv; break;
    default: *sn = -cv; *cs = sv; break;
    }
}

// ------------------------------------------------------------- RANLUX ----
constexpr uint64_t kMask48 = (uint64_t(1) << 48) - 1;

// gsl_rng_set seeding of ranlxd: 31-bit shift register (taps 0 and 18),
// 48 complemented register bits per word.  Host form: the seed is held in a
// C int exactly as there.
inline void ranlux_seed_host(uint64_t x[12], unsigned long seed) {
    if (seed == 0) seed = 1;
    int i = (int)(uint32_t)(seed & 0xFFFFFFFFUL);
    int xbit[31];
    for (int k = 0; k < 31; ++k) { xbit[k] = i % 2; i /= 2; }
    int ibit = 0, jbit = 18;
    for (int k = 0; k < 12; ++k) {
        uint64_t w = 0;
        for (int l = 1; l <= 48; ++l) {
            const int y = (xbit[ibit] + 1) % 2;
            w = 2 * w + (uint64_t)(int64_t)y;
            xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
            ibit = (ibit + 1) % 31;
            jbit = (jbit + 1) % 31;
        }
        x[k] = w & kMask48;
    }
}

// Device form for seeds in [0, 2^31): the register is a 31-bit mask so that
// nothing is indexed dynamically (the state stays in VGPRs).
__device__ inline void ranlux_seed_dev(uint64_t x[12], uint32_t seed) {
    uint32_t bits = (seed == 0 ? 1u : seed) & 0x7fffffffu;
    int ibit = 0, jbit = 18;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        uint64_t w = 0;
        for (int l = 0; l < 48; ++l) {
            const uint32_t bi = (bits >> ibit) & 1u, bj = (bits >> jbit) & 1u;
            w = 2 * w + (uint64_t)(1u - bi);
            bits = (bits & ~(1u << ibit)) | ((bi ^ bj) << ibit);
            ibit = ibit == 30 ? 0 : ibit + 1;
            jbit = jbit == 30 ? 0 : jbit + 1;
        }
        x[k] = w;
    }
}

// x_ir <- x_jr - x_ir - c  (mod 2^48), jr = ir + 7 (mod 12)
#define MFIC_SWB(IR)                                                          \
    {                                                                         \
        uint64_t d_ = x[((IR) + 7) % 12] + (uint64_t(1) << 48) - x[(IR)] - c; \
        c = 1 - (int)(d_ >> 48);                                              \
        x[(IR)] = d_ & kMask48;                                               \
    }

// one luxury-202 cycle starting at word 0; afterwards the words are renamed
// so that the 12 outputs (the last 12 produced, oldest first) are x[0..11]
// and the next cycle again starts at word 0
__host__ __device__ inline void ranlux_cycle202(uint64_t x[12], int& c) {
    for (int blk = 0; blk < 16; ++blk) {
#pragma unroll
        for (int r = 0; r < 12; ++r) MFIC_SWB(r)
    }
#pragma unroll
    for (int r = 0; r < 10; ++r) MFIC_SWB(r)
    const uint64_t a = x[10], b = x[11];
#pragma unroll
    for (int r = 11; r >= 2; --r) x[r] = x[r - 2];
    x[0] = a;
    x[1] = b;
}
#undef MFIC_SWB

__device__ inline uint64_t sel12(const uint64_t x[12], int c) {
    uint64_t v = x[0];
#pragma unroll
    for (int r = 1; r < 12; ++r) v = (c == r) ? x[r] : v;
    return v;
}

// FP32 output fast path: y approximates the correctly rounded double result
// within a few double ulps.  When every double within 64 ulps of y rounds
// to the same float, that float is also the rounding of the correctly
// rounded result, so the FP32 noise stays bit-identical.
__device__ __forceinline__ bool float_rounding_settled(double y) {
    const double d = fabs(y) * 0x1p-46;
    const float f = (float)y;
    return (float)(y - d) == f && (float)(y + d) == f;
}

// one value of the white-noise field from a (phase, amplitude) draw pair
template <typename T>
__device__ __forceinline__ cx<T> ngenic_value(double phase, double ampl, const CrTables& tabs) {
    double zr = 0.0, zi = 0.0;
    bool done = false;
    if constexpr (sizeof(T) == 4) {
        const double a = sqrt(-log(ampl));
        double sn, cs;
        sincos(phase, &sn, &cs);
        zr = a * cs;
        zi = a * sn;
        done = float_rounding_settled(zr) && float_rounding_settled(zi);
    }
    if (!done) {
        const double a = sqrt(-cr_log(ampl, tabs));
        double sn, cs;
        cr_sincos(phase, &sn, &cs, tabs);
        zr = a * cs;
        zi = a * sn;
    }
    return cx<T>{(T)zr, (T)zi};
}

// One thread per (i, j) column draws its kz = 0 .. n/2 values in stream
// order.  kz = 0 values (and their Hermitian mirrors) are stored directly;
// kz >= 1 values go through an LDS tile of NGENIC_KT draws per column and
// leave the workgroup as runs of NGENIC_KT consecutive elements per column,
// so every cache line is written whole.
constexpr int NGENIC_BS = 64, NGENIC_KT = 16;
// PHI: store phi1 = sg * w / volfac * F(|n|^2) instead of the noise w (the
// values k_phi1_tab forms from the stored noise, from the same T-rounded w),
// with F the FP64 -A(k)/k^2 table per integer |n|^2 of a cubic box and w
// normalised to |w| = 1 when fixing
struct Phi1Fold {
    const double* tab = nullptr;
    double sg = 1.0, volfac = 1.0;
    int fixing = 0;
};
template <typename T>
__device__ __forceinline__ cx<T> phi1_fold(cx<T> w, int si, int sj, int k, const Phi1Fold& f) {
    double wr = w.x, wi = w.y;
    if (f.fixing) {
        const double a = hypot(wr, wi);
        if (a != 0.0) { wr /= a; wi /= a; }
    }
    const double F = f.tab[si * si + sj * sj + k * k];
    return cx<T>{(T)(f.sg * wr / f.volfac * F), (T)(f.sg * wi / f.volfac * F)};
}

template <typename T, bool PHI = false>
__global__ void __launch_bounds__(NGENIC_BS)
k_ngenic_fill(cx<T>* __restrict__ g, const uint32_t* __restrict__ table, int n, CrTables tabs, Phi1Fold pf) {
    __shared__ cx<T> tile[NGENIC_BS][NGENIC_KT + 1];
    const long cols = (long)n * n;
    const long col0 = (long)blockIdx.x * NGENIC_BS;
    const long col = col0 + threadIdx.x;
    const bool active = col < cols;
    const int i = active ? (int)(col / n) : 0, j = active ? (int)(col % n) : 0;
    const int nc = n / 2 + 1, nh = n / 2;
    const int ii = i > 0 ? n - i : 0, jj = j > 0 ? n - j : 0;
    const bool nyq_col = i == nh || j == nh;
    uint64_t x[12];
    int carry = 0, pos = 12;  // pos == 12: the next draw runs a cycle
    if (active) ranlux_seed_dev(x, table[(long)i * n + j]);
    auto draw = [&](double& phase, double& ampl) {
        if (pos == 12) { ranlux_cycle202(x, carry); pos = 0; }
        phase = (double)sel12(x, pos++) * (1.0 / 281474976710656.0) * 2 * 3.14159265358979323846;
        do {
            if (pos == 12) { ranlux_cycle202(x, carry); pos = 0; }
            ampl = (double)sel12(x, pos++) * (1.0 / 281474976710656.0);
        } while (ampl == 0 || ampl == 1);
    };
    if (active) {  // kz = 0: the draw fills (i, j) and its mirror (-i, -j)
        double phase, ampl;
        draw(phase, ampl);
        if (!nyq_col && !(i == 0 && j == 0) && (i == 0 ? j < nh : i < nh)) {
            cx<T> z = ngenic_value<T>(phase, ampl, tabs);
            // the mirror (-i, -j) has the same |n|^2: its value is conj(z) either way
            if constexpr (PHI) z = phi1_fold(z, i - (i > nh) * n, j - (j > nh) * n, 0, pf);
            g[((long)i * n + j) * nc] = z;
            g[((long)ii * n + jj) * nc] = cx<T>{z.x, -z.y};
        }
    }
    for (int k0 = 1; k0 < nc; k0 += NGENIC_KT) {
        if (active) {
            for (int kk = 0; kk < NGENIC_KT; ++kk) {
                const int k = k0 + kk;
                if (k >= nc) break;
                double phase, ampl;
                draw(phase, ampl);
                cx<T> z = (nyq_col || k == nh) ? cx<T>{0, 0} : ngenic_value<T>(phase, ampl, tabs);
                if constexpr (PHI) {
                    if (!(nyq_col || k == nh)) z = phi1_fold(z, i - (i > nh) * n, j - (j > nh) * n, k, pf);
                }
                tile[threadIdx.x][kk] = z;
            }
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < NGENIC_BS * NGENIC_KT; idx += NGENIC_BS) {
            const int c = idx / NGENIC_KT, kk = idx % NGENIC_KT;
            const long cc = col0 + c;
            const int k = k0 + kk;
            if (cc < cols && k < nc) g[cc * nc + k] = tile[c][kk];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- host -----

namespace {
struct HostRanlux {
    uint64_t x[12];
    int carry = 0, pos = 12;
    explicit HostRanlux(unsigned long seed) { ranlux_seed_host(x, seed); }
    double uniform() {
        if (pos == 12) { ranlux_cycle202(x, carry); pos = 0; }
        return (double)x[pos++] * (1.0 / 281474976710656.0);
    }
};
}  // namespace

void ngenic_seedtable(long seed, size_t n, uint32_t* T) {
    HostRanlux r((unsigned long)seed);
    auto draw = [&r]() { return (uint32_t)(2147483647.0 * r.uniform()); };
    for (size_t q = 0; q < n * n; ++q) T[q] = 0;
    for (size_t i = 0; i < n / 2; ++i) {
        for (size_t j = 0; j < i; j++) T[i * n + j] = draw();
        for (size_t j = 0; j < i + 1; j++) T[j * n + i] = draw();
        for (size_t j = 0; j < i; j++) T[(n - 1 - i) * n + j] = draw();
        for (size_t j = 0; j < i + 1; j++) T[(n - 1 - j) * n + i] = draw();
        for (size_t j = 0; j < i; j++) T[i * n + (n - 1 - j)] = draw();
        for (size_t j = 0; j < i + 1; j++) T[j * n + (n - 1 - i)] = draw();
        for (size_t j = 0; j < i; j++) T[(n - 1 - i) * n + (n - 1 - j)] = draw();
        for (size_t j = 0; j < i + 1; j++) T[(n - 1 - j) * n + (n - 1 - i)] = draw();
    }
}

void ngenic_fill(mfic_grid g, const uint32_t* dev_table) {
    MFIC_CHECK(g->n[0] == g->n[1] && g->n[1] == g->n[2], MFIC_ERR_SIZE, "NGenIC needs a cubic grid");
    MFIC_CHECK(g->n[0] % 2 == 0, MFIC_ERR_SIZE, "NGenIC needs an even grid size");
    static const CrTables tabs = make_tables();
    const int n = (int)g->n[0];
    const long cols = (long)n * n;
    const int bs = NGENIC_BS;
    ProfScope prof(g->ctx, "ngenic_fill", (double)cols * (n / 2 + 1) * (g->prec == MFIC_F64 ? 16 : 8));
    if (g->prec == MFIC_F64)
        hipLaunchKernelGGL(k_ngenic_fill<double>, dim3((unsigned)((cols + bs - 1) / bs)), dim3(bs), 0, g->ctx->stream,
                           (cx<double>*)g->d, dev_table, n, tabs, Phi1Fold{});
    else
        hipLaunchKernelGGL(k_ngenic_fill<float>, dim3((unsigned)((cols + bs - 1) / bs)), dim3(bs), 0, g->ctx->stream,
                           (cx<float>*)g->d, dev_table, n, tabs, Phi1Fold{});
    MFIC_LAUNCH_CHECK();
}

// zero the kz = 0 plane of a complex [n][n][n/2+1] grid (the only entries
// the noise kernel does not write to every column of)
template <typename T> __global__ void k_zero_kz0(cx<T>* __restrict__ g, long cols, int nc) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cols) g[c * nc] = cx<T>{0, 0};
}

void ngenic_phi1_fill(mfic_grid phi, const uint32_t* dev_table, const double* phitab, double volfac, int fixing,
                      int inversion) {
    MFIC_CHECK(phi->n[0] == phi->n[1] && phi->n[1] == phi->n[2], MFIC_ERR_SIZE, "NGenIC needs a cubic grid");
    MFIC_CHECK(phi->n[0] % 2 == 0, MFIC_ERR_SIZE, "NGenIC needs an even grid size");
    static const CrTables tabs = make_tables();
    const int n = (int)phi->n[0], nc = n / 2 + 1;
    const long cols = (long)n * n;
    const int bs = NGENIC_BS;
    Phi1Fold pf;
    pf.tab = phitab;
    pf.sg = inversion ? -1.0 : 1.0;
    pf.volfac = volfac;
    pf.fixing = fixing;
    ProfScope prof(phi->ctx, "ngenic_phi1", (double)cols * nc * (phi->prec == MFIC_F64 ? 16 : 8));
    const dim3 zgrid((unsigned)((cols + 255) / 256)), grid((unsigned)((cols + bs - 1) / bs));
    if (phi->prec == MFIC_F64) {
        hipLaunchKernelGGL(k_zero_kz0<double>, zgrid, dim3(256), 0, phi->ctx->stream, (cx<double>*)phi->d, cols, nc);
        hipLaunchKernelGGL((k_ngenic_fill<double, true>), grid, dim3(bs), 0, phi->ctx->stream, (cx<double>*)phi->d,
                           dev_table, n, tabs, pf);
    } else {
        hipLaunchKernelGGL(k_zero_kz0<float>, zgrid, dim3(256), 0, phi->ctx->stream, (cx<float>*)phi->d, cols, nc);
        hipLaunchKernelGGL((k_ngenic_fill<float, true>), grid, dim3(bs), 0, phi->ctx->stream, (cx<float>*)phi->d,
                           dev_table, n, tabs, pf);
    }
    MFIC_LAUNCH_CHECK();
}

}  // namespace mfic
