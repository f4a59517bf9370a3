// Pruned and fused 3/2-padded transforms for the Orszag products
// (convolution.hh:495-651) on gfx950.
//
// The padded spectrum of an N-grid operand is nonzero only for source rows
// (i, j in [0, N/2] U (N, 3N/2)) and kz <= N/2, and the truncation keeps
// exactly those modes again.  So:
//   c2r  A: axis-0 pass over the (j, kz <= N/2) source columns, reading the
//           N-grid operand (Hessian / gradient / identity evaluated on the fly,
//           x (3/2)^{3/2}) and writing the full padded axis;
//        B: axis-1 pass over (i', kz <= N/2), the zero band of j' is not read;
//        C: contiguous-axis c2r rows, kz > N/2 implicitly zero, real rows out.
//   r2c  C': contiguous-axis r2c rows whose input is a sum of products of
//           padded real fields (FP64 accumulation), only kz <= N/2 written;
//        B': axis-1 pass, only the kept j' written;
//        A': axis-0 pass, only the kept i' written straight into the N-grid
//           with the unpad rules (Nyquist planes zeroed, / (3/2)^{3/2}), the
//           output operator and optionally -1/k^2.
// Padded k-space intermediates use a compact [Np][Np][N/2+1] layout; padded
// real fields use [Np][Np][Np].  Every skipped element is an exact zero of
// the unpruned algorithm, and every dropped output is one the unpad
// discards, so results equal the full transforms up to rounding order.
#include <algorithm>
#include <cmath>

#include "fft_core.hpp"
#include "mfic_internal.hpp"
#include "pruned.hpp"

namespace mfic {

// ---------------------------------------------------------- loaders / stores
//
// Every padded pass runs pencils of length L = 3N/2 and every loader and
// store below is instantiated with that L, so N, N/2 + 1 and the strides are
// compile-time constants.  Element e of a pencil sits on lane q at position
// pos = q + P e; once the element loop is unrolled, which side of the zero
// band it lies on is a constant for all lanes except the two elements that
// straddle N/2 and N, so band tests and gradient factors cost nothing for the
// others.

template <int L, int V = 0> struct PadGeo {
    static constexpr int P = Plan<L, V>::P, N = 2 * L / 3, NH = N / 2, KC = NH + 1;
    static_assert(3 * N == 2 * L, "padded length must be 3N/2");
    // 0: every lane below N/2, 1: every lane above N, 2: every lane in the
    // zero band N/2 < pos <= N, 3: mixed
    static __device__ __forceinline__ constexpr int side(int e) {
        return P * e + P - 1 < NH ? 0 : (P * e > N ? 1 : (P * e > NH && P * e + P - 1 <= N ? 2 : 3));
    }
    static __device__ __forceinline__ bool band(int q, int e) {
        const int pos = q + P * e;
        const int sd = side(e);
        return sd == 2 || (sd == 3 && pos > NH && pos <= N);
    }
    // source index behind pos (convolution.hh:562-571), -1 in the band
    static __device__ __forceinline__ int src(int q, int e) {
        const int pos = q + P * e;
        switch (side(e)) {
        case 0: return pos;
        case 1: return pos - NH;
        case 2: return -1;
        default: return pos <= NH ? pos : (pos > N ? pos - NH : -1);
        }
    }
    // gradient factor of the source mode behind pos: signed index times kf,
    // 0 at the source Nyquist N/2 (never used in the band)
    template <typename T> static __device__ __forceinline__ T grad(int q, int e, T kf) {
        const int pos = q + P * e;
        switch (side(e)) {
        case 0: return (T)q * kf + (T)(P * e) * kf;
        case 1: return (T)q * kf + (T)(P * e - L) * kf;
        default: return pos == NH ? T(0) : (T)(pos - (pos > NH ? L : 0)) * kf;
        }
    }
};

// gradient factor of an unpadded index
template <typename T> __device__ __forceinline__ T grad_t(int idx, int n, T kf) {
    return idx == n / 2 ? T(0) : (T)(idx - (idx > n / 2) * n) * kf;
}

// pass A: operand of an N-grid k-space field.  Along a pencil only the
// axis-0 gradient g_i varies, so every operand multiplier is a polynomial
// s(g_i) = (alpha g_i + beta) g_i + gamma fixed per column (imaginary for a
// single gradient); the (3/2)^{3/2} pad factor is folded in.
template <typename T> struct LdOperand {
    const cx<T>* f;
    int n, kc, kind, d0, d1, d2, d3;
    double kf, scale;
    struct Col { const cx<T>* p; double al, be, ga; bool imag; };
    // product -(g_a g_b) as polynomial coefficients in g_i
    __device__ __forceinline__ static void prod2(int a, int b, double gj, double gk, double sg, double& al, double& be,
                                                 double& ga) {
        const double ga_ = a == 1 ? gj : gk, gb_ = b == 1 ? gj : gk;
        if (a == 0 && b == 0) al += -sg;
        else if (a == 0) be += -sg * gb_;
        else if (b == 0) be += -sg * ga_;
        else ga += -sg * (ga_ * gb_);
    }
    template <int L> __device__ __forceinline__ Col column(long col) const {
        using G = PadGeo<L>;
        Col c;
        const int j = (int)(col / G::KC), k = (int)(col % G::KC);
        c.p = f + col;
        const double gj = grad_t<double>(j, G::N, kf), gk = grad_t<double>(k, G::N, kf);
        c.al = c.be = c.ga = 0.0;
        c.imag = false;
        switch (kind) {
        case MFIC_OPND_HESS: prod2(d0, d1, gj, gk, 1.0, c.al, c.be, c.ga); break;
        case MFIC_OPND_SUM_HESS:
            prod2(d0, d1, gj, gk, 1.0, c.al, c.be, c.ga);
            prod2(d2, d3, gj, gk, 1.0, c.al, c.be, c.ga);
            break;
        case MFIC_OPND_DIFF_HESS:
            prod2(d0, d1, gj, gk, 1.0, c.al, c.be, c.ga);
            prod2(d2, d3, gj, gk, -1.0, c.al, c.be, c.ga);
            break;
        case MFIC_OPND_GRAD:
            c.imag = true;
            if (d0 == 0) c.be = 1.0; else c.ga = d0 == 1 ? gj : gk;
            break;
        default: c.ga = 1.0; break;
        }
        c.al *= scale;
        c.be *= scale;
        c.ga *= scale;
        return c;
    }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> load(const Col& c, int q, int e) const {
        using G = PadGeo<L, V>;
        const int i = G::src(q, e);
        return i < 0 ? cx<T>{0, 0} : c.p[(long)i * (G::N * G::KC)];
    }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> fix(const Col& c, int q, int e, cx<T> v) const {
        using G = PadGeo<L, V>;
        if (G::band(q, e)) return v;
        const double g = G::template grad<double>(q, e, kf);
        const double s = (c.al * g + c.be) * g + c.ga;
        return c.imag ? cx<T>{(T)(-s * (double)v.y), (T)(s * (double)v.x)}
                      : cx<T>{(T)(s * (double)v.x), (T)(s * (double)v.y)};
    }
};
// pass A of the factorised Hessians: the N-grid field times g_i^p (p = 0..2,
// the number of x derivatives) and the pad factor
template <typename T> struct LdPow {
    const cx<T>* f;
    int n, kc, p;
    T kf, scale;
    struct Col { const cx<T>* p; };
    template <int L> __device__ __forceinline__ Col column(long col) const { return Col{f + col}; }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> load(const Col& c, int q, int e) const {
        using G = PadGeo<L, V>;
        const int i = G::src(q, e);
        return i < 0 ? cx<T>{0, 0} : c.p[(long)i * (G::N * G::KC)];
    }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> fix(const Col&, int q, int e, cx<T> v) const {
        const T g = PadGeo<L, V>::template grad<T>(q, e, kf);
        const T m = scale * (p > 0 ? g : T(1)) * (p > 1 ? g : T(1));
        return cx<T>{m * v.x, m * v.y};
    }
};
// pass B of the factorised Hessians: g_j^q along the pencil (zero band
// skipped) times the per-column factor sign * g_k^r
template <typename T> struct LdAx1Mul {
    const cx<T>* kb;
    int n, np, kc, q, r;
    T kf, sign;
    struct Col { const cx<T>* p; T m; };
    template <int L> __device__ __forceinline__ Col column(long col) const {
        using G = PadGeo<L>;
        const long ip = col / G::KC;
        const int k = (int)(col % G::KC);
        const T gk = grad_t<T>(k, G::N, kf);
        return Col{kb + ip * (L * G::KC) + k, r == 0 ? sign : (r == 1 ? sign * gk : sign * (gk * gk))};
    }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> load(const Col& c, int ln, int e) const {
        using G = PadGeo<L, V>;
        return G::band(ln, e) ? cx<T>{0, 0} : c.p[(long)(ln + G::P * e) * G::KC];
    }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> fix(const Col& c, int ln, int e, cx<T> v) const {
        const T g = PadGeo<L, V>::template grad<T>(ln, e, kf);
        const T m = c.m * (q > 0 ? g : T(1)) * (q > 1 ? g : T(1));
        return cx<T>{m * v.x, m * v.y};
    }
};
// pass A store: padded column (j' = pad(j)) of the compact k buffer
template <typename T> struct StPadCol {
    cx<T>* kb;
    int n, np, kc;
    struct Col { cx<T>* p; };
    template <int L> __device__ __forceinline__ Col column(long col) const {
        using G = PadGeo<L>;
        const int j = (int)(col / G::KC), k = (int)(col % G::KC);
        return Col{kb + (long)(j > G::NH ? j + G::NH : j) * G::KC + k};
    }
    template <int L, int V = 0> __device__ __forceinline__ void store(const Col& c, int q, int e, cx<T> v) const {
        using G = PadGeo<L, V>;
        c.p[(long)(q + G::P * e) * (L * G::KC)] = v;
    }
};
// passes B / B': axis-1 pencils of the compact buffer, column = (i', kz)
template <typename T, bool SKIP_BAND> struct LdAx1 {
    const cx<T>* kb;
    struct Col { const cx<T>* p; };
    template <int L> __device__ __forceinline__ Col column(long col) const {
        using G = PadGeo<L>;
        const long ip = col / G::KC, k = col % G::KC;
        return Col{kb + ip * (L * G::KC) + k};
    }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> load(const Col& c, int q, int e) const {
        using G = PadGeo<L, V>;
        if (SKIP_BAND && G::band(q, e)) return cx<T>{0, 0};
        return c.p[(long)(q + G::P * e) * G::KC];
    }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> fix(const Col&, int, int, cx<T> v) const { return v; }
};
template <typename T, bool DROP_BAND> struct StAx1 {
    cx<T>* kb;
    struct Col { cx<T>* p; };
    template <int L> __device__ __forceinline__ Col column(long col) const {
        using G = PadGeo<L>;
        const long ip = col / G::KC, k = col % G::KC;
        return Col{kb + ip * (L * G::KC) + k};
    }
    template <int L, int V = 0> __device__ __forceinline__ void store(const Col& c, int q, int e, cx<T> v) const {
        using G = PadGeo<L, V>;
        if (DROP_BAND && G::band(q, e)) return;
        c.p[(long)(q + G::P * e) * G::KC] = v;
    }
};
// pass A' load: padded column of the compact buffer for kept j
template <typename T> struct LdPadCol {
    const cx<T>* kb;
    int n, np, kc;
    struct Col { const cx<T>* p; };
    template <int L> __device__ __forceinline__ Col column(long col) const {
        using G = PadGeo<L>;
        const int j = (int)(col / G::KC), k = (int)(col % G::KC);
        return Col{kb + (long)(j > G::NH ? j + G::NH : j) * G::KC + k};
    }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> load(const Col& c, int q, int e) const {
        using G = PadGeo<L, V>;
        return c.p[(long)(q + G::P * e) * (L * G::KC)];
    }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> fix(const Col&, int, int, cx<T> v) const { return v; }
};
// pass A' store: unpad rules + output operator (+ optional -1/k^2) into the N-grid
template <typename T> struct StUnpad {
    cx<T>* out;
    int n, kc, op, invlap;
    double scale, coef, kf;
    struct Col { cx<T>* p; double kj, kk; bool nyq; };
    template <int L> __device__ __forceinline__ Col column(long col) const {
        using G = PadGeo<L>;
        Col c;
        const int j = (int)(col / G::KC), k = (int)(col % G::KC);
        c.p = out + col;
        c.kj = ((double)j - (double)(j > G::NH) * G::N) * kf;
        c.kk = ((double)k - (double)(k > G::NH) * G::N) * kf;
        c.nyq = (j == G::NH) || (k == G::NH);
        return c;
    }
    template <int L, int V = 0> __device__ __forceinline__ void store(const Col& c, int q, int e, cx<T> v) const {
        using G = PadGeo<L, V>;
        const int i = G::src(q, e);
        if (i < 0) return;
        double re = 0.0, im = 0.0;
        if (!(c.nyq || i == G::NH)) {
            double f = scale;
            if (invlap) {  // one FP64 division per element, shared by both parts
                const double ki = G::template grad<double>(q, e, kf);
                const double k2 = ki * ki + c.kj * c.kj + c.kk * c.kk;
                f = k2 > 0.0 ? -scale / k2 : 0.0;
            }
            re = (double)v.x * f;
            im = (double)v.y * f;
        }
        cx<T>* dst = c.p + (long)i * (G::N * G::KC);
        if (op == MFIC_OP_ASSIGN) { *dst = cx<T>{(T)re, (T)im}; return; }
        const cx<T> o = *dst;
        if (op == MFIC_OP_ADD) *dst = cx<T>{(T)((double)o.x + re), (T)((double)o.y + im)};
        else if (op == MFIC_OP_SUB) *dst = cx<T>{(T)((double)o.x - re), (T)((double)o.y - im)};
        else *dst = cx<T>{(T)((double)o.x + re * coef), (T)((double)o.y + im * coef)};
    }
};
// ------------------------------------------------------------- kernels ----

template <typename T, int L, int TK, bool FWD, class LD, class ST, int V = 0>
__global__ void __launch_bounds__((Plan<L, V>::P * TK))
k_pencils_f(LD ld, ST st, long ncols, const cx<T>* __restrict__ tw, int tkp, int xcd) {
    constexpr int E = Plan<L, V>::E;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int t = threadIdx.x % TK, q = threadIdx.x / TK;
    const long col = (xcd ? xcd_tile(blockIdx.x, gridDim.x) : (long)blockIdx.x) * TK + t;
    const bool active = col < ncols;
    const long c = active ? col : 0;
    const auto lc = ld.template column<L>(c);
    cx<T> v[E];
    // every load of the pencil is issued before any multiplier is applied;
    // lanes past the last column read column 0 and store nothing, so no load
    // carries a branch
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = ld.template load<L, V>(lc, q, e);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = ld.template fix<L, V>(lc, q, e, v[e]);
    fft_pencil<T, L, FWD, true, V>(v, q, lds, LdsMap{t, tkp}, tw);
    if (active) {
        const auto sc = st.template column<L>(c);
#pragma unroll
        for (int e = 0; e < E; ++e) st.template store<L, V>(sc, q, e, v[e]);
    }
}

// Orszag truncation followed by re-padding along one axis of a compact
// padded buffer that is real-space along that axis: forward transform, keep
// the source modes (not the zero band, not the source Nyquist N/2), inverse
// transform.  Equal to the unpad (convolution.hh:600-651) then pad_insert
// (convolution.hh:552-594) pair of a two-stage product restricted to that
// axis; the column mask zeroes kz = N/2 and `scale` carries the pad factors.
template <typename T, int L, int TK, class LD, class ST, int V = 0>
__global__ void __launch_bounds__((Plan<L, V>::P * TK))
k_pencils_filter(LD ld, ST st, long ncols, const cx<T>* __restrict__ tw, int tkp, int xcd, int zero_kz_nyquist,
                 T scale) {
    using G = PadGeo<L, V>;
    constexpr int E = Plan<L, V>::E;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int t = threadIdx.x % TK, q = threadIdx.x / TK;
    const long col = (xcd ? xcd_tile(blockIdx.x, gridDim.x) : (long)blockIdx.x) * TK + t;
    const bool active = col < ncols;
    const long c = active ? col : 0;
    const auto lc = ld.template column<L>(c);
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = ld.template load<L, V>(lc, q, e);
    fft_pencil<T, L, true, false, V>(v, q, lds, LdsMap{t, tkp}, tw);
    const bool colzero = zero_kz_nyquist && (int)(c % G::KC) == G::NH;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        // keep the source modes: not the zero band, not the source Nyquist
        const bool keep = !colzero && !G::band(q, e) && q + G::P * e != G::NH;
        v[e] = keep ? cscale(v[e], scale) : cx<T>{0, 0};
    }
    fft_pencil<T, L, false, false, V>(v, q, lds, LdsMap{t, tkp}, tw);
    if (active) {
        const auto sc = st.template column<L>(c);
#pragma unroll
        for (int e = 0; e < E; ++e) st.template store<L, V>(sc, q, e, v[e]);
    }
}

// axis-0 pencils of the compact buffer over every row j', column = (j', kz)
template <typename T> struct LdAx0 {
    const cx<T>* kb;
    int np, kc;
    struct Col { const cx<T>* p; };
    template <int L> __device__ __forceinline__ Col column(long col) const { return Col{kb + col}; }
    template <int L, int V = 0> __device__ __forceinline__ cx<T> load(const Col& c, int q, int e) const {
        using G = PadGeo<L, V>;
        return c.p[(long)(q + G::P * e) * (L * G::KC)];
    }
};
template <typename T> struct StAx0 {
    cx<T>* kb;
    int np, kc;
    struct Col { cx<T>* p; };
    template <int L> __device__ __forceinline__ Col column(long col) const { return Col{kb + col}; }
    template <int L, int V = 0> __device__ __forceinline__ void store(const Col& c, int q, int e, cx<T> v) const {
        using G = PadGeo<L, V>;
        c.p[(long)(q + G::P * e) * (L * G::KC)] = v;
    }
};

// compact half-spectrum rows (kc <= M+1 values, the rest zero) -> real rows
template <typename T, int M, int RW>
__global__ void __launch_bounds__(Plan<M>::P * RW)
k_rows_c2r_pruned(const cx<T>* __restrict__ kin, int kc, T* __restrict__ rout, long rpitch, long nrows, T scale,
                  const cx<T>* __restrict__ tw) {
    constexpr int E = Plan<M>::E, P = Plan<M>::P, RP = row_lds_pitch<M>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int q = threadIdx.x % P, rw = threadIdx.x / P;
    const long row = (long)blockIdx.x * RW + rw;
    const bool active = row < nrows;
    const cx<T>* src = kin + (active ? row : 0) * kc;
    const LdsRowMap lm{rw * RP};
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = q + P * e;
        cx<T> x = (active && k < kc) ? src[k] : cx<T>{0, 0};
        if (k == 0) x.y = 0;
        lds[lm.at(k)] = x;
    }
    if (q == 0) lds[lm.at(M)] = (active && M < kc) ? cx<T>{src[M].x, (T)0} : cx<T>{0, 0};
    __syncthreads();
    const cx<T>* uw = tw + plan_tw_size<M>();
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = q + P * e;
        const cx<T> xk = lds[lm.at(k)], xm = lds[lm.at(M - k)];
        const cx<T> A = {xk.x + xm.x, xk.y - xm.y};
        const cx<T> D = {xk.x - xm.x, xk.y + xm.y};
        const cx<T> B = cmulc(D, uw[k]);
        v[e] = cx<T>{A.x - B.y, A.y + B.x};
    }
    __syncthreads();
    fft_regs_row<T, M, false>(v, q, lds, rw * RP, tw);
    if (active) {
        cx<T>* dst = reinterpret_cast<cx<T>*>(rout + row * rpitch);
#pragma unroll
        for (int e = 0; e < E; ++e) dst[q + P * e] = cscale(v[e], scale);
    }
}

// real rows = sum_t c_t A_t B_t (FP64 sums; an array used by several terms
// is re-read from cache) -> truncated half-spectrum rows (first kc values)
struct RowProd {
    const void* a[16];
    const void* b[16];
    double c[16];
    int nterm, narr;
};

template <typename T, int M, int RW>
__global__ void __launch_bounds__(Plan<M>::P * RW)
k_rows_r2c_prod(RowProd pr, long rpitch, cx<T>* __restrict__ kout, int kc, long nrows, T scale,
                const cx<T>* __restrict__ tw) {
    constexpr int E = Plan<M>::E, P = Plan<M>::P, RP = row_lds_pitch<M>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int q = threadIdx.x % P, rw = threadIdx.x / P;
    const long row = (long)blockIdx.x * RW + rw;
    const bool active = row < nrows;
    const long off = (active ? row : 0) * rpitch;
    cx<T> v[E];
    // terms outermost, CH elements per chunk: all loads of a chunk are in
    // flight together
    constexpr int CH = (E % 8 == 0) ? 8 : (E % 6 == 0 ? 6 : (E % 4 == 0 ? 4 : E));
#pragma unroll
    for (int e0 = 0; e0 < E; e0 += CH) {
        double s0[CH], s1[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) s0[u] = s1[u] = 0.0;
        if (active)
            for (int t = 0; t < pr.nterm; ++t) {
                const cx<T>* A = reinterpret_cast<const cx<T>*>(static_cast<const T*>(pr.a[t]) + off);
                const cx<T>* B = reinterpret_cast<const cx<T>*>(static_cast<const T*>(pr.b[t]) + off);
                cx<T> x[CH], y[CH];
#pragma unroll
                for (int u = 0; u < CH; ++u) { x[u] = A[q + P * (e0 + u)]; y[u] = B[q + P * (e0 + u)]; }
                const double ct = pr.c[t];
#pragma unroll
                for (int u = 0; u < CH; ++u) {
                    s0[u] += ct * ((double)x[u].x * (double)y[u].x);
                    s1[u] += ct * ((double)x[u].y * (double)y[u].y);
                }
            }
#pragma unroll
        for (int u = 0; u < CH; ++u) v[e0 + u] = cx<T>{(T)s0[u], (T)s1[u]};
    }
    fft_regs_row<T, M, true>(v, q, lds, rw * RP, tw);
    const LdsRowMap lm{rw * RP};
#pragma unroll
    for (int e = 0; e < E; ++e) lds[lm.at(q + P * e)] = v[e];
    __syncthreads();
    const cx<T>* uw = tw + plan_tw_size<M>();
    if (active) {
        cx<T>* dst = kout + row * kc;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int k = q + P * e;
            if (k >= kc) continue;
            const cx<T> zk = v[e], zm = lds[lm.at(k == 0 ? 0 : M - k)];
            const cx<T> Ek = {(zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5};
            const cx<T> Ok = {(zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5};
            dst[k] = cscale(cadd(Ek, cmul(uw[k], Ok)), scale);
        }
        if (q == 0 && M < kc) {
            const cx<T> z0 = lds[lm.at(0)];
            dst[M] = cx<T>{(z0.x - z0.y) * scale, (T)0};
        }
    }
}

// ------------------------------------------------------------- launchers --

// tile of a padded pass: the override of its own kind, else the override of
// every strided pass of the length (kind 0), else the widest tile whose LDS
// fits the default budget
template <typename T, int L, bool SPLIT, int V>
static int pick_tile(mfic_ctx ctx, int pass) {
    constexpr int P = Plan<L, V>::P;
    // the axis-1 truncation filter measured best with 8-pencil tiles at
    // 768-point FP32 (2.87 against 3.04 ms for 16 per 512^3 step): three
    // workgroups per CU instead of one
    int tk = pass == MFIC_PASS_TRUNC_AX1 ? 8 : 64;
    size_t cap = pencil_lds_cap<SPLIT>();
    const int prec = sizeof(T) == 8 ? MFIC_F64 : MFIC_F32;
    auto ov = ctx->tiles.find(std::make_tuple(L, pass, prec));
    if (ov == ctx->tiles.end()) ov = ctx->tiles.find(std::make_tuple(L, 0, prec));
    if (ov != ctx->tiles.end()) { tk = ov->second.first; cap = 139264; }
    while (tk > 1 && ((size_t)L * tk * pencil_lds_elem<T, SPLIT>() > cap || P * tk > 1024)) tk /= 2;
    return V != 0 && tk < 4 ? 4 : tk;  // alternative plans are built for tiles of 4 and more
}

template <typename T, int L, bool FWD, class LD, class ST, int V>
static void launch_fv(mfic_ctx ctx, int pass, const LD& ld, const ST& st, long ncols, const char* name, double bytes,
                      bool xcd) {
    constexpr int P = Plan<L, V>::P;
    const int tk = pick_tile<T, L, true, V>(ctx, pass);
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, L, 0, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32, V));
    const size_t lds = (size_t)L * tk * pencil_lds_elem<T, true>();
    // axis-1 passes: 16-column tiles of rows with an odd pitch share cache
    // lines with their neighbours, which the XCD-aware order keeps in one L2
    const long ntiles = (ncols + tk - 1) / tk;
    dim3 grid((unsigned)(xcd ? xcd_grid(ntiles) : ntiles));
    ProfScope prof(ctx, std::string(name) + "_L" + std::to_string(L) + (sizeof(T) == 8 ? "_f64" : "_f32"), bytes);
    switch (tk) {
#define TKC(K)                                                                                                  \
    case K:                                                                                                     \
        if constexpr (P * K <= 1024 && (V == 0 || K >= 4))                                                      \
            hipLaunchKernelGGL((k_pencils_f<T, L, K, FWD, LD, ST, V>), grid, dim3(P * K), lds, ctx->stream, ld, \
                               st, ncols, tw, K, xcd ? 1 : 0);                                                  \
        break;
        TKC(1) TKC(2) TKC(4) TKC(8) TKC(16) TKC(32) TKC(64)
#undef TKC
    default:
        throw Error(MFIC_ERR_SIZE, "bad tile");
    }
    MFIC_LAUNCH_CHECK();
}

// a strided padded pass with the radix plan selected for its kind (pass)
template <typename T, int L, bool FWD, class LD, class ST>
static void launch_f(mfic_ctx ctx, int pass, const LD& ld, const ST& st, long ncols, const char* name, double bytes,
                     bool xcd = false) {
    if (plan_variant(ctx, L, pass, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32) == 1)
        launch_fv<T, L, FWD, LD, ST, alt_plan<L>()>(ctx, pass, ld, st, ncols, name, bytes, xcd);
    else
        launch_fv<T, L, FWD, LD, ST, 0>(ctx, pass, ld, st, ncols, name, bytes, xcd);
}

template <typename T, int M>
static int pick_rows(mfic_ctx ctx, bool r2c) {
    constexpr int P = Plan<M>::P;
    int rw = std::max(1, std::min(64, (r2c ? 512 : 256) / P / (M >= 256 ? 4 : 1)));
    int p2 = 1;
    while (p2 * 2 <= rw) p2 *= 2;
    rw = p2;
    const auto ov = ctx->tiles.find(std::make_tuple(M, 1, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32));
    if (ov != ctx->tiles.end()) rw = ov->second.first;
    while (rw > 1 && ((size_t)row_lds_pitch<M>() * rw * sizeof(cx<T>) > 139264 || P * rw > 1024)) rw /= 2;
    return rw;
}

template <typename T, int M>
static void rows_c2r_M(mfic_ctx ctx, const cx<T>* kin, int kc, T* rout, long rpitch, long nrows, T scale) {
    constexpr int P = Plan<M>::P;
    const int rw = pick_rows<T, M>(ctx, false);
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, M, 1, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32));
    const size_t lds = (size_t)row_lds_pitch<M>() * rw * sizeof(cx<T>);
    dim3 grid((unsigned)((nrows + rw - 1) / rw));
    ProfScope prof(ctx, std::string("fft_c2rC_n") + std::to_string(2 * M) + (sizeof(T) == 8 ? "_f64" : "_f32"),
                   (double)nrows * ((double)kc * sizeof(cx<T>) + 2.0 * M * sizeof(T)));
    switch (rw) {
#define RWC(K)                                                                                                   \
    case K:                                                                                                      \
        if constexpr (P * K <= 1024)                                                                             \
            hipLaunchKernelGGL((k_rows_c2r_pruned<T, M, K>), grid, dim3(P * K), lds, ctx->stream, kin, kc, rout, \
                               rpitch, nrows, scale, tw);                                                        \
        break;
        RWC(1) RWC(2) RWC(4) RWC(8) RWC(16) RWC(32) RWC(64)
#undef RWC
    default:
        throw Error(MFIC_ERR_SIZE, "bad row tile");
    }
    MFIC_LAUNCH_CHECK();
}

template <typename T, int M>
static void rows_r2c_M(mfic_ctx ctx, const RowProd& pr, long rpitch, cx<T>* kout, int kc, long nrows, T scale) {
    constexpr int P = Plan<M>::P;
    const int rw = pick_rows<T, M>(ctx, true);
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, M, 1, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32));
    const size_t lds = (size_t)row_lds_pitch<M>() * rw * sizeof(cx<T>);
    dim3 grid((unsigned)((nrows + rw - 1) / rw));
    ProfScope prof(ctx, std::string("fft_r2cC_n") + std::to_string(2 * M) + "_a" + std::to_string(pr.narr) +
                            (sizeof(T) == 8 ? "_f64" : "_f32"),
                   (double)nrows * (2.0 * M * sizeof(T) * pr.narr + (double)kc * sizeof(cx<T>)));
    switch (rw) {
#define RWC(K)                                                                                                 \
    case K:                                                                                                    \
        if constexpr (P * K <= 1024)                                                                           \
            hipLaunchKernelGGL((k_rows_r2c_prod<T, M, K>), grid, dim3(P * K), lds, ctx->stream, pr, rpitch, kout, \
                               kc, nrows, scale, tw);                                                          \
        break;
        RWC(1) RWC(2) RWC(4) RWC(8) RWC(16) RWC(32) RWC(64)
#undef RWC
    default:
        throw Error(MFIC_ERR_SIZE, "bad row tile");
    }
    MFIC_LAUNCH_CHECK();
}

#define MFIC_PADDED_SIZES(X) X(12) X(18) X(24) X(36) X(48) X(72) X(96) X(144) X(192) X(288) X(384) X(576) X(768)

template <typename T>
static void c2r_from_operand_t(mfic_ctx ctx, const OperandDesc& o, size_t n, double L, void* kbuf, void* rout) {
    const int N = (int)n, np = 3 * N / 2, kc = N / 2 + 1;
    const double rfac = std::pow(1.5, 1.5);
    const double norm = 1.0 / std::sqrt((double)np * (double)np * (double)np);
    const double kf = 2.0 * 3.14159265358979323846 / L;
    LdOperand<T> la{static_cast<const cx<T>*>(o.f), N, kc, o.kind, o.d[0], o.d[1], o.d[2], o.d[3], kf, rfac};
    StPadCol<T> sa{static_cast<cx<T>*>(kbuf), N, np, kc};
    LdAx1<T, true> lb{static_cast<const cx<T>*>(kbuf)};
    StAx1<T, false> sb{static_cast<cx<T>*>(kbuf)};
    const double cs = sizeof(cx<T>);
    const long colsA = (long)N * kc, colsB = (long)np * kc;
    switch (np) {
#define X(S)                                                                                                 \
    case S:                                                                                                  \
        launch_f<T, S, false>(ctx, MFIC_PASS_PAD_C2R_A, la, sa, colsA, "fft_c2rA", (double)colsA * (N + np) * cs);                    \
        launch_f<T, S, false>(ctx, MFIC_PASS_PAD_C2R_B, lb, sb, colsB, "fft_c2rB", (double)colsB * (N + np) * cs, true);                    \
        rows_c2r_M<T, S / 2>(ctx, static_cast<const cx<T>*>(kbuf), kc, static_cast<T*>(rout), np, (long)np * np, \
                             (T)norm);                                                                       \
        break;
        MFIC_PADDED_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported padded size " + std::to_string(np));
    }
}

template <typename T>
static void r2c_products_t(mfic_ctx ctx, const std::vector<ProdTerm>& terms, size_t n, double L, void* kbuf, void* out,
                           int op, double coef, bool invlap) {
    const int N = (int)n, np = 3 * N / 2, kc = N / 2 + 1;
    RowProd pr{};
    std::vector<const void*> arrs;
    MFIC_CHECK(!terms.empty() && terms.size() <= 16, MFIC_ERR_ARG, "1..16 product terms");
    auto slot = [&](const void* p) {
        auto it = std::find(arrs.begin(), arrs.end(), p);
        if (it != arrs.end()) return (int)(it - arrs.begin());
        arrs.push_back(p);
        return (int)arrs.size() - 1;
    };
    for (size_t t = 0; t < terms.size(); ++t) {
        slot(terms[t].a);
        slot(terms[t].b);
        pr.a[t] = terms[t].a;
        pr.b[t] = terms[t].b;
        pr.c[t] = terms[t].c;
    }
    pr.narr = (int)arrs.size();
    pr.nterm = (int)terms.size();
    const double rfac = std::sqrt((double)np * np * np) / std::sqrt((double)N * N * N);
    const double norm = 1.0 / std::sqrt((double)np * (double)np * (double)np);
    const double kf = 2.0 * 3.14159265358979323846 / L;
    LdAx1<T, false> lb{static_cast<const cx<T>*>(kbuf)};
    StAx1<T, true> sb{static_cast<cx<T>*>(kbuf)};
    LdPadCol<T> la{static_cast<const cx<T>*>(kbuf), N, np, kc};
    StUnpad<T> sa{static_cast<cx<T>*>(out), N, kc, op, invlap ? 1 : 0, norm / rfac, coef, kf};
    const double cs = sizeof(cx<T>);
    const long colsA = (long)N * kc, colsB = (long)np * kc;
    switch (np) {
#define X(S)                                                                                                        \
    case S:                                                                                                         \
        rows_r2c_M<T, S / 2>(ctx, pr, np, static_cast<cx<T>*>(kbuf), kc, (long)np * np, (T)1);                      \
        launch_f<T, S, true>(ctx, MFIC_PASS_PAD_R2C_B, lb, sb, colsB, "fft_r2cB", (double)colsB * (np + N) * cs, true);                            \
        launch_f<T, S, true>(ctx, MFIC_PASS_PAD_R2C_A, la, sa, colsA, "fft_r2cA", (double)colsA * (np + N * (op == MFIC_OP_ASSIGN ? 1 : 2)) * cs); \
        break;
        MFIC_PADDED_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported padded size " + std::to_string(np));
    }
}

template <typename T>
static void c2r_to_kbuf_t(mfic_ctx ctx, const OperandDesc& o, size_t n, double L, void* kbuf) {
    const int N = (int)n, np = 3 * N / 2, kc = N / 2 + 1;
    const double rfac = std::pow(1.5, 1.5);
    const double kf = 2.0 * 3.14159265358979323846 / L;
    LdOperand<T> la{static_cast<const cx<T>*>(o.f), N, kc, o.kind, o.d[0], o.d[1], o.d[2], o.d[3], kf, rfac};
    StPadCol<T> sa{static_cast<cx<T>*>(kbuf), N, np, kc};
    LdAx1<T, true> lb{static_cast<const cx<T>*>(kbuf)};
    StAx1<T, false> sb{static_cast<cx<T>*>(kbuf)};
    const double cs = sizeof(cx<T>);
    const long colsA = (long)N * kc, colsB = (long)np * kc;
    switch (np) {
#define X(S)                                                                                   \
    case S:                                                                                    \
        launch_f<T, S, false>(ctx, MFIC_PASS_PAD_C2R_A, la, sa, colsA, "fft_c2rA", (double)colsA * (N + np) * cs); \
        launch_f<T, S, false>(ctx, MFIC_PASS_PAD_C2R_B, lb, sb, colsB, "fft_c2rB", (double)colsB * (N + np) * cs, true); \
        break;
        MFIC_PADDED_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported padded size " + std::to_string(np));
    }
}

template <typename T>
static void r2c_from_kbuf_t(mfic_ctx ctx, size_t n, double L, void* kbuf, void* out, int op, double coef, bool invlap) {
    const int N = (int)n, np = 3 * N / 2, kc = N / 2 + 1;
    const double rfac = std::sqrt((double)np * np * np) / std::sqrt((double)N * N * N);
    const double norm = 1.0 / std::sqrt((double)np * (double)np * (double)np);
    const double kf = 2.0 * 3.14159265358979323846 / L;
    LdAx1<T, false> lb{static_cast<const cx<T>*>(kbuf)};
    StAx1<T, true> sb{static_cast<cx<T>*>(kbuf)};
    LdPadCol<T> la{static_cast<const cx<T>*>(kbuf), N, np, kc};
    StUnpad<T> sa{static_cast<cx<T>*>(out), N, kc, op, invlap ? 1 : 0, norm / rfac, coef, kf};
    const double cs = sizeof(cx<T>);
    const long colsA = (long)N * kc, colsB = (long)np * kc;
    switch (np) {
#define X(S)                                                                                                        \
    case S:                                                                                                         \
        launch_f<T, S, true>(ctx, MFIC_PASS_PAD_R2C_B, lb, sb, colsB, "fft_r2cB", (double)colsB * (np + N) * cs, true);                            \
        launch_f<T, S, true>(ctx, MFIC_PASS_PAD_R2C_A, la, sa, colsA, "fft_r2cA", (double)colsA * (np + N * (op == MFIC_OP_ASSIGN ? 1 : 2)) * cs); \
        break;
        MFIC_PADDED_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported padded size " + std::to_string(np));
    }
}

// The six Hessians (xx yy zz xy xz yz) of an N-grid field, factorised per
// axis: -g_a g_b = -(g_i^p)(g_j^q)(g_k^r).  Pass A runs once per x power
// (3 launches into `xbuf`, three consecutive compact buffers), pass B applies
// the y power along its pencils and the sign and z power per column.
template <typename T>
static void hessians_to_kbuf_t(mfic_ctx ctx, const void* f, size_t n, double L, void* xbuf, void* out, long ostride) {
    const int N = (int)n, np = 3 * N / 2, kc = N / 2 + 1;
    const double rfac = std::pow(1.5, 1.5);
    const double kf = 2.0 * 3.14159265358979323846 / L;
    const long stride = (long)np * np * kc;
    cx<T>* xb = static_cast<cx<T>*>(xbuf);
    cx<T>* ob = static_cast<cx<T>*>(out);
    static const int PQR[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    const double cs = sizeof(cx<T>);
    const long colsA = (long)N * kc, colsB = (long)np * kc;
    switch (np) {
#define X(S)                                                                                                   \
    case S:                                                                                                    \
        for (int p = 2; p >= 0; --p) {                                                                         \
            LdPow<T> la{static_cast<const cx<T>*>(f), N, kc, p, (T)kf, (T)rfac};                               \
            StPadCol<T> sa{xb + p * stride, N, np, kc};                                                        \
            launch_f<T, S, false>(ctx, MFIC_PASS_PAD_C2R_A, la, sa, colsA, "fft_c2rAh", (double)colsA * (N + np) * cs);            \
            for (int h = 0; h < 6; ++h) {                                                                      \
                if (PQR[h][0] != p) continue;                                                                  \
                LdAx1Mul<T> lb{xb + p * stride, N, np, kc, PQR[h][1], PQR[h][2], (T)kf, (T)-1};                \
                StAx1<T, false> sb{ob + h * ostride};                                                   \
                launch_f<T, S, false>(ctx, MFIC_PASS_PAD_C2R_B, lb, sb, colsB, "fft_c2rBh", (double)colsB * (N + np) * cs, true);        \
            }                                                                                                  \
        }                                                                                                      \
        break;
        MFIC_PADDED_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported padded size " + std::to_string(np));
    }
}

void padded_hessians_to_kbuf(mfic_ctx ctx, int prec, const void* f, size_t n, double L, void* xbuf, void* out) {
    const long stride = (long)(3 * n / 2) * (3 * n / 2) * (n / 2 + 1);
    if (prec == MFIC_F64) hessians_to_kbuf_t<double>(ctx, f, n, L, xbuf, out, stride);
    else hessians_to_kbuf_t<float>(ctx, f, n, L, xbuf, out, stride);
}

template <typename T, int L, class LD, class ST, int V>
static void launch_filter_v(mfic_ctx ctx, int pass, const LD& ld, const ST& st, long ncols, bool xcd, int zk,
                            double scale, const char* name) {
    constexpr int P = Plan<L, V>::P;
    const int tk = pick_tile<T, L, false, V>(ctx, pass);
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, L, 0, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32, V));
    const size_t lds = (size_t)L * tk * pencil_lds_elem<T, false>();
    const long ntiles = (ncols + tk - 1) / tk;
    dim3 grid((unsigned)(xcd ? xcd_grid(ntiles) : ntiles));
    ProfScope prof(ctx, std::string(name) + "_L" + std::to_string(L) + (sizeof(T) == 8 ? "_f64" : "_f32"),
                   2.0 * (double)ncols * L * sizeof(cx<T>));
    switch (tk) {
#define TKC(K)                                                                                                  \
    case K:                                                                                                     \
        if constexpr (P * K <= 1024 && (V == 0 || K >= 4))                                                      \
            hipLaunchKernelGGL((k_pencils_filter<T, L, K, LD, ST, V>), grid, dim3(P * K), lds, ctx->stream, ld, \
                               st, ncols, tw, K, xcd ? 1 : 0, zk, (T)scale);                                    \
        break;
        TKC(1) TKC(2) TKC(4) TKC(8) TKC(16) TKC(32) TKC(64)
#undef TKC
    default:
        throw Error(MFIC_ERR_SIZE, "bad tile");
    }
    MFIC_LAUNCH_CHECK();
}

template <typename T, int L, class LD, class ST>
static void launch_filter(mfic_ctx ctx, int pass, const LD& ld, const ST& st, long ncols, bool xcd, int zk,
                          double scale, const char* name) {
    if (plan_variant(ctx, L, pass, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32) == 1)
        launch_filter_v<T, L, LD, ST, alt_plan<L>()>(ctx, pass, ld, st, ncols, xcd, zk, scale, name);
    else
        launch_filter_v<T, L, LD, ST, 0>(ctx, pass, ld, st, ncols, xcd, zk, scale, name);
}

template <typename T>
static void truncate_repad_t(mfic_ctx ctx, size_t n, void* kbuf) {
    const int N = (int)n, np = 3 * N / 2, kc = N / 2 + 1;
    const double norm = 1.0 / std::sqrt((double)np * (double)np * (double)np);
    cx<T>* kb = static_cast<cx<T>*>(kbuf);
    LdAx1<T, false> l1{kb};
    StAx1<T, false> s1{kb};
    LdAx0<T> l0{kb, np, kc};
    StAx0<T> s0{kb, np, kc};
    const long cols = (long)np * kc;
    switch (np) {
#define X(S)                                                                                       \
    case S:                                                                                        \
        launch_filter<T, S>(ctx, MFIC_PASS_TRUNC_AX1, l1, s1, cols, true, 1, norm, "fft_truncate_ax1");       \
        launch_filter<T, S>(ctx, MFIC_PASS_TRUNC_AX0, l0, s0, cols, false, 0, 1.0, "fft_truncate_ax0");       \
        break;
        MFIC_PADDED_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported padded size " + std::to_string(np));
    }
}

void padded_truncate_repad(mfic_ctx ctx, int prec, size_t n, void* kbuf) {
    if (prec == MFIC_F64) truncate_repad_t<double>(ctx, n, kbuf);
    else truncate_repad_t<float>(ctx, n, kbuf);
}

void padded_c2r_to_kbuf(mfic_ctx ctx, int prec, const OperandDesc& o, size_t n, double L, void* kbuf) {
    if (prec == MFIC_F64) c2r_to_kbuf_t<double>(ctx, o, n, L, kbuf);
    else c2r_to_kbuf_t<float>(ctx, o, n, L, kbuf);
}

void padded_r2c_from_kbuf(mfic_ctx ctx, int prec, size_t n, double L, void* kbuf, void* out, int op, double coef,
                          bool invlap) {
    if (prec == MFIC_F64) r2c_from_kbuf_t<double>(ctx, n, L, kbuf, out, op, coef, invlap);
    else r2c_from_kbuf_t<float>(ctx, n, L, kbuf, out, op, coef, invlap);
}

void padded_c2r_from_operand(mfic_ctx ctx, int prec, const OperandDesc& o, size_t n, double L, void* kbuf, void* rout) {
    if (prec == MFIC_F64) c2r_from_operand_t<double>(ctx, o, n, L, kbuf, rout);
    else c2r_from_operand_t<float>(ctx, o, n, L, kbuf, rout);
}

void padded_r2c_products(mfic_ctx ctx, int prec, const std::vector<ProdTerm>& terms, size_t n, double L, void* kbuf,
                         void* out, int op, double coef, bool invlap) {
    if (prec == MFIC_F64) r2c_products_t<double>(ctx, terms, n, L, kbuf, out, op, coef, invlap);
    else r2c_products_t<float>(ctx, terms, n, L, kbuf, out, op, coef, invlap);
}

}  // namespace mfic
