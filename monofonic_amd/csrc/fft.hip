// 3D real<->complex FFTs on the Grid_FFT layout [n0][n1][n2+2] (gfx950).
//
// Forward (r2c, FFTW sign e^{-ikx}): rows pass along the contiguous axis
// (real rows of n2 -> n2/2+1 complex via a length-n2/2 complex FFT and an
// untangle step), then two strided pencil passes along axes 1 and 0.
// Backward (c2r): pencils along axes 0 and 1, then the rows pass.  The
// unitary scale 1/sqrt(n0 n1 n2) (grid_fft.cc:69, 199-205) is folded into the
// last pass.  Each pass reads and writes the array once.
#include <cmath>
#include <complex>

#include "fft_core.hpp"
#include "mfic_internal.hpp"

namespace mfic {

// ------------------------------------------------------------ kernels -----

// Transforms of length L along a strided axis.  Block = TK neighbouring
// pencils (contiguous t) x P threads per pencil; a wavefront reads TK
// consecutive complex values per pencil position.
template <typename T, int L, int TK, bool FWD, int V = 0>
__global__ void __launch_bounds__((Plan<L, V>::P * TK))
k_pencils(cx<T>* __restrict__ data, long ps, long os, int tlen, T scale,
          const cx<T>* __restrict__ tw, int tkp, int gx, long ntiles, int xcd) {
    constexpr int E = Plan<L, V>::E, P = Plan<L, V>::P;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int t = threadIdx.x % TK, q = threadIdx.x / TK;
    // one-dimensional grid of gx tiles per pencil group (XCD-aware order for
    // the narrow groups of the axis-1 pass, whose tiles share cache lines)
    const long tile = xcd ? xcd_tile(blockIdx.x, gridDim.x) : (long)blockIdx.x;
    const long outer = tile / gx;
    const int tg = (int)(tile - outer * gx) * TK + t;
    const bool active = tg < tlen && tile < ntiles;
    cx<T>* base = data + (active ? outer : 0) * os + tg;
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = active ? base[(long)(q + P * e) * ps] : cx<T>{0, 0};
    fft_pencil<T, L, FWD, true, V>(v, q, lds, LdsMap{t, tkp}, tw);
    if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) base[(long)(q + P * e) * ps] = cscale(v[e], scale);
    }
}

// Axis-0 c2r pass of a k-space component that also produces the component
// of the (1/2, 1/2, 1/2)-cell shifted sub-lattice: the shifted copy is the
// column times exp(i pi (i' + j' + k') / n) (signed indices, shift_field of
// grid_fft.hh:832-842 as particle_generator.hh:518-522 uses it), transformed
// from the same loads.  Column t = j * nc + k of an [n0][n1][nc] grid.
// GEN: both outputs carry a shift of wa / 2 and wb / 2 cells per axis
// (weights 0, 1 or 2: the secondary lattices of particle_generator.hh:58-63
// and their sub-lattices), the in-place output included.
template <typename T, int L, int TK, int V = 0, bool GEN = false>
__global__ void __launch_bounds__((Plan<L, V>::P * TK))
k_pencils_ax0_shift(cx<T>* __restrict__ data, cx<T>* __restrict__ shifted, int n1, int nc, T scale,
                    const cx<T>* __restrict__ tw, const cx<double>* __restrict__ half, int tkp, int3 wa, int3 wb) {
    constexpr int E = Plan<L, V>::E, P = Plan<L, V>::P;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int t = threadIdx.x % TK, q = threadIdx.x / TK;
    const int tlen = n1 * nc;
    const int tg = (int)blockIdx.x * TK + t;
    const bool active = tg < tlen;
    const int col = active ? tg : 0;
    const long ps = (long)tlen;
    const int j = col / nc, k = col - (col / nc) * nc;
    const int js = j - (j > n1 / 2) * n1;  // kz index k <= n/2 is its own signed index
    const int mc = GEN ? wb.y * js + wb.z * k : js + k;
    const int ma = GEN ? wa.y * js + wa.z * k : 0;
    cx<T> v[E], w[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = data[col + (long)(q + P * e) * ps];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = q + P * e, is = i - (i > L / 2) * L;
        if (GEN) {
            // exp(i pi m / L) with m = w . (signed i, j, k), reduced mod 2L
            int m = (wb.x * is + mc) % (2 * L);
            m += m < 0 ? 2 * L : 0;
            const cx<double> ph = half[m];
            w[e] = cmul(v[e], cx<T>{(T)ph.x, (T)ph.y});
            int m0 = (wa.x * is + ma) % (2 * L);
            m0 += m0 < 0 ? 2 * L : 0;
            const cx<double> p0 = half[m0];
            v[e] = cmul(v[e], cx<T>{(T)p0.x, (T)p0.y});
        } else {
            int m = is + mc;
            m = m < 0 ? m + 2 * L : (m >= 2 * L ? m - 2 * L : m);
            const cx<double> ph = half[m];
            w[e] = cmul(v[e], cx<T>{(T)ph.x, (T)ph.y});
        }
    }
    // both transforms in lockstep: shared twiddle loads and barriers
    fft_pencil2<T, L, false, V>(v, w, q, reinterpret_cast<T*>(lds), LdsMap{t, tkp}, LdsMap{L * tkp + t, tkp}, tw);
    if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) data[col + (long)(q + P * e) * ps] = cscale(v[e], scale);
#pragma unroll
        for (int e = 0; e < E; ++e) shifted[col + (long)(q + P * e) * ps] = cscale(w[e], scale);
    }
}

// Real rows of length 2M -> M+1 complex, in place (row pitch in reals).
template <typename T, int M, int RW>
__global__ void __launch_bounds__(Plan<M>::P * RW)
k_rows_r2c(T* __restrict__ data, long nrows, long pitch, T scale, const cx<T>* __restrict__ tw) {
    constexpr int E = Plan<M>::E, P = Plan<M>::P, RP = row_lds_pitch<M>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int q = threadIdx.x % P, rw = threadIdx.x / P;
    const LdsRowMap lm{rw * RP};
    const long row = (long)blockIdx.x * RW + rw;
    const bool active = row < nrows;
    cx<T>* r = reinterpret_cast<cx<T>*>(data + (active ? row : 0) * pitch);
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = active ? r[q + P * e] : cx<T>{0, 0};
    fft_regs_row<T, M, true>(v, q, lds, rw * RP, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) lds[lm.at(q + P * e)] = v[e];
    __syncthreads();
    const cx<T>* uw = tw + plan_tw_size<M>();
    if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int k = q + P * e;
            const cx<T> zk = v[e], zm = lds[lm.at(k == 0 ? 0 : M - k)];
            const cx<T> Ek = {(zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5};
            const cx<T> Ok = {(zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5};
            const cx<T> X = cadd(Ek, cmul(uw[k], Ok));
            r[k] = cscale(X, scale);
        }
        if (q == 0) {
            const cx<T> z0 = lds[lm.at(0)];
            r[M] = cx<T>{(z0.x - z0.y) * scale, (T)0};
        }
    }
}

// M+1 complex -> real rows of length 2M, in place.  The imaginary parts of
// the kz = 0 and kz = M terms are discarded (Hermitian projection).
template <typename T, int M, int RW>
__global__ void __launch_bounds__(Plan<M>::P * RW)
k_rows_c2r(T* __restrict__ data, long nrows, long pitch, T scale, const cx<T>* __restrict__ tw) {
    constexpr int E = Plan<M>::E, P = Plan<M>::P, RP = row_lds_pitch<M>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int q = threadIdx.x % P, rw = threadIdx.x / P;
    const LdsRowMap lm{rw * RP};
    const long row = (long)blockIdx.x * RW + rw;
    const bool active = row < nrows;
    cx<T>* r = reinterpret_cast<cx<T>*>(data + (active ? row : 0) * pitch);
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        cx<T> x = active ? r[q + P * e] : cx<T>{0, 0};
        if (q + P * e == 0) x.y = 0;
        lds[lm.at(q + P * e)] = x;
        v[e] = x;
    }
    if (q == 0) lds[lm.at(M)] = active ? cx<T>{r[M].x, (T)0} : cx<T>{0, 0};
    __syncthreads();
    const cx<T>* uw = tw + plan_tw_size<M>();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = q + P * e;
        const cx<T> xk = v[e], xm = lds[lm.at(M - k)];
        // A = X_k + conj(X_{M-k}),  B = (X_k - conj(X_{M-k})) conj(W^k)
        const cx<T> A = {xk.x + xm.x, xk.y - xm.y};
        const cx<T> D = {xk.x - xm.x, xk.y + xm.y};
        const cx<T> B = cmulc(D, uw[k]);
        v[e] = cx<T>{A.x - B.y, A.y + B.x};  // A + i B
    }
    __syncthreads();
    fft_regs_row<T, M, false>(v, q, lds, rw * RP, tw);
    if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) r[q + P * e] = cscale(v[e], scale);
        if (q == 0) r[M] = cx<T>{0, 0};
    }
}

// ------------------------------------------------------------ tables ------

template <int L, int V = 0> static void fill_tw(std::vector<std::complex<long double>>& out) {
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int s = 0; s < Plan<L, V>::S; ++s) {
        const int R = Plan<L, V>::r(s), NS = plan_ns<L, V>(s);
        for (int k = 0; k < NS; ++k)
            for (int r = 1; r < R; ++r) {
                const long double a = -2.0L * PI * (long double)((long)r * k % ((long)NS * R)) / (long double)(NS * R);
                out.emplace_back(cosl(a), sinl(a));
            }
    }
}

#define MFIC_FOR_SIZES(X) \
    X(4) X(6) X(8) X(9) X(12) X(16) X(18) X(24) X(32) X(36) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(576) X(768)

bool fft_length_supported(size_t n) {
    switch (n) {
#define X(L) case L:
        MFIC_FOR_SIZES(X)
#undef X
        return true;
    default:
        return false;
    }
}

int fft_plan_count(int L) {
    switch (L) {
#define X(S) case S: return plan_variants<S>();
        MFIC_FOR_SIZES(X)
#undef X
    default:
        return 0;
    }
}

template <int L, int V> static void plan_info_t(int* elems, int radices[4]) {
    *elems = Plan<L, V>::E;
    for (int s = 0; s < 4; ++s) radices[s] = s < Plan<L, V>::S ? Plan<L, V>::r(s) : 0;
}

void fft_plan_info(int L, int variant, int* elems, int radices[4]) {
    MFIC_CHECK(variant >= 0 && variant < fft_plan_count(L), MFIC_ERR_ARG,
               "no radix plan " + std::to_string(variant) + " for length " + std::to_string(L));
    switch (L) {
#define X(S)                                                                         \
    case S:                                                                          \
        if (variant == 1) plan_info_t<S, alt_plan<S>()>(elems, radices);             \
        else plan_info_t<S, 0>(elems, radices);                                      \
        break;
        MFIC_FOR_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported FFT length " + std::to_string(L));
    }
}

int plan_variant(mfic_ctx ctx, int L, int pass, int prec) {
    if (ctx->plans.empty()) return 0;
    const auto it = ctx->plans.find(std::make_tuple(L, pass, prec));
    return it == ctx->plans.end() ? 0 : it->second;
}

static std::vector<std::complex<long double>> make_table(int L, int kind, int variant) {
    std::vector<std::complex<long double>> tw;
    if (kind == 2) {  // half-cell shift phases exp(+i pi m / L), m = 0 .. 2L-1
        const long double PI = 3.141592653589793238462643383279502884L;
        for (int m = 0; m < 2 * L; ++m) {
            const long double a = PI * (long double)m / (long double)L;
            tw.emplace_back(cosl(a), sinl(a));
        }
        return tw;
    }
    switch (L) {
#define X(S)                                                     \
    case S:                                                      \
        if (variant == 1) fill_tw<S, plan_or_default<S, 1>()>(tw);      \
        else if (variant == 2) fill_tw<S, plan_or_default<S, 2>()>(tw); \
        else if (variant == 3) fill_tw<S, plan_or_default<S, 3>()>(tw); \
        else fill_tw<S>(tw);                                            \
        break;
        MFIC_FOR_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported FFT length " + std::to_string(L));
    }
    if (kind == 1) {  // untangle factors exp(-2 pi i k / 2L), k = 0..L
        const long double PI = 3.141592653589793238462643383279502884L;
        for (int k = 0; k <= L; ++k) {
            if (k == 0) { tw.emplace_back(1.0L, 0.0L); continue; }
            if (k == L) { tw.emplace_back(-1.0L, 0.0L); continue; }
            if (2 * k == L) { tw.emplace_back(0.0L, -1.0L); continue; }
            const long double a = -PI * (long double)k / (long double)L;
            tw.emplace_back(cosl(a), sinl(a));
        }
    }
    return tw;
}

const void* twiddles(mfic_ctx ctx, int L, int kind, int prec, int variant) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    MFIC_CHECK(variant >= 0 && variant <= 3 && kind >= 0 && kind <= 2 && plan_exists(L, variant), MFIC_ERR_ARG,
               "no such radix plan");
    auto key = std::make_tuple(L, kind + 4 * variant, prec);
    auto it = ctx->tables.find(key);
    if (it != ctx->tables.end()) return it->second.dev;
    MFIC_CHECK(!ctx->capturing, MFIC_ERR_STATE,
               "twiddle table for length " + std::to_string(L) + " created during graph capture: run the sequence once before capturing");
    auto tw = make_table(L, kind, variant);
    TwTable t;
    t.count = tw.size();
    size_t bytes = std::max<size_t>(1, t.count) * (prec == MFIC_F64 ? 16 : 8);
    MFIC_HIP(hipMalloc(&t.dev, bytes));
    if (prec == MFIC_F64) {
        std::vector<cx<double>> h(t.count);
        for (size_t i = 0; i < t.count; ++i) h[i] = {(double)tw[i].real(), (double)tw[i].imag()};
        MFIC_HIP(hipMemcpy(t.dev, h.data(), t.count * sizeof(cx<double>), hipMemcpyHostToDevice));
    } else {
        std::vector<cx<float>> h(t.count);
        for (size_t i = 0; i < t.count; ++i) h[i] = {(float)tw[i].real(), (float)tw[i].imag()};
        MFIC_HIP(hipMemcpy(t.dev, h.data(), t.count * sizeof(cx<float>), hipMemcpyHostToDevice));
    }
    ctx->tables[key] = t;
    return t.dev;
}

// ------------------------------------------------------------ launchers ---

static int pow2_floor(int x) {
    int p = 1;
    while (p * 2 <= x) p *= 2;
    return p;
}
static size_t max_lds(mfic_ctx ctx) {
    if (ctx->max_lds == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess || v <= 0)
            v = 65536;
        ctx->max_lds = (size_t)v;
    }
    return ctx->max_lds;
}

template <typename T, int L, bool FWD, int V>
static void launch_pencils_LV(mfic_ctx ctx, cx<T>* data, long ps, long os, int tlen, long nouter, T scale,
                              const char* tag) {
    constexpr int P = Plan<L, V>::P;
    // widest tile (pencils per block) whose LDS fits ~136 KB: 128-byte row
    // segments per pencil position and one resident block per CU measured
    // best on MI355X for the 512 / 768 passes
    int tk = 64;
    while (tk > 4 && ((size_t)L * tk * pencil_lds_elem<T, true>() > pencil_lds_cap<true>() || P * tk > 1024)) tk /= 2;
    int pad = 0;
    const auto ov = ctx->tiles.find(std::make_tuple(L, 0, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32));
    if (ov != ctx->tiles.end()) { tk = ov->second.first; pad = ov->second.second; }
    while (tk > 1 && (size_t)L * (tk + pad) * pencil_lds_elem<T, true>() > max_lds(ctx)) tk /= 2;
    while (tk > 1 && P * tk > 1024) tk /= 2;
    if (V != 0 && tk < 4) tk = 4;  // alternative plans are built for tiles of 4 and more
    const int tkp = tk + pad;
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, L, 0, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32, V));
    const int gx = (tlen + tk - 1) / tk;
    const long ntiles = (long)gx * nouter;
    const bool xcd = nouter > 1;
    dim3 grid((unsigned)(xcd ? xcd_grid(ntiles) : ntiles));
    size_t lds = (size_t)L * tkp * pencil_lds_elem<T, true>();
    ProfScope prof(ctx, "fft_pencils_L" + std::to_string(L) + (FWD ? "_fwd" : "_inv") + tag + (sizeof(T) == 8 ? "_f64" : "_f32"),
                   2.0 * (double)tlen * L * nouter * sizeof(cx<T>));
    switch (tk) {
#define TKCASE(K)                                                                                               \
    case K:                                                                                                     \
        if constexpr (P * K <= 1024 && (V == 0 || K >= 4))                                                      \
            hipLaunchKernelGGL((k_pencils<T, L, K, FWD, V>), grid, dim3(P * K), lds, ctx->stream, data, ps, os, \
                               tlen, scale, tw, tkp, gx, ntiles, xcd ? 1 : 0);                                  \
        break;
        TKCASE(1) TKCASE(2) TKCASE(4) TKCASE(8) TKCASE(16) TKCASE(32) TKCASE(64)
#undef TKCASE
    default:
        throw Error(MFIC_ERR_SIZE, "bad tile");
    }
    MFIC_LAUNCH_CHECK();
}

template <typename T, int L, bool FWD>
static void launch_pencils_L(mfic_ctx ctx, cx<T>* data, long ps, long os, int tlen, long nouter, T scale,
                             const char* tag) {
    if (plan_variant(ctx, L, MFIC_PASS_PENCILS, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32) == 1)
        launch_pencils_LV<T, L, FWD, alt_plan<L>()>(ctx, data, ps, os, tlen, nouter, scale, tag);
    else
        launch_pencils_LV<T, L, FWD, 0>(ctx, data, ps, os, tlen, nouter, scale, tag);
}

template <typename T>
void fft_pencils(mfic_ctx ctx, void* data, int L, long ps, long os, int tlen, long nouter, bool fwd,
                 double scale, const char* tag) {
    cx<T>* d = static_cast<cx<T>*>(data);
    switch (L) {
#define X(S)                                                                                \
    case S:                                                                                 \
        if (fwd) launch_pencils_L<T, S, true>(ctx, d, ps, os, tlen, nouter, (T)scale, tag); \
        else launch_pencils_L<T, S, false>(ctx, d, ps, os, tlen, nouter, (T)scale, tag);    \
        break;
        MFIC_FOR_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported FFT length " + std::to_string(L));
    }
}

template <typename T, int M, bool R2C>
static void launch_rows_M(mfic_ctx ctx, T* data, size_t nrows, long pitch, T scale) {
    constexpr int P = Plan<M>::P;
    int rw = std::max(1, std::min(64, pow2_floor((R2C ? 512 : 256) / P / (M >= 256 ? 4 : 1))));
    const auto ov = ctx->tiles.find(std::make_tuple(M, 1, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32));
    if (ov != ctx->tiles.end()) rw = ov->second.first;
    while (rw > 1 && ((size_t)row_lds_pitch<M>() * rw * sizeof(cx<T>) > max_lds(ctx) || P * rw > 1024)) rw /= 2;
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, M, 1, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32));
    dim3 grid((unsigned)((nrows + rw - 1) / rw));
    size_t lds = (size_t)row_lds_pitch<M>() * rw * sizeof(cx<T>);
    ProfScope prof(ctx, std::string(R2C ? "fft_rows_r2c_n" : "fft_rows_c2r_n") + std::to_string(2 * M) + (sizeof(T) == 8 ? "_f64" : "_f32"),
                   (double)nrows * (2.0 * M + 2.0 * (M + 1)) * sizeof(T));
    switch (rw) {
#define RWCASE(K)                                                                                             \
    case K:                                                                                                   \
        if constexpr (P * K <= 1024) {                                                                        \
            if (R2C)                                                                                          \
                hipLaunchKernelGGL((k_rows_r2c<T, M, K>), grid, dim3(P * K), lds, ctx->stream, data, (long)nrows, \
                                   pitch, scale, tw);                                                         \
            else                                                                                              \
                hipLaunchKernelGGL((k_rows_c2r<T, M, K>), grid, dim3(P * K), lds, ctx->stream, data, (long)nrows, \
                                   pitch, scale, tw);                                                         \
        }                                                                                                     \
        break;
        RWCASE(1) RWCASE(2) RWCASE(4) RWCASE(8) RWCASE(16) RWCASE(32) RWCASE(64)
#undef RWCASE
    default:
        throw Error(MFIC_ERR_SIZE, "bad row tile");
    }
    MFIC_LAUNCH_CHECK();
}

template <typename T, bool R2C>
static void launch_rows(mfic_ctx ctx, T* data, size_t nrows, int n2, double scale) {
    MFIC_CHECK(n2 % 2 == 0, MFIC_ERR_SIZE, "contiguous axis must be even");
    const int M = n2 / 2;
    const long pitch = n2 + 2;
    switch (M) {
#define X(S) case S: launch_rows_M<T, S, R2C>(ctx, data, nrows, pitch, (T)scale); break;
        MFIC_FOR_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported contiguous-axis length " + std::to_string(n2));
    }
}

template <typename T> void fft_rows_r2c(mfic_ctx ctx, T* data, size_t nrows, int n2, double scale) {
    launch_rows<T, true>(ctx, data, nrows, n2, scale);
}
template <typename T> void fft_rows_c2r(mfic_ctx ctx, T* data, size_t nrows, int n2, double scale) {
    launch_rows<T, false>(ctx, data, nrows, n2, scale);
}

static void check_sizes(const size_t n[3]) {
    MFIC_CHECK(fft_length_supported(n[0]) && fft_length_supported(n[1]), MFIC_ERR_SIZE,
               "unsupported FFT size (supported: 4 6 8 9 12 16 18 24 32 36 48 64 72 96 128 144 192 256 288 384 512 576 768)");
    MFIC_CHECK(n[2] % 2 == 0 && fft_length_supported(n[2] / 2) && n[2] <= 1536, MFIC_ERR_SIZE,
               "unsupported contiguous FFT size (need even, n/2 = 2^a 3^b)");
}

template <typename T> void fft_r2c(mfic_ctx ctx, T* data, const size_t n[3], double scale) {
    check_sizes(n);
    const long nc = (long)(n[2] / 2 + 1);
    fft_rows_r2c<T>(ctx, data, n[0] * n[1], (int)n[2], 1.0);
    fft_pencils<T>(ctx, data, (int)n[1], nc, (long)n[1] * nc, (int)nc, (long)n[0], true, 1.0, "_ax1");
    // axis 0: the (j, k) plane is contiguous, so tiles run over it flat
    fft_pencils<T>(ctx, data, (int)n[0], (long)n[1] * nc, 0, (int)(n[1] * nc), 1, true, scale, "_ax0");
}

template <typename T> void fft_c2r(mfic_ctx ctx, T* data, const size_t n[3], double scale) {
    check_sizes(n);
    const long nc = (long)(n[2] / 2 + 1);
    fft_pencils<T>(ctx, data, (int)n[0], (long)n[1] * nc, 0, (int)(n[1] * nc), 1, false, 1.0, "_ax0");
    fft_pencils<T>(ctx, data, (int)n[1], nc, (long)n[1] * nc, (int)nc, (long)n[0], false, 1.0, "_ax1");
    fft_rows_c2r<T>(ctx, data, n[0] * n[1], (int)n[2], scale);
}

// the two strided c2r passes only (axis 0, axis 1): the contiguous-axis pass
// is left to a fused row kernel
template <typename T, int L, int V, bool GEN>
static void launch_ax0_shift_v(mfic_ctx ctx, cx<T>* data, cx<T>* shifted, int n1, int nc, int3 wa, int3 wb) {
    constexpr int P = Plan<L, V>::P;
    // at most 16 pencils per tile by default: at 512^3 FP32 the 16-pencil
    // tile takes 0.46 ms per launch, the 32-pencil tile 1.37 ms (two outputs
    // per pencil double the registers of the 1024-thread workgroups)
    int tk = 16;
    // two exchange regions (both transforms in lockstep)
    while (tk > 4 && (2 * (size_t)L * tk * pencil_lds_elem<T, true>() > pencil_lds_cap<true>() || P * tk > 1024)) tk /= 2;
    // its own tuning key (kind 2): this pass keeps two outputs per pencil, so
    // a tile tuned for the plain pencils (kind 0) does not carry over
    const auto ov = ctx->tiles.find(std::make_tuple(L, 2, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32));
    if (ov != ctx->tiles.end()) tk = ov->second.first;
    while (tk > 1 && (2 * (size_t)L * tk * pencil_lds_elem<T, true>() > max_lds(ctx) || P * tk > 1024)) tk /= 2;
    if (V != 0 && tk < 4) tk = 4;  // alternative plans are built for tiles of 4 and more
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, L, 0, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32, V));
    const cx<double>* half = static_cast<const cx<double>*>(twiddles(ctx, L, 2, MFIC_F64));
    const int tlen = n1 * nc;
    dim3 grid((unsigned)((tlen + tk - 1) / tk));
    const size_t lds = 2 * (size_t)L * tk * pencil_lds_elem<T, true>();
    ProfScope prof(ctx, "fft_pencils_L" + std::to_string(L) + "_inv_ax0_shift" + (sizeof(T) == 8 ? "_f64" : "_f32"),
                   3.0 * (double)tlen * L * sizeof(cx<T>));
    switch (tk) {
#define TKCASE(K)                                                                                               \
    case K:                                                                                                     \
        if constexpr (P * K <= 1024 && (V == 0 || K >= 4))                                                      \
            hipLaunchKernelGGL((k_pencils_ax0_shift<T, L, K, V, GEN>), grid, dim3(P * K), lds, ctx->stream,     \
                               data, shifted, n1, nc, (T)1, tw, half, K, wa, wb);                               \
        break;
        TKCASE(1) TKCASE(2) TKCASE(4) TKCASE(8) TKCASE(16) TKCASE(32) TKCASE(64)
#undef TKCASE
    default:
        throw Error(MFIC_ERR_SIZE, "bad tile");
    }
    MFIC_LAUNCH_CHECK();
}

template <typename T, int L>
static void launch_ax0_shift(mfic_ctx ctx, cx<T>* data, cx<T>* shifted, int n1, int nc, const int* wa,
                             const int* wb) {
    const int3 a = wa ? make_int3(wa[0], wa[1], wa[2]) : make_int3(0, 0, 0);
    const int3 b = wb ? make_int3(wb[0], wb[1], wb[2]) : make_int3(1, 1, 1);
    const bool gen = wa || wb;
    if (plan_variant(ctx, L, MFIC_PASS_AX0_SHIFT, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32) == 1) {
        if (gen) launch_ax0_shift_v<T, L, alt_plan<L>(), true>(ctx, data, shifted, n1, nc, a, b);
        else launch_ax0_shift_v<T, L, alt_plan<L>(), false>(ctx, data, shifted, n1, nc, a, b);
    } else {
        if (gen) launch_ax0_shift_v<T, L, 0, true>(ctx, data, shifted, n1, nc, a, b);
        else launch_ax0_shift_v<T, L, 0, false>(ctx, data, shifted, n1, nc, a, b);
    }
}

template <typename T>
void fft_c2r_cols_shift(mfic_ctx ctx, T* data, T* shifted, const size_t n[3], const int* wa, const int* wb) {
    check_sizes(n);
    for (int a = 0; a < 3; ++a)
        MFIC_CHECK((!wa || (wa[a] >= 0 && wa[a] <= 2)) && (!wb || (wb[a] >= 0 && wb[a] <= 2)), MFIC_ERR_ARG,
                   "sub-lattice shift weights must be 0, 1 or 2 half cells");
    MFIC_CHECK(n[0] == n[1] && n[1] == n[2], MFIC_ERR_SIZE, "the shifted sub-lattice pass needs a cubic grid");
    const long nc = (long)(n[2] / 2 + 1);
    switch ((int)n[0]) {
#define X(S)                                                                                                      \
    case S:                                                                                                       \
        launch_ax0_shift<T, S>(ctx, reinterpret_cast<cx<T>*>(data), reinterpret_cast<cx<T>*>(shifted), (int)n[1], \
                               (int)nc, wa, wb);                                                                  \
        break;
        MFIC_FOR_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported FFT length " + std::to_string(n[0]));
    }
    for (T* d : {data, shifted})
        fft_pencils<T>(ctx, d, (int)n[1], nc, (long)n[1] * nc, (int)nc, (long)n[0], false, 1.0, "_ax1");
}
template void fft_c2r_cols_shift<float>(mfic_ctx, float*, float*, const size_t*, const int*, const int*);
template void fft_c2r_cols_shift<double>(mfic_ctx, double*, double*, const size_t*, const int*, const int*);

template <typename T> void fft_c2r_cols(mfic_ctx ctx, T* data, const size_t n[3]) {
    check_sizes(n);
    const long nc = (long)(n[2] / 2 + 1);
    fft_pencils<T>(ctx, data, (int)n[0], (long)n[1] * nc, 0, (int)(n[1] * nc), 1, false, 1.0, "_ax0");
    fft_pencils<T>(ctx, data, (int)n[1], nc, (long)n[1] * nc, (int)nc, (long)n[0], false, 1.0, "_ax1");
}
template void fft_c2r_cols<float>(mfic_ctx, float*, const size_t*);
template void fft_c2r_cols<double>(mfic_ctx, double*, const size_t*);

template void fft_r2c<float>(mfic_ctx, float*, const size_t*, double);
template void fft_r2c<double>(mfic_ctx, double*, const size_t*, double);
template void fft_c2r<float>(mfic_ctx, float*, const size_t*, double);
template void fft_c2r<double>(mfic_ctx, double*, const size_t*, double);
template void fft_rows_r2c<float>(mfic_ctx, float*, size_t, int, double);
template void fft_rows_r2c<double>(mfic_ctx, double*, size_t, int, double);
template void fft_rows_c2r<float>(mfic_ctx, float*, size_t, int, double);
template void fft_rows_c2r<double>(mfic_ctx, double*, size_t, int, double);
template void fft_pencils<float>(mfic_ctx, void*, int, long, long, int, long, bool, double, const char*);
template void fft_pencils<double>(mfic_ctx, void*, int, long, long, int, long, bool, double, const char*);

}  // namespace mfic
