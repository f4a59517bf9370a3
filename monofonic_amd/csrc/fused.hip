// Fused padded row pass for the LPT sources (ic_generator.cc:523-629 with
// the Orszag 3/2 rule of convolution.hh:495-532).
//
// One workgroup takes one or a few padded rows (i', j').  Each input field's
// compact half-spectrum row (kz <= N/2, after the axis-0 and axis-1 c2r
// passes) is brought to real space in LDS by its own thread group, the
// source products are formed pointwise from LDS with FP64 sums, and every
// output row is transformed back (r2c) and written as a compact
// half-spectrum row for the axis-1 / axis-0 r2c passes.  The padded real
// fields therefore never reach HBM: per row the pass reads NI and writes NO
// compact rows.
//
// Each transform's threads lie inside one wavefront (groups start on
// power-of-two lane boundaries), so Stockham stages synchronise per
// wavefront and idle output groups simply skip the r2c; workgroup barriers
// separate the three phases.  Two forms: one half-length transform per row
// (k_rows_fused) or two rows per full-length complex transform
// (k_rows_fused_pair, the default), chosen per length and precision at run
// time.
#include <cmath>

#include "fft_core.hpp"
#include "fused.hpp"
#include "mfic_internal.hpp"

namespace mfic {

// Hessian component order in every source: xx yy zz xy xz yz
// 2LPT source (ic_generator.cc:536-541)
struct ProdSrc2 {
    static constexpr int NI = 6, NO = 1, IPG = 1, WAVES = 3;
    static constexpr bool OUT32 = false;
    __device__ static __forceinline__ void eval(const double* h, double* o) {
        o[0] = h[0] * (h[1] + h[2]) + h[1] * h[2] - h[3] * h[3] - h[4] * h[4] - h[5] * h[5];
    }
};
// 2LPT source plus the inner products of the 3a terms grouped by their outer
// factor: Qzz = xx yy - xy^2, Qyz = xy xz, Qxx = yz^2, Qyy = xz^2
// (ic_generator.cc:581-585)
struct ProdSrc2Q {
    static constexpr int NI = 6, NO = 5, IPG = 1, WAVES = 3;
    static constexpr bool OUT32 = false;
    __device__ static __forceinline__ void eval(const double* h, double* o) {
        ProdSrc2::eval(h, o);
        o[1] = h[0] * h[1] - h[3] * h[3];
        o[2] = h[3] * h[4];
        o[3] = h[5] * h[5];
        o[4] = h[4] * h[4];
    }
};
// phi3 source (3a outer products with the truncated Q fields + 3b,
// ic_generator.cc:581-609) and A3 (ic_generator.cc:614-627).  Inputs: phi1
// Hessians a[0..5], phi2 Hessians b[6..11], Qzz Qyz Qxx Qyy [12..15]
struct ProdSrc3 {
    // two inputs per lane group: 4-wave workgroups; registers for two waves
    // per SIMD (no spills), i.e. two workgroups per CU
    static constexpr int NI = 16, NO = 4, IPG = 1, WAVES = 2;
    // the four output transforms on 32 threads each (two waves) instead of
    // 16 (one wave): the output phase was a third of the pass
    static constexpr bool OUT32 = true;
    __device__ static __forceinline__ void eval(const double* h, double* o) {
        const double* a = h;
        const double* b = h + 6;
        const double* Q = h + 12;
        const double c5 = -5.0 / 7.0, c10 = 10.0 / 7.0;
        o[0] = Q[0] * a[2] + 2.0 * (Q[1] * a[5]) - Q[2] * a[0] - Q[3] * a[1] +
               c5 * (a[0] * (b[1] + b[2]) + a[1] * (b[2] + b[0]) + a[2] * (b[0] + b[1])) +
               c10 * (a[3] * b[3] + a[4] * b[4] + a[5] * b[5]);
        // A3[d] = b_{dd'} a_{dd''} - b_{dd''} a_{dd'} + a_{d'd''} (b_{d'd'} - b_{d''d''})
        //         - b_{d'd''} (a_{d'd'} - a_{d''d''}),  d' = d+1, d'' = d+2 (mod 3)
        o[1] = b[3] * a[4] - b[4] * a[3] + a[5] * (b[1] - b[2]) - b[5] * (a[1] - a[2]);
        o[2] = b[5] * a[3] - b[3] * a[5] + a[4] * (b[2] - b[0]) - b[4] * (a[2] - a[0]);
        o[3] = b[4] * a[5] - b[5] * a[4] + a[3] * (b[0] - b[1]) - b[3] * (a[0] - a[1]);
    }
};

template <int M, int NI> constexpr int fused_rows_per_block() {
    constexpr int tpr = NI * lane_slot<Plan<M>::P>();
    return tpr >= 256 ? 1 : (256 / tpr > 8 ? 8 : 256 / tpr);
}

// FP32: at least PR::WAVES waves per SIMD (register budget 512 / WAVES)
template <typename T, int M, class PR>
__global__ void __launch_bounds__((fused_rows_per_block<M, PR::NI / PR::IPG>() * (PR::NI / PR::IPG) * lane_slot<Plan<M>::P>()))
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? PR::WAVES : 1, 8)))
k_rows_fused(const cx<T>* __restrict__ in, long istride, cx<T>* __restrict__ out, long ostride, int kc, long nrows,
             T scale, const cx<T>* __restrict__ tw, const cx<T>* __restrict__ tw2) {
    constexpr int NI = PR::NI, NO = PR::NO, IPG = PR::IPG, NG = NI / IPG;
    constexpr int E = Plan<M>::E, P = Plan<M>::P, SL = lane_slot<P>();
    constexpr int RP = row_lds_pitch<M>(), RB = fused_rows_per_block<M, NG>(), TPR = NG * SL;
    constexpr int NPOS = (M + TPR - 1) / TPR;
    static_assert(P <= 64 && 64 % SL == 0, "a transform must fit in one wavefront");
    static_assert(NI % IPG == 0, "inputs per group must divide the inputs");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int tr = threadIdx.x % TPR, rb = threadIdx.x / TPR;
    const int g = tr / SL, q = tr % SL;
    const bool lane = q < P;
    const long row = xcd_tile(blockIdx.x, gridDim.x) * RB + rb;
    const bool active = row < nrows;
    const cx<T>* uw = tw + plan_tw_size<M>();

    // 1. c2r of inputs g*IPG .. g*IPG+IPG-1: half-spectrum rows (kz < kc, kz = 0
    // and the padded Nyquist real) -> M complex = 2M real values, each kept in
    // its input's region; every row load is issued before the first transform
    if (lane) {
        // kc = N/2 + 1 with N = 4M/3: positions below KC carry data, the rest
        // of the padded row is the zero band
        constexpr int KC = 2 * M / 3 + 1;
        cx<T> raw[IPG][E];
#pragma unroll
        for (int t = 0; t < IPG; ++t) {
            const cx<T>* src = in + (g * IPG + t) * istride + (active ? row : 0) * (long)KC;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = q + P * e;
                if (P * e + P <= KC) raw[t][e] = src[k];
                else if (P * e < KC) raw[t][e] = k < KC ? src[k] : cx<T>{0, 0};
                else raw[t][e] = cx<T>{0, 0};
            }
        }
#pragma unroll
        for (int t = 0; t < IPG; ++t) {
            const LdsRowMap lm{(rb * NI + g * IPG + t) * RP};
#pragma unroll
            for (int e = 0; e < E; ++e) {
                cx<T> x = raw[t][e];
                if (q + P * e == 0) x.y = 0;
                lds[lm.at(q + P * e)] = x;
            }
            if (q == 0) lds[lm.at(M)] = cx<T>{0, 0};
            WaveSync::sync();
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
            WaveSync::sync();
            fft_regs_row<T, M, false, WaveSync>(v, q, lds, lm.base, tw);
#pragma unroll
            for (int e = 0; e < E; ++e) lds[lm.at(q + P * e)] = cscale(v[e], scale);
        }
    }
    __syncthreads();

    // 2. source products at each complex slot (two real points), FP64 sums
    // rounded once to T.  Slot p of every row is read and written by one
    // thread only, so outputs overwrite the first NO input rows in place.
#pragma unroll
    for (int u = 0; u < NPOS; ++u) {
        const int p = tr + u * TPR;
        if (p < M) {
            cx<T> x[NI];
#pragma unroll
            for (int f = 0; f < NI; ++f) x[f] = lds[(rb * NI + f) * RP + p];
            T res[NO][2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                double h[NI], o[NO];
#pragma unroll
                for (int f = 0; f < NI; ++f) h[f] = (double)(c ? x[f].y : x[f].x);
                PR::eval(h, o);
#pragma unroll
                for (int r = 0; r < NO; ++r) res[r][c] = (T)o[r];
            }
#pragma unroll
            for (int o = 0; o < NO; ++o) lds[(rb * NI + o) * RP + p] = cx<T>{res[o][0], res[o][1]};
        }
    }
    __syncthreads();

    // 3. r2c of outputs g, g + NG, ... < NO, first kc half-spectrum values written
    constexpr int RV = row_alt_plan<M>();
    constexpr bool O32 = PR::OUT32 && RV == 2 && RB == 1 && NO * Plan<M, RV>::P <= NG * SL;
    if constexpr (O32) {
        constexpr int E2 = Plan<M, RV>::E, P2 = Plan<M, RV>::P;
        const int o = tr / P2, q2 = tr % P2;
        const cx<T>* uw2 = tw2 + plan_tw_size<M, RV>();
        if (o < NO) {
            const int base = (rb * NI + o) * RP;
            cx<T> v[E2];
#pragma unroll
            for (int e = 0; e < E2; ++e) v[e] = lds[base + q2 + P2 * e];
            WaveSync::sync();
            fft_regs_row<T, M, true, WaveSync, RV>(v, q2, lds, base, tw2);
#pragma unroll
            for (int e = 0; e < E2; ++e) lds[base + q2 + P2 * e] = v[e];
            WaveSync::sync();
            if (active) {
                cx<T>* dst = out + o * ostride + row * (long)kc;
#pragma unroll
                for (int e = 0; e < E2; ++e) {
                    const int k = q2 + P2 * e;
                    if (k >= kc) continue;
                    const cx<T> zk = v[e], zm = lds[base + (k == 0 ? 0 : M - k)];
                    const cx<T> Ek = {(zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5};
                    const cx<T> Ok = {(zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5};
                    dst[k] = cadd(Ek, cmul(uw2[k], Ok));
                }
            }
        }
    } else if (lane) {
        for (int o = g; o < NO; o += NG) {
            const LdsRowMap lm{(rb * NI + o) * RP};
            cx<T> v[E];
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = lds[lm.at(q + P * e)];
            WaveSync::sync();
            fft_regs_row<T, M, true, WaveSync>(v, q, lds, lm.base, tw);
#pragma unroll
            for (int e = 0; e < E; ++e) lds[lm.at(q + P * e)] = v[e];
            WaveSync::sync();
            if (active) {
                cx<T>* dst = out + o * ostride + row * (long)kc;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int k = q + P * e;
                    if (k >= kc) continue;
                    const cx<T> zk = v[e], zm = lds[lm.at(k == 0 ? 0 : M - k)];
                    const cx<T> Ek = {(zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5};
                    const cx<T> Ok = {(zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5};
                    dst[k] = cadd(Ek, cmul(uw[k], Ok));
                }
            }
        }
    }
}

// Paired form of the fused row pass: two real rows per complex transform of
// the full padded length L2 = 2M.  Inputs 2g and 2g+1 of a row are brought
// to real space together, z = x_2g + i x_2g+1, from the Hermitian spectrum
// Z[k] = X_2g[k] + i X_2g+1[k] (k <= N/2) and conj(X_2g[L2-k]) +
// i conj(X_2g+1[L2-k]) (k >= L2 - N/2), read straight from the compact rows:
// no half-length untangle, no LDS round trip before the transform.  Outputs
// are paired the same way, w = o_2p + i o_2p+1, and separated after the
// forward transform as O_2p[k] = W[k] + conj(W[L2-k]),
// O_2p+1[k] = -i (W[k] - conj(W[L2-k])) (the 1/2 and both normalisations
// are one FP64 factor on the products).
// input transforms on the 32-lane plan; FP64 on the 64-lane output plan
// (twice the wavefronts per row: the FP64 passes hold one or two workgroups
// per CU, so their latency wants the waves)
template <typename T, int L2> constexpr int pair_in_variant() {
    return sizeof(T) == 8 && pair_out_plan<L2>() != 0 ? pair_out_plan<L2>() : pair_in_plan<L2>();
}
template <typename T, int L2, int GI> constexpr int pair_rows_per_block() {
    constexpr int tpr = GI * lane_slot<Plan<L2, pair_in_variant<T, L2>()>::P>();
    return tpr >= 256 ? 1 : (256 / tpr > 8 ? 8 : 256 / tpr);
}
template <typename T, int L2, class PR> constexpr int pair_block_threads() {
    constexpr int GI = PR::NI / 2;
    return pair_rows_per_block<T, L2, GI>() * GI * lane_slot<Plan<L2, pair_in_variant<T, L2>()>::P>();
}
// output transforms on 64 lanes (plan pair_out_plan) where every output
// pair of the block gets a wavefront, else on the input transforms' lanes
template <typename T, int L2, class PR> constexpr int pair_out_variant() {
    constexpr int GO = (PR::NO + 1) / 2, RB = pair_rows_per_block<T, L2, PR::NI / 2>();
    constexpr int VO = pair_out_plan<L2>();
    return VO != 0 && RB * GO * lane_slot<Plan<L2, VO>::P>() <= pair_block_threads<T, L2, PR>() ? VO
                                                                                               : pair_in_variant<T, L2>();
}

// LDS slots per pair row: the first pitch of at least L2 that is 32 modulo
// 64, so the two rows a wavefront covers (two 32-lane groups) fall on
// opposite halves of the 64 bank slots
template <int L2> constexpr int pair_row_pitch() { return L2 + ((32 - L2 % 64) % 64 + 64) % 64; }

template <typename T, int L2, class PR>
__global__ void __launch_bounds__((pair_block_threads<T, L2, PR>()))
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? PR::WAVES : 1, 8)))
k_rows_fused_pair(const cx<T>* __restrict__ in, long istride, cx<T>* __restrict__ out, long ostride, long nrows,
                  double scale2, const cx<T>* __restrict__ tw, const cx<T>* __restrict__ two) {
    constexpr int NI = PR::NI, NO = PR::NO, GI = NI / 2, GO = (NO + 1) / 2;
    constexpr int VI = pair_in_variant<T, L2>(), VO = pair_out_variant<T, L2, PR>();
    constexpr int E = Plan<L2, VI>::E, P = Plan<L2, VI>::P, SL = lane_slot<P>();
    constexpr int EO = Plan<L2, VO>::E, PO = Plan<L2, VO>::P, SLO = lane_slot<PO>();
    constexpr int RB = pair_rows_per_block<T, L2, GI>(), TPR = GI * SL, RP = pair_row_pitch<L2>();
    constexpr int KC = L2 / 3 + 1;  // N/2 + 1 values per compact row, N = 2 L2 / 3
    constexpr int NPOS = (L2 + TPR - 1) / TPR;
    static_assert(NI % 2 == 0, "inputs come in pairs");
    static_assert(P <= 64 && PO <= 64 && 64 % SL == 0 && 64 % SLO == 0, "a transform must fit in one wavefront");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int tr = threadIdx.x % TPR, rb = threadIdx.x / TPR;
    const int g = tr / SL, q = tr % SL;
    const long tile = xcd_tile(blockIdx.x, gridDim.x);
    const long row = tile * RB + rb;
    const bool active = row < nrows;

    // 1. inverse transform of input pair g: all compact-row loads first, the
    // zero band (N/2 < k < L2 - N/2) never loaded
    if (q < P) {
        const long r0 = (active ? row : 0) * (long)KC;
        const cx<T>* s0 = in + (2 * g) * istride + r0;
        const cx<T>* s1 = in + (2 * g + 1) * istride + r0;
        cx<T> a[E], b[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int k = q + P * e;
            a[e] = b[e] = cx<T>{0, 0};
            if (P * e + P <= KC) {
                a[e] = s0[k];
                b[e] = s1[k];
            } else if (P * e > L2 - KC) {
                a[e] = s0[L2 - k];
                b[e] = s1[L2 - k];
            } else if (P * e < KC || P * e + P > L2 - KC + 1) {
                if (k < KC) {
                    a[e] = s0[k];
                    b[e] = s1[k];
                } else if (k > L2 - KC) {
                    a[e] = s0[L2 - k];
                    b[e] = s1[L2 - k];
                }
            }
        }
        cx<T> v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int k = q + P * e;
            if (P * e > L2 - KC || (P * e + P > L2 - KC + 1 && k > L2 - KC)) {
                v[e] = cx<T>{a[e].x + b[e].y, b[e].x - a[e].y};  // conj(A) + i conj(B)
            } else {
                if (e == 0 && k == 0) a[e].y = b[e].y = 0;  // the DC imaginary parts are not data
                v[e] = cx<T>{a[e].x - b[e].y, a[e].y + b[e].x};  // A + i B
            }
        }
        const int base = (rb * GI + g) * RP;
        fft_regs_row<T, L2, false, WaveSync, VI>(v, q, lds, base, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) lds[base + q + P * e] = v[e];
    }
    __syncthreads();

    // 2. products at every real point n from the NI/2 pair rows, in FP64,
    // times the normalisation; outputs paired over the first pair rows (a
    // point is read and written by one thread only)
#pragma unroll
    for (int u = 0; u < NPOS; ++u) {
        const int n = tr + u * TPR;
        if (n < L2) {
            double h[NI], o[NO];
#pragma unroll
            for (int p = 0; p < GI; ++p) {
                const cx<T> z = lds[(rb * GI + p) * RP + n];
                h[2 * p] = (double)z.x;
                h[2 * p + 1] = (double)z.y;
            }
            PR::eval(h, o);
#pragma unroll
            for (int p = 0; p < GO; ++p) {
                const T lo = (T)(o[2 * p] * scale2);
                const T hi = 2 * p + 1 < NO ? (T)(o[2 * p + 1] * scale2) : (T)0;
                lds[(rb * GI + p) * RP + n] = cx<T>{lo, hi};
            }
        }
    }
    __syncthreads();

    // 3. forward transform of output pair p, separated into the two outputs'
    // first KC modes
    {
        const int t = threadIdx.x / SLO, qo = threadIdx.x % SLO;
        const int rbo = t / GO, p = t % GO;
        if (t < RB * GO && qo < PO) {
            const int base = (rbo * GI + p) * RP;
            cx<T> v[EO];
#pragma unroll
            for (int e = 0; e < EO; ++e) v[e] = lds[base + qo + PO * e];
            WaveSync::sync();
            fft_regs_row<T, L2, true, WaveSync, VO>(v, qo, lds, base, two);
            // partners W[L2 - k] of the kept modes sit at L2 - KC + 1 .. L2 - 1
#pragma unroll
            for (int e = 0; e < EO; ++e)
                if (PO * e + PO > L2 - KC + 1) lds[base + qo + PO * e] = v[e];
            WaveSync::sync();
            const long rowo = tile * RB + rbo;
            if (rowo < nrows) {
                cx<T>* d0 = out + (2 * p) * ostride + rowo * (long)KC;
                cx<T>* d1 = out + (2 * p + 1 < NO ? 2 * p + 1 : 2 * p) * ostride + rowo * (long)KC;
#pragma unroll
                for (int e = 0; e < EO; ++e) {
                    const int k = qo + PO * e;
                    if (PO * e >= KC || k >= KC) continue;
                    const cx<T> w = v[e], wm = k == 0 ? v[e] : lds[base + L2 - k];
                    d0[k] = cx<T>{w.x + wm.x, w.y - wm.y};
                    if (2 * p + 1 < NO) d1[k] = cx<T>{w.y + wm.y, wm.x - w.x};
                }
            }
        }
    }
}

template <typename T, int L2, class PR>
static void launch_fused_pair(mfic_ctx ctx, const void* in, void* out, long stride, int kc, long nrows, double scale,
                              const char* name) {
    constexpr int NI = PR::NI, NO = PR::NO, GI = NI / 2;
    constexpr int RB = pair_rows_per_block<T, L2, GI>(), VI = pair_in_variant<T, L2>(), VO = pair_out_variant<T, L2, PR>();
    MFIC_CHECK(3 * (kc - 1) == L2, MFIC_ERR_SIZE, "half-spectrum rows must hold N/2 + 1 values of a 3N/2 padded row");
    const size_t lds = (size_t)RB * GI * pair_row_pitch<L2>() * sizeof(cx<T>);
    MFIC_CHECK(lds <= (size_t)ctx->max_lds, MFIC_ERR_SIZE, "fused row pass exceeds the LDS of a workgroup");
    const int prec = sizeof(T) == 8 ? MFIC_F64 : MFIC_F32;
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, L2, 0, prec, VI));
    const cx<T>* two = static_cast<const cx<T>*>(twiddles(ctx, L2, 0, prec, VO));
    dim3 grid((unsigned)xcd_grid((nrows + RB - 1) / RB));
    ProfScope prof(ctx, std::string(name) + "_n" + std::to_string(L2) + (sizeof(T) == 8 ? "_f64" : "_f32"),
                   (double)nrows * (NI + NO) * kc * sizeof(cx<T>));
    hipLaunchKernelGGL((k_rows_fused_pair<T, L2, PR>), grid, dim3(pair_block_threads<T, L2, PR>()), lds, ctx->stream,
                       static_cast<const cx<T>*>(in), stride, static_cast<cx<T>*>(out), stride, nrows,
                       0.5 * scale * scale, tw, two);
    MFIC_LAUNCH_CHECK();
}

template <typename T, int M, class PR>
static void launch_fused(mfic_ctx ctx, const void* in, void* out, long stride, int kc, long nrows, double scale,
                         const char* name) {
    constexpr int NI = PR::NI, NO = PR::NO, NG = NI / PR::IPG, SL = lane_slot<Plan<M>::P>();
    constexpr int RB = fused_rows_per_block<M, NG>(), RP = row_lds_pitch<M>();
    MFIC_CHECK(3 * (kc - 1) == 2 * M, MFIC_ERR_SIZE, "half-spectrum rows must hold N/2 + 1 values of a 3N/2 padded row");
    const size_t lds = (size_t)RB * NI * RP * sizeof(cx<T>);
    MFIC_CHECK(lds <= (size_t)ctx->max_lds, MFIC_ERR_SIZE, "fused row pass exceeds the LDS of a workgroup");
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, M, 1, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32));
    // the output transforms' own plan (row plan 2) where the pass uses one
    const cx<T>* tw2 = (PR::OUT32 && row_alt_plan<M>() == 2 && RB == 1)
                           ? static_cast<const cx<T>*>(twiddles(ctx, M, 1, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32, 2))
                           : tw;
    dim3 grid((unsigned)xcd_grid((nrows + RB - 1) / RB));
    ProfScope prof(ctx, std::string(name) + "_n" + std::to_string(2 * M) + (sizeof(T) == 8 ? "_f64" : "_f32"),
                   (double)nrows * (NI + NO) * kc * sizeof(cx<T>));
    hipLaunchKernelGGL((k_rows_fused<T, M, PR>), grid, dim3(RB * NG * SL), lds, ctx->stream,
                       static_cast<const cx<T>*>(in), stride, static_cast<cx<T>*>(out), stride, kc, nrows, (T)scale,
                       tw, tw2);
    MFIC_LAUNCH_CHECK();
}

// Last c2r pass of the three displacement (or velocity) components fused
// with the lattice readout (particle_generator.hh:499-544, 597-639): each
// group transforms one component row in LDS, then the row's particles are
// written as interleaved FP64 [p][3], q + Psi for positions
// ((ijk + shift)/N * lunit per axis, grid_fft.hh:242-252, not wrapped) or Psi
// for velocities.
template <typename T, int M>
__global__ void __launch_bounds__((fused_rows_per_block<M, 3>() * 3 * lane_slot<Plan<M>::P>()))
k_rows_c2r_particles(const cx<T>* __restrict__ f0, const cx<T>* __restrict__ f1, const cx<T>* __restrict__ f2,
                     long nrows, T scale, const cx<T>* __restrict__ tw, double* __restrict__ dst, int is_pos,
                     double lunit, double3 shift, size_t pbase) {
    constexpr int E = Plan<M>::E, P = Plan<M>::P, SL = lane_slot<P>(), N = 2 * M;
    constexpr int RP = row_lds_pitch<M>(), RB = fused_rows_per_block<M, 3>(), TPR = 3 * SL;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int tr = threadIdx.x % TPR, rb = threadIdx.x / TPR;
    const int g = tr / SL, q = tr % SL;
    const long row = (long)blockIdx.x * RB + rb;
    const bool active = row < nrows;
    const LdsRowMap lm{(rb * 3 + g) * RP};
    const cx<T>* uw = tw + plan_tw_size<M>();
    if (q < P) {
        // the untangle's partner M - k is loaded from global memory beside k
        // (same cache lines), so it runs in registers without an LDS round
        // trip (4.96 -> 4.13 ms for the four 512^3 launches)
        const cx<T>* src = (g == 0 ? f0 : (g == 1 ? f1 : f2)) + (active ? row : 0) * (long)(M + 1);
        cx<T> raw[E], rev[E], v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int k = q + P * e;
            raw[e] = active ? src[k] : cx<T>{0, 0};
            rev[e] = active ? src[M - k] : cx<T>{0, 0};
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int k = q + P * e;
            cx<T> xk = raw[e], xm = rev[e];
            if (k == 0) { xk.y = 0; xm.y = 0; }  // kz = 0 and the Nyquist M are real
            const cx<T> A = {xk.x + xm.x, xk.y - xm.y};
            const cx<T> D = {xk.x - xm.x, xk.y + xm.y};
            const cx<T> B = cmulc(D, uw[k]);
            v[e] = cx<T>{A.x - B.y, A.y + B.x};
        }
        fft_regs_row<T, M, false, WaveSync>(v, q, lds, lm.base, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) lds[lm.at(q + P * e)] = cscale(v[e], scale);
    }
    __syncthreads();
    if (!active) return;
    const int i = (int)(row / N), j = (int)(row % N);
    double* out = dst + 3 * (pbase + (size_t)row * N);
    for (int idx = tr; idx < 3 * N; idx += TPR) {
        const int p = idx / 3, c = idx - 3 * p;
        const cx<T> w = lds[(rb * 3 + c) * RP + (p >> 1)];
        double val = (double)((p & 1) ? w.y : w.x);
        if (is_pos) {
            const int ijk = c == 0 ? i : (c == 1 ? j : p);
            const double sh = c == 0 ? shift.x : (c == 1 ? shift.y : shift.z);
            val = (((double)ijk + sh) / (double)N) * lunit + val;
        }
        out[idx] = val;
    }
}

template <typename T, int M>
static void launch_particles(mfic_ctx ctx, const void* const f[3], int is_pos, double lunit, const double shift[3],
                             size_t pbase, double* dst) {
    constexpr int SL = lane_slot<Plan<M>::P>(), RB = fused_rows_per_block<M, 3>(), RP = row_lds_pitch<M>();
    const long nrows = (long)(2 * M) * (2 * M);
    const size_t lds = (size_t)RB * 3 * RP * sizeof(cx<T>);
    MFIC_CHECK(lds <= (size_t)ctx->max_lds, MFIC_ERR_SIZE, "particle row pass exceeds the LDS of a workgroup");
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, M, 1, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32));
    const double N = 2.0 * M;
    ProfScope prof(ctx, std::string("rows_c2r_particles_n") + std::to_string(2 * M) + (sizeof(T) == 8 ? "_f64" : "_f32"),
                   (double)nrows * (3.0 * (M + 1) * sizeof(cx<T>) + 3.0 * 2 * M * 8.0));
    dim3 grid((unsigned)((nrows + RB - 1) / RB));
    hipLaunchKernelGGL((k_rows_c2r_particles<T, M>), grid, dim3(RB * 3 * SL), lds, ctx->stream,
                       static_cast<const cx<T>*>(f[0]), static_cast<const cx<T>*>(f[1]),
                       static_cast<const cx<T>*>(f[2]), nrows, (T)(1.0 / std::sqrt(N * N * N)), tw, dst, is_pos,
                       lunit, make_double3(shift[0], shift[1], shift[2]), pbase);
    MFIC_LAUNCH_CHECK();
}

#define MFIC_PARTICLE_ROWS(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384)

void rows_c2r_particles(mfic_ctx ctx, int prec, const void* const f[3], int n, int is_pos, double lunit,
                        const double shift[3], size_t pbase, double* dst) {
    switch (n / 2) {
#define X(M)                                                                                  \
    case M:                                                                                   \
        if (prec == MFIC_F64) launch_particles<double, M>(ctx, f, is_pos, lunit, shift, pbase, dst); \
        else launch_particles<float, M>(ctx, f, is_pos, lunit, shift, pbase, dst);            \
        break;
        MFIC_PARTICLE_ROWS(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported particle grid size " + std::to_string(n));
    }
}

#define MFIC_FUSED_SIZES(X) X(12) X(18) X(24) X(36) X(48) X(72) X(96) X(144) X(192) X(288) X(384) X(576) X(768)

template <typename T, class PR>
static void fused_t(mfic_ctx ctx, size_t n, const void* in, void* out, const char* name) {
    const int N = (int)n, np = 3 * N / 2, kc = N / 2 + 1;
    const long stride = (long)np * np * kc;
    const double norm = 1.0 / std::sqrt((double)np * (double)np * (double)np);
    // form 0 (default): two rows per full-length transform; form 1: one
    // half-length transform per row
    const bool paired = plan_variant(ctx, np, MFIC_PASS_FUSED_ROWS, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32) == 0;
    switch (np) {
#define X(S)                                                                              \
    case S:                                                                               \
        if (paired) launch_fused_pair<T, S, PR>(ctx, in, out, stride, kc, (long)np * np, norm, name); \
        else launch_fused<T, S / 2, PR>(ctx, in, out, stride, kc, (long)np * np, norm, name); \
        break;
        MFIC_FUSED_SIZES(X)
#undef X
    default:
        throw Error(MFIC_ERR_SIZE, "unsupported padded size " + std::to_string(np));
    }
}

void fused_source_rows(mfic_ctx ctx, int prec, FusedKind kind, size_t n, const void* in, void* out) {
    const bool f64 = prec == MFIC_F64;
    switch (kind) {
    case FusedKind::Src2:
        f64 ? fused_t<double, ProdSrc2>(ctx, n, in, out, "fused_src2") : fused_t<float, ProdSrc2>(ctx, n, in, out, "fused_src2");
        break;
    case FusedKind::Src2Q:
        f64 ? fused_t<double, ProdSrc2Q>(ctx, n, in, out, "fused_src2q") : fused_t<float, ProdSrc2Q>(ctx, n, in, out, "fused_src2q");
        break;
    case FusedKind::Src3:
        f64 ? fused_t<double, ProdSrc3>(ctx, n, in, out, "fused_src3") : fused_t<float, ProdSrc3>(ctx, n, in, out, "fused_src3");
        break;
    }
}

int fused_inputs(FusedKind kind) { return kind == FusedKind::Src3 ? ProdSrc3::NI : ProdSrc2::NI; }
int fused_outputs(FusedKind kind) {
    return kind == FusedKind::Src3 ? ProdSrc3::NO : kind == FusedKind::Src2Q ? ProdSrc2Q::NO : ProdSrc2::NO;
}

}  // namespace mfic
