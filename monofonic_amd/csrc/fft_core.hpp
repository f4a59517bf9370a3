// Device-side FFT building blocks for gfx950 (CDNA4).
//
// A 1D complex transform of length L is executed by P = L/E cooperating
// threads; each thread keeps E elements in registers (positions q + P*e,
// q = thread index within the transform).  A Stockham autosort stage of radix
// R makes every thread compute E/R radix-R butterflies on its own registers;
// the outputs are exchanged through LDS between stages.  The last stage writes
// straight back to the thread's own positions, so loads and stores of a pass
// touch the same addresses (one global read + one global write per element).
//
// Radix plans are compile-time per length (all lengths 2^a 3^b up to 768);
// every radix of a plan divides E.  The two padded-pass lengths 512 and 768
// carry an alternative plan (twice the elements per thread, half the
// threads per transform) that strided pencil passes can select at run time.  Inter-stage twiddles come from a table
// computed on the host in long double (rounded once to the working type);
// intra-butterfly constants are 48th roots of unity.
#pragma once
#include <hip/hip_runtime.h>

namespace mfic {

template <typename T> struct alignas(2 * sizeof(T)) cx { T x, y; };

template <typename T> __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {  // a * conj(b)
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T> __device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool FWD> __device__ __forceinline__ cx<T> rotq(cx<T> a) {
    return FWD ? cx<T>{a.y, -a.x} : cx<T>{-a.y, a.x};
}

// cos(2 pi a / 48), a = 0..12
__device__ constexpr double kCos48[13] = {
    1.0, 0.9914448613738104111446, 0.9659258262890682867497, 0.9238795325112867561282,
    0.8660254037844386467637, 0.7933533402912351645798, 0.7071067811865475244008,
    0.6087614290087206394161, 0.5, 0.3826834323650897717285, 0.2588190451025207623489,
    0.1305261922200515915484, 0.0};

__device__ __forceinline__ constexpr double cos48(int a) {
    a = ((a % 48) + 48) % 48;
    if (a <= 12) return kCos48[a];
    if (a <= 24) return -kCos48[24 - a];
    if (a <= 36) return -kCos48[a - 24];
    return kCos48[48 - a];
}
__device__ __forceinline__ constexpr double sin48(int a) { return cos48(a - 12); }

// multiply by exp(-+ 2 pi i m / R) (sign by direction), R | 48
template <typename T, int R, bool FWD>
__device__ __forceinline__ cx<T> rot_const(cx<T> v, int m) {
    const int a = (m * (48 / R)) % 48;
    if (a == 0) return v;
    if (a == 24) return {-v.x, -v.y};
    if (a == 12) return rotq<T, FWD>(v);
    if (a == 36) return rotq<T, !FWD>(v);
    const T c = (T)cos48(a);
    const T s = FWD ? -(T)sin48(a) : (T)sin48(a);
    return {v.x * c - v.y * s, v.x * s + v.y * c};
}

// ------------------------------------------------------------ radix DFTs --
template <typename T, int R, bool FWD> struct Dft;

template <typename T, bool FWD> struct Dft<T, 2, FWD> {
    static __device__ __forceinline__ void run(cx<T>* a) {
        cx<T> t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    }
};
template <typename T, bool FWD> struct Dft<T, 3, FWD> {
    static __device__ __forceinline__ void run(cx<T>* a) {
        const T s3 = (T)0.8660254037844386467637;
        cx<T> t1 = cadd(a[1], a[2]), t2 = csub(a[1], a[2]);
        cx<T> m = {a[0].x - (T)0.5 * t1.x, a[0].y - (T)0.5 * t1.y};
        cx<T> n = rotq<T, FWD>(cscale(t2, s3));
        a[0] = cadd(a[0], t1);
        a[1] = cadd(m, n);
        a[2] = csub(m, n);
    }
};
template <typename T, bool FWD> struct Dft<T, 4, FWD> {
    static __device__ __forceinline__ void run(cx<T>* a) {
        cx<T> t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        cx<T> t2 = cadd(a[1], a[3]), t3 = rotq<T, FWD>(csub(a[1], a[3]));
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    }
};
// composite R = R1 * R2 (decimation in time: n = n2 + R2 n1, k = k1 + R1 k2)
template <typename T, int R1, int R2, bool FWD>
__device__ __forceinline__ void dft_composite(cx<T>* a) {
    constexpr int R = R1 * R2;
    cx<T> y[R];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
        cx<T> t[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) t[n1] = a[n2 + R2 * n1];
        Dft<T, R1, FWD>::run(t);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) y[n2 * R1 + k1] = rot_const<T, R, FWD>(t[k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
        cx<T> t[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) t[n2] = y[n2 * R1 + k1];
        Dft<T, R2, FWD>::run(t);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) a[k1 + R1 * k2] = t[k2];
    }
}
template <typename T, bool FWD> struct Dft<T, 6, FWD> {
    static __device__ __forceinline__ void run(cx<T>* a) { dft_composite<T, 3, 2, FWD>(a); }
};
template <typename T, bool FWD> struct Dft<T, 8, FWD> {
    static __device__ __forceinline__ void run(cx<T>* a) { dft_composite<T, 2, 4, FWD>(a); }
};
template <typename T, bool FWD> struct Dft<T, 12, FWD> {
    static __device__ __forceinline__ void run(cx<T>* a) { dft_composite<T, 4, 3, FWD>(a); }
};
template <typename T, bool FWD> struct Dft<T, 16, FWD> {
    static __device__ __forceinline__ void run(cx<T>* a) { dft_composite<T, 4, 4, FWD>(a); }
};
template <typename T, bool FWD> struct Dft<T, 24, FWD> {
    static __device__ __forceinline__ void run(cx<T>* a) { dft_composite<T, 8, 3, FWD>(a); }
};

// ------------------------------------------------------------ size plans --
// radices r0 r1 r2 r3 (0 = unused), E = elements per thread; V numbers the
// alternative plans of a length (0 = the default)
template <int L, int V = 0> struct Plan;
#define MFIC_PLAN_V(L_, V_, E_, S_, A, B, C, D)                           \
    template <> struct Plan<L_, V_> {                                      \
        static constexpr int L = L_, V = V_, E = E_, S = S_, P = L_ / E_;  \
        static constexpr int r(int s) { return s == 0 ? A : s == 1 ? B : s == 2 ? C : D; } \
    };
#define MFIC_PLAN(L_, E_, S_, A, B, C, D) MFIC_PLAN_V(L_, 0, E_, S_, A, B, C, D)
MFIC_PLAN(4, 4, 1, 4, 0, 0, 0)
MFIC_PLAN(6, 6, 1, 6, 0, 0, 0)
MFIC_PLAN(8, 8, 1, 8, 0, 0, 0)
MFIC_PLAN(9, 3, 2, 3, 3, 0, 0)
MFIC_PLAN(12, 12, 1, 12, 0, 0, 0)
MFIC_PLAN(16, 16, 1, 16, 0, 0, 0)
MFIC_PLAN(18, 6, 2, 3, 6, 0, 0)
MFIC_PLAN(24, 12, 2, 2, 12, 0, 0)
MFIC_PLAN(32, 8, 2, 4, 8, 0, 0)
MFIC_PLAN(36, 6, 2, 6, 6, 0, 0)
MFIC_PLAN(48, 12, 2, 4, 12, 0, 0)
MFIC_PLAN(64, 8, 2, 8, 8, 0, 0)
MFIC_PLAN(72, 12, 2, 6, 12, 0, 0)
MFIC_PLAN(96, 12, 3, 4, 4, 6, 0)
MFIC_PLAN(128, 16, 2, 8, 16, 0, 0)
MFIC_PLAN(144, 12, 2, 12, 12, 0, 0)
MFIC_PLAN(192, 12, 3, 4, 4, 12, 0)
MFIC_PLAN(256, 16, 2, 16, 16, 0, 0)
MFIC_PLAN(288, 12, 3, 2, 12, 12, 0)
MFIC_PLAN(384, 24, 3, 24, 8, 2, 0)
MFIC_PLAN(512, 16, 3, 16, 8, 4, 0)
MFIC_PLAN(576, 24, 3, 8, 6, 12, 0)
MFIC_PLAN(768, 24, 3, 12, 8, 8, 0)
// alternatives of the padded-pass lengths: 16 threads per transform
MFIC_PLAN_V(512, 1, 32, 3, 16, 8, 4, 0)
MFIC_PLAN_V(768, 1, 48, 3, 16, 12, 4, 0)
// row plan (V 2) for the output transforms of the 16-input fused pass: 32
// threads per 384-point row
MFIC_PLAN_V(384, 2, 12, 4, 12, 4, 4, 2)
// row plans of the paired fused passes (two real rows per 768-point complex
// transform): V 3 on 32 threads for the inputs, V 2 on 64 threads for the
// outputs.  Radix orders with power-of-two spans before the last stage
// (shift-and-mask twiddle indices) whose exchanges have conflict-free
// swizzles.
MFIC_PLAN_V(768, 2, 12, 4, 4, 4, 4, 12)
MFIC_PLAN_V(768, 3, 24, 3, 8, 8, 12, 0)
#undef MFIC_PLAN
#undef MFIC_PLAN_V
// number of radix plans compiled for a length
template <int L> constexpr int plan_variants() { return (L == 512 || L == 768) ? 2 : 1; }
// plan 1 of a length that has one, else plan 0 (lets run-time dispatch name
// plan 1 for every length without instantiating a missing plan)
template <int L> constexpr int alt_plan() { return plan_variants<L>() > 1 ? 1 : 0; }
// the row plan 2 of a length that has one, else plan 0
template <int L> constexpr int row_alt_plan() { return L == 384 ? 2 : 0; }
// whether plan V of length L is compiled (run-time form for table checks)
__host__ __device__ constexpr bool plan_exists(int L, int V) {
    return V == 0 || (V == 1 && (L == 512 || L == 768)) || (V == 2 && (L == 384 || L == 768)) || (V == 3 && L == 768);
}
// plan V of L if compiled, else plan 0 (compile-time dispatch helper)
template <int L, int V> constexpr int plan_or_default() { return plan_exists(L, V) ? V : 0; }
// plans of the paired fused passes: input transforms (one group of lanes
// per pair of input rows) and output transforms (one per pair of outputs)
template <int L> constexpr int pair_in_plan() { return L == 768 ? 3 : 0; }
template <int L> constexpr int pair_out_plan() { return L == 768 ? 2 : 0; }

template <int L, int V = 0> constexpr int plan_ns(int s) {  // span before stage s
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= Plan<L, V>::r(i);
    return ns;
}
template <int L, int V = 0> constexpr int plan_tw_offset(int s) {  // twiddle table offset
    int off = 0;
    for (int i = 0; i < s; ++i) off += plan_ns<L, V>(i) * (Plan<L, V>::r(i) - 1);
    return off;
}
template <int L, int V = 0> constexpr int plan_tw_size() { return plan_tw_offset<L, V>(Plan<L, V>::S); }

// LDS addressing for a thread's transform: address(pos) = base + pos * stride
struct LdsMap {
    int base, stride;
    __device__ __forceinline__ int at(int pos) const { return base + pos * stride; }
};

// XCD-aware tile order: workgroups are dispatched round-robin over the 8
// XCDs (each with its own L2), so workgroup b runs on XCD b % 8.  Giving
// each XCD a contiguous range of tiles keeps neighbouring tiles, which share
// cache lines at their edges, in one L2.  nb must be a multiple of 8; tiles
// past the real count come out inactive in the kernels.
__device__ __forceinline__ long xcd_tile(long b, long nb) { return (b % 8) * (nb / 8) + b / 8; }
__host__ __device__ constexpr long xcd_grid(long nb) { return (nb + 7) / 8 * 8; }

// LDS addressing for a transform whose threads all work on one row: one pad
// slot every 16 elements spreads the stride-R butterfly writes of the early
// Stockham stages over the LDS banks
struct LdsRowMap {
    int base;
    __device__ __forceinline__ int at(int pos) const { return base + pos; }
};
// power-of-two lane slot of a transform of P threads
template <int P> constexpr int lane_slot() { return P <= 1 ? 1 : P <= 2 ? 2 : P <= 4 ? 4 : P <= 8 ? 8 : P <= 16 ? 16 : P <= 32 ? 32 : 64; }
// LDS elements between consecutive rows of M + 1 complex values.  Rows
// whose transforms share a 32-lane group sit lane_slot banks (8-byte slots
// modulo 32) apart, so contiguous accesses of neighbouring rows interleave
// without bank conflicts.
template <int M> constexpr int row_lds_pitch() {
    constexpr int SL = lane_slot<Plan<M>::P>();
    return SL >= 32 ? M + 1 : M + 1 + (((SL - (M + 1)) % 32 + 32) % 32);
}

// Per-stage LDS swizzle of a row transform's Stockham exchanges: output x of
// stage S is stored at base + swz(x), a bijection of [0, L) that flips the
// low `a` bits of x by a multiple of its high bits,
//   swz(x) = x ^ (((x >> a) * k1 ^ (x >> b) * k2) & (2^a - 1)),
// chosen per plan so that every strided write and contiguous read of the
// stage hits distinct 8-byte bank slots within each 16- and 32-lane group
// (exhaustive search over a, b, k1, k2 against the access pattern of the
// plan).  kind 0: identity; kind 2: rotation instead of xor.
template <int L, int S, int V = 0> struct RowSwz { static constexpr int kind = 0, a = 0, k1 = 0, b = 0, k2 = 0; };
#define MFIC_SWZ_V(L_, S_, V_, KIND, A, K1, B, K2) \
    template <> struct RowSwz<L_, S_, V_> { static constexpr int kind = KIND, a = A, k1 = K1, b = B, k2 = K2; };
#define MFIC_SWZ(L_, S_, KIND, A, K1, B, K2) MFIC_SWZ_V(L_, S_, 0, KIND, A, K1, B, K2)
MFIC_SWZ(24, 0, 1, 1, 1, 1, 0)
MFIC_SWZ(32, 0, 1, 2, 1, 2, 0)
MFIC_SWZ(36, 0, 1, 1, 0, 3, 1)
MFIC_SWZ(48, 0, 1, 2, 1, 2, 0)
MFIC_SWZ(64, 0, 1, 3, 1, 3, 0)
MFIC_SWZ(72, 0, 1, 1, 0, 3, 1)
MFIC_SWZ(96, 0, 1, 2, 0, 3, 1)
MFIC_SWZ(96, 1, 1, 3, 0, 4, 4)
MFIC_SWZ(128, 0, 1, 3, 1, 3, 0)
MFIC_SWZ(144, 0, 1, 2, 0, 4, 1)
MFIC_SWZ(192, 0, 1, 2, 0, 4, 1)
MFIC_SWZ(192, 1, 1, 4, 4, 4, 0)
MFIC_SWZ(256, 0, 1, 4, 1, 4, 0)
MFIC_SWZ(288, 0, 1, 2, 0, 5, 1)
MFIC_SWZ(288, 1, 1, 3, 0, 4, 7)
MFIC_SWZ(384, 0, 1, 3, 0, 4, 1)
MFIC_SWZ(384, 1, 1, 4, 0, 6, 3)
MFIC_SWZ(512, 0, 1, 4, 1, 5, 1)
MFIC_SWZ(512, 1, 1, 5, 0, 6, 8)
// paired fused-pass plans (one transform per 32 or 64 lanes): searched
// against 64-bit LDS accesses served 16 lanes at a time by 16 8-byte slots
// (equivalently a wavefront by 64 slots), the model that reproduces the
// SQ_LDS_BANK_CONFLICT counts measured on MI355X, with two 32-lane
// transforms per wavefront sitting 800 slots apart
MFIC_SWZ_V(768, 0, 3, 1, 3, 2, 5, 1)
MFIC_SWZ_V(768, 1, 3, 1, 5, 0, 5, 4)
MFIC_SWZ_V(768, 0, 2, 1, 4, 1, 6, 1)
MFIC_SWZ_V(768, 1, 2, 1, 4, 1, 4, 4)
MFIC_SWZ_V(768, 2, 2, 1, 6, 16, 6, 0)
#undef MFIC_SWZ
#undef MFIC_SWZ_V
__host__ __device__ constexpr bool pow2_or_0(int k) { return (k & (k - 1)) == 0; }
template <int L, int S, int V = 0> struct LdsSwzMap {
    using W = RowSwz<L, S, V>;
    static_assert(W::kind == 0 || L % (1 << W::a) == 0, "swizzle must stay inside the row");
    int base;
    __host__ __device__ static constexpr int swz(int x) {
        constexpr int A = (1 << W::a) - 1;
        if constexpr (W::kind == 0) return x;
        else if constexpr (W::kind == 1) return x ^ ((((x >> W::a) * W::k1) ^ ((x >> W::b) * W::k2)) & A);
        else return (x & ~A) | ((x + (x >> W::a) * W::k1 + (x >> W::b) * W::k2) & A);
    }
    // an xor swizzle whose multipliers are 0 or powers of two is linear over
    // GF(2): swz(u + c) = swz(u) ^ swz(c) when u and c share no bits, so an
    // address is one per-thread value xor a compile-time constant
    static constexpr bool LINEAR = W::kind == 0 || (W::kind == 1 && pow2_or_0(W::k1) && pow2_or_0(W::k2));
    __device__ __forceinline__ int at(int x) const { return base + swz(x); }
};
template <class MAP> struct map_linear { static constexpr bool value = false; };
template <int L, int S, int V> struct map_linear<LdsSwzMap<L, S, V>> {
    static constexpr bool value = LdsSwzMap<L, S, V>::LINEAR;
};

// LDS exchange synchronisation between Stockham stages
struct BlockSync {
    __device__ static __forceinline__ void sync() { __syncthreads(); }
};
// for a transform whose threads all sit in one wavefront: a wavefront's LDS
// operations are performed in issue order, so it suffices to keep the
// compiler from moving memory operations across this point
struct WaveSync {
    __device__ static __forceinline__ void sync() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

// one Stockham stage on registers v[E]; exchanges through LDS unless last.
// SPLIT exchanges the real parts, then the imaginary parts, through an LDS
// array of T: half the LDS of a complex exchange (more resident workgroups
// for the wide pencil tiles) for twice the LDS instructions and barriers.
template <typename T, int L, int S, bool FWD, class MAP, class SYNC = BlockSync, bool SPLIT = false, int V = 0>
__device__ __forceinline__ void fft_stage(cx<T>* v, int q, cx<T>* lds, MAP m, const cx<T>* __restrict__ tw) {
    using PL = Plan<L, V>;
    constexpr int E = PL::E, P = PL::P, R = PL::r(S), B = E / R, NS = plan_ns<L, V>(S);
    constexpr bool LAST = (S == PL::S - 1);
    constexpr bool LINW = map_linear<MAP>::value && pow2_or_0(NS) && pow2_or_0(R);
    constexpr bool LINR = map_linear<MAP>::value && pow2_or_0(P);
    const cx<T>* tws = tw + plan_tw_offset<L, V>(S);
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = q + P * b;
        const int k = j % NS;
        cx<T> a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = v[b + r * B];
        if (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cx<T> w = tws[k * (R - 1) + (r - 1)];
                a[r] = FWD ? cmul(a[r], w) : cmulc(a[r], w);
            }
        }
        Dft<T, R, FWD>::run(a);
        if (LAST || SPLIT) {  // outputs stay on this thread's registers
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + r * B] = a[r];
        } else if constexpr (LINW) {  // r * NS and the thread's part share no bits
            const int fu = MAP::swz((j / NS) * NS * R + k);
#pragma unroll
            for (int r = 0; r < R; ++r) lds[m.base + (fu ^ MAP::swz(r * NS))] = a[r];
        } else {
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[m.at(base + r * NS)] = a[r];
        }
    }
    if constexpr (!LAST && SPLIT) {
        T* f = reinterpret_cast<T*>(lds);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = q + P * b;
            const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
            for (int r = 0; r < R; ++r) f[m.at(base + r * NS)] = v[b + r * B].x;
        }
        SYNC::sync();
#pragma unroll
        for (int e = 0; e < E; ++e) v[e].x = f[m.at(q + P * e)];
        SYNC::sync();
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = q + P * b;
            const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
            for (int r = 0; r < R; ++r) f[m.at(base + r * NS)] = v[b + r * B].y;
        }
        SYNC::sync();
#pragma unroll
        for (int e = 0; e < E; ++e) v[e].y = f[m.at(q + P * e)];
        SYNC::sync();
    } else if constexpr (!LAST && LINR) {  // q < P and P * e share no bits
        SYNC::sync();
        const int fq = MAP::swz(q);
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = lds[m.base + (fq ^ MAP::swz(P * e))];
        SYNC::sync();
    } else if constexpr (!LAST) {
        SYNC::sync();
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = lds[m.at(q + P * e)];
        SYNC::sync();
    }
}

// full transform of length L on registers v[E] (positions q + P e)
template <typename T, int L, bool FWD, class MAP, class SYNC = BlockSync, bool SPLIT = false, int V = 0>
__device__ __forceinline__ void fft_regs(cx<T>* v, int q, cx<T>* lds, MAP m, const cx<T>* __restrict__ tw) {
    constexpr int S = Plan<L, V>::S;
    fft_stage<T, L, 0, FWD, MAP, SYNC, SPLIT, V>(v, q, lds, m, tw);
    if constexpr (S > 1) fft_stage<T, L, 1 % S, FWD, MAP, SYNC, SPLIT, V>(v, q, lds, m, tw);
    if constexpr (S > 2) fft_stage<T, L, 2 % S, FWD, MAP, SYNC, SPLIT, V>(v, q, lds, m, tw);
    if constexpr (S > 3) fft_stage<T, L, 3 % S, FWD, MAP, SYNC, SPLIT, V>(v, q, lds, m, tw);
}

// Pencil passes (TK neighbouring columns per workgroup) with one transform
// exchange real and imaginary parts separately (SPLIT): with 16-column
// tiles of 768-point pencils this lets three workgroups share a CU instead
// of one.  Passes with two transforms per pencil (filters) keep complex
// exchanges, whose barriers are half as many.
template <typename T, bool SPLIT> constexpr size_t pencil_lds_elem() { return SPLIT ? sizeof(T) : sizeof(cx<T>); }
// default LDS budget of a pencil tile: 128-byte row segments per pencil
// position (16 complex FP32 columns) at 768 points, 32 at 512
template <bool SPLIT> constexpr size_t pencil_lds_cap() { return SPLIT ? 69632 : 139264; }
template <typename T, int L, bool FWD, bool SPLIT = true, int V = 0>
__device__ __forceinline__ void fft_pencil(cx<T>* v, int q, cx<T>* lds, LdsMap m, const cx<T>* __restrict__ tw) {
    fft_regs<T, L, FWD, LdsMap, BlockSync, SPLIT, V>(v, q, lds, m, tw);
}

// Two transforms of the same length and positions (registers v and w) in
// lockstep: each stage applies one set of twiddle loads to both, and the two
// exchanges (through LDS regions m and m2) share their barriers.  Real and
// imaginary parts are exchanged in two rounds, as with SPLIT.
template <typename T, int L, int S, bool FWD, int V>
__device__ __forceinline__ void fft_stage2(cx<T>* v, cx<T>* w, int q, T* f, LdsMap m, LdsMap m2,
                                           const cx<T>* __restrict__ tw) {
    using PL = Plan<L, V>;
    constexpr int E = PL::E, P = PL::P, R = PL::r(S), B = E / R, NS = plan_ns<L, V>(S);
    constexpr bool LAST = (S == PL::S - 1);
    const cx<T>* tws = tw + plan_tw_offset<L, V>(S);
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = q + P * b;
        const int k = j % NS;
        cx<T> a[R], c[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            a[r] = v[b + r * B];
            c[r] = w[b + r * B];
        }
        if (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const cx<T> t = tws[k * (R - 1) + (r - 1)];
                a[r] = FWD ? cmul(a[r], t) : cmulc(a[r], t);
                c[r] = FWD ? cmul(c[r], t) : cmulc(c[r], t);
            }
        }
        Dft<T, R, FWD>::run(a);
        Dft<T, R, FWD>::run(c);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v[b + r * B] = a[r];
            w[b + r * B] = c[r];
        }
    }
    if constexpr (!LAST) {
#pragma unroll
        for (int part = 0; part < 2; ++part) {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = q + P * b;
                const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    f[m.at(base + r * NS)] = part ? v[b + r * B].y : v[b + r * B].x;
                    f[m2.at(base + r * NS)] = part ? w[b + r * B].y : w[b + r * B].x;
                }
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (part) {
                    v[e].y = f[m.at(q + P * e)];
                    w[e].y = f[m2.at(q + P * e)];
                } else {
                    v[e].x = f[m.at(q + P * e)];
                    w[e].x = f[m2.at(q + P * e)];
                }
            }
            __syncthreads();
        }
    }
}
template <typename T, int L, bool FWD, int V = 0>
__device__ __forceinline__ void fft_pencil2(cx<T>* v, cx<T>* w, int q, T* f, LdsMap m, LdsMap m2,
                                            const cx<T>* __restrict__ tw) {
    constexpr int S = Plan<L, V>::S;
    fft_stage2<T, L, 0, FWD, V>(v, w, q, f, m, m2, tw);
    if constexpr (S > 1) fft_stage2<T, L, 1 % S, FWD, V>(v, w, q, f, m, m2, tw);
    if constexpr (S > 2) fft_stage2<T, L, 2 % S, FWD, V>(v, w, q, f, m, m2, tw);
    if constexpr (S > 3) fft_stage2<T, L, 3 % S, FWD, V>(v, w, q, f, m, m2, tw);
}

// full transform of a row held by one thread group, exchanges through the
// per-stage swizzled LDS slots starting at `base`
template <typename T, int L, bool FWD, class SYNC = BlockSync, int V = 0>
__device__ __forceinline__ void fft_regs_row(cx<T>* v, int q, cx<T>* lds, int base, const cx<T>* __restrict__ tw) {
    constexpr int S = Plan<L, V>::S;
    using M0 = LdsSwzMap<L, 0, V>;
    using M1 = LdsSwzMap<L, 1 % S, V>;
    using M2 = LdsSwzMap<L, 2 % S, V>;
    using M3 = LdsSwzMap<L, 3 % S, V>;
    fft_stage<T, L, 0, FWD, M0, SYNC, false, V>(v, q, lds, M0{base}, tw);
    if constexpr (S > 1) fft_stage<T, L, 1 % S, FWD, M1, SYNC, false, V>(v, q, lds, M1{base}, tw);
    if constexpr (S > 2) fft_stage<T, L, 2 % S, FWD, M2, SYNC, false, V>(v, q, lds, M2{base}, tw);
    if constexpr (S > 3) fft_stage<T, L, 3 % S, FWD, M3, SYNC, false, V>(v, q, lds, M3{base}, tw);
}

}  // namespace mfic
