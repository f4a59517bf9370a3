// Orszag 3/2-rule dealiased products (convolution.hh:412-651) and the LPT
// source block of ic_generator::run (ic_generator.cc:523-629).
//
// Two entry levels:
//  * convolve2 / convolve3: the reference's per-product schedule (pad both
//    operands, two c2r, product, r2c, unpad + output operator), kept for the
//    drop-in OrszagConvolver interface;
//  * lpt_sources: the reorganised schedule used by the pipeline.  Each
//    Hessian of phi1 and phi2 goes through the axis-0 / axis-1 padded c2r
//    passes once; one fused row pass per LPT order brings all of them to
//    real space in LDS, sums every product sharing one truncation (in FP64)
//    and returns the r2c'd rows, so padded real fields never reach HBM.  The
//    cubic 3a terms keep their two-stage truncation trunc(trunc(a b) c)
//    (convolution.hh:525-532) with the inner products grouped by their outer
//    factor.  Equal to the reference schedule in exact arithmetic (every step
//    is linear in the grouped quantity).
#include <algorithm>
#include <cmath>

#include "fft_core.hpp"
#include "kspace.hpp"
#include "mfic_internal.hpp"
#include "orszag.hpp"
#include "fused.hpp"
#include "pruned.hpp"

namespace mfic {

static inline unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct Dims {
    int n0, n1, n2, nc;       // N-grid
    int p0, p1, p2, pc;       // padded grid
    double kf0, kf1, kf2;     // 2 pi / L
};

// Grid_FFT::gradient imaginary part (grid_fft.hh:376-385): 0 at index n/2
__device__ __forceinline__ double grad_comp(int idx, int n, double kf) {
    return idx == n / 2 ? 0.0 : ((double)idx - (double)(idx > n / 2) * n) * kf;
}

template <typename T> struct DevOperand {
    const cx<T>* f;
    int kind, d0, d1, d2, d3;
};

// operand value at N-grid mode (i,j,k) (convolution.hh:74-275)
template <typename T>
__device__ __forceinline__ void operand_val(const DevOperand<T>& o, const Dims& D, int i, int j, int k, size_t idx,
                                            double& re, double& im) {
    const cx<T> v = o.f[idx];
    const double fr = (double)v.x, fi = (double)v.y;
    double g[3];
    g[0] = grad_comp(i, D.n0, D.kf0);
    g[1] = grad_comp(j, D.n1, D.kf1);
    g[2] = grad_comp(k, D.n2, D.kf2);
    double s;
    switch (o.kind) {
    case MFIC_OPND_HESS: s = -(g[o.d0] * g[o.d1]); break;
    case MFIC_OPND_SUM_HESS: s = -(g[o.d0] * g[o.d1]) + -(g[o.d2] * g[o.d3]); break;
    case MFIC_OPND_DIFF_HESS: s = -(g[o.d0] * g[o.d1]) - -(g[o.d2] * g[o.d3]); break;
    case MFIC_OPND_GRAD: re = -g[o.d0] * fi; im = g[o.d0] * fr; return;
    default: re = fr; im = fi; return;
    }
    re = s * fr;
    im = s * fi;
}

// padded index -> source index, -1 inside the zero band (pad_insert,
// convolution.hh:562-571: i > N/2 goes to i + N/2)
__device__ __forceinline__ int pad_src(int ip, int n) {
    if (ip <= n / 2) return ip;
    if (ip > n) return ip - n / 2;
    return -1;
}

// write the whole padded k-space array: operand * (3/2)^{3/2} or 0
template <typename T>
__global__ void k_pad_insert(cx<T>* __restrict__ fp, DevOperand<T> o, Dims D, double rfac, size_t ntot) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ntot) return;
    const int kp = (int)(idx % D.pc);
    const size_t r = idx / D.pc;
    const int jp = (int)(r % D.p1), ip = (int)(r / D.p1);
    const int i = pad_src(ip, D.n0), j = pad_src(jp, D.n1);
    cx<T> out = {0, 0};
    if (i >= 0 && j >= 0 && kp <= D.n2 / 2) {
        double re, im;
        operand_val(o, D, i, j, kp, ((size_t)i * D.n1 + j) * D.nc + kp, re, im);
        out = cx<T>{(T)(re * rfac), (T)(im * rfac)};
    }
    fp[idx] = out;
}

// unpad + Nyquist zeroing + output operator (convolution.hh:600-636,
// operators.hh:25-35)
template <typename T>
__global__ void k_unpad(const cx<T>* __restrict__ fp, cx<T>* __restrict__ out, Dims D, double inv_rfac, int op,
                        double coef, size_t ntot) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ntot) return;
    const int k = (int)(idx % D.nc);
    const size_t r = idx / D.nc;
    const int j = (int)(r % D.n1), i = (int)(r / D.n1);
    const int ip = i > D.n0 / 2 ? i + D.n0 / 2 : i, jp = j > D.n1 / 2 ? j + D.n1 / 2 : j;
    double re = 0.0, im = 0.0;
    if (!(i == D.n0 / 2 || j == D.n1 / 2 || k == D.n2 / 2)) {
        const cx<T> v = fp[((size_t)ip * D.p1 + jp) * D.pc + k];
        re = (double)v.x * inv_rfac;
        im = (double)v.y * inv_rfac;
    }
    cx<T> o = out[idx];
    switch (op) {
    case MFIC_OP_ASSIGN: o = cx<T>{(T)re, (T)im}; break;
    case MFIC_OP_ADD: o = cx<T>{(T)((double)o.x + re), (T)((double)o.y + im)}; break;
    case MFIC_OP_SUB: o = cx<T>{(T)((double)o.x - re), (T)((double)o.y - im)}; break;
    default: o = cx<T>{(T)((double)o.x + re * coef), (T)((double)o.y + im * coef)}; break;
    }
    out[idx] = o;
}

template <typename T> __global__ void k_mul(T* __restrict__ a, const T* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] *= b[i];
}

// out = sum_t c_t A_t B_t over real padded elements, accumulated in FP64
struct ProdTerms {
    const void* a[16];
    const void* b[16];
    double c[16];
    int n;
};
template <typename T>
__global__ void k_prodsum(T* __restrict__ out, ProdTerms P, size_t ntot) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot) return;
    double acc = 0.0;
    for (int t = 0; t < P.n; ++t)
        acc += P.c[t] * ((double)static_cast<const T*>(P.a[t])[i] * (double)static_cast<const T*>(P.b[t])[i]);
    out[i] = (T)acc;
}

// ------------------------------------------------------------- host -----

static Dims dims_of(const size_t n[3], const double L[3]) {
    Dims D;
    D.n0 = (int)n[0]; D.n1 = (int)n[1]; D.n2 = (int)n[2]; D.nc = D.n2 / 2 + 1;
    D.p0 = 3 * D.n0 / 2; D.p1 = 3 * D.n1 / 2; D.p2 = 3 * D.n2 / 2; D.pc = D.p2 / 2 + 1;
    const double pi2 = 2.0 * 3.14159265358979323846;
    D.kf0 = pi2 / L[0]; D.kf1 = pi2 / L[1]; D.kf2 = pi2 / L[2];
    return D;
}

ConvCore::ConvCore(mfic_ctx c, const size_t n[3], const double L[3], int p) : ctx(c), prec(p) {
    for (int i = 0; i < 3; ++i) {
        MFIC_CHECK(n[i] % 2 == 0, MFIC_ERR_SIZE, "Orszag convolver for LPTorder>1 requires even grid size!");
        N[i] = n[i];
        Np[i] = 3 * n[i] / 2;
        len[i] = L[i];
    }
    MFIC_CHECK(Np[2] % 2 == 0, MFIC_ERR_SIZE, "padded contiguous axis 3N/2 must be even (N multiple of 4)");
    for (int i = 0; i < 3; ++i)
        MFIC_CHECK(fft_length_supported(N[i]) && fft_length_supported(Np[i]), MFIC_ERR_SIZE,
                   "unsupported grid size for the Orszag convolver");
}

size_t ConvCore::padded_reals() const { return Np[0] * Np[1] * (Np[2] + 2); }
size_t ConvCore::grid_reals() const { return N[0] * N[1] * (N[2] + 2); }
size_t ConvCore::esize() const { return prec == MFIC_F64 ? 8 : 4; }

template <typename T>
static void pad_insert_t(const ConvCore& c, const OperandDesc& o, void* fp) {
    Dims D = dims_of(c.N, c.len);
    DevOperand<T> od{static_cast<const cx<T>*>(o.f), o.kind, o.d[0], o.d[1], o.d[2], o.d[3]};
    const size_t ntot = (size_t)D.p0 * D.p1 * D.pc;
    const double rfac = std::pow(1.5, 1.5);
    ProfScope prof(c.ctx, "pad_insert", (double)ntot * sizeof(cx<T>) + (double)D.n0 * D.n1 * D.nc * sizeof(cx<T>));
    hipLaunchKernelGGL(k_pad_insert<T>, dim3(nblk(ntot, 256)), dim3(256), 0, c.ctx->stream, (cx<T>*)fp, od, D, rfac, ntot);
    MFIC_LAUNCH_CHECK();
}

template <typename T>
static void unpad_t(const ConvCore& c, const void* fp, void* out, int op, double coef) {
    Dims D = dims_of(c.N, c.len);
    const size_t ntot = (size_t)D.n0 * D.n1 * D.nc;
    const double rfac = std::sqrt((double)(c.Np[0] * c.Np[1] * c.Np[2])) / std::sqrt((double)(c.N[0] * c.N[1] * c.N[2]));
    ProfScope prof(c.ctx, "unpad", (double)ntot * sizeof(cx<T>) * (op == MFIC_OP_ASSIGN ? 2.0 : 3.0));
    hipLaunchKernelGGL(k_unpad<T>, dim3(nblk(ntot, 256)), dim3(256), 0, c.ctx->stream, (const cx<T>*)fp, (cx<T>*)out, D,
                       1.0 / rfac, op, coef, ntot);
    MFIC_LAUNCH_CHECK();
}

void ConvCore::pad_insert(const OperandDesc& o, void* fp) const {
    if (prec == MFIC_F64) pad_insert_t<double>(*this, o, fp);
    else pad_insert_t<float>(*this, o, fp);
}
void ConvCore::unpad(const void* fp, void* out, int op, double coef) const {
    if (prec == MFIC_F64) unpad_t<double>(*this, fp, out, op, coef);
    else unpad_t<float>(*this, fp, out, op, coef);
}
void ConvCore::c2r_padded(void* fp) const {
    const double s = 1.0 / std::sqrt((double)Np[0] * (double)Np[1] * (double)Np[2]);
    if (prec == MFIC_F64) fft_c2r<double>(ctx, (double*)fp, Np, s);
    else fft_c2r<float>(ctx, (float*)fp, Np, s);
}
void ConvCore::r2c_padded(void* fp) const {
    const double s = 1.0 / std::sqrt((double)Np[0] * (double)Np[1] * (double)Np[2]);
    if (prec == MFIC_F64) fft_r2c<double>(ctx, (double*)fp, Np, s);
    else fft_r2c<float>(ctx, (float*)fp, Np, s);
}
void ConvCore::mul(void* a, const void* b) const {
    const size_t n = padded_reals();
    ProfScope prof(ctx, "mul", 3.0 * n * esize());
    if (prec == MFIC_F64)
        hipLaunchKernelGGL(k_mul<double>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, (double*)a, (const double*)b, n);
    else
        hipLaunchKernelGGL(k_mul<float>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, (float*)a, (const float*)b, n);
    MFIC_LAUNCH_CHECK();
}
void ConvCore::prodsum(void* out, const std::vector<Term>& terms) const {
    MFIC_CHECK(terms.size() <= 16, MFIC_ERR_ARG, "too many product terms");
    ProdTerms P{};
    P.n = (int)terms.size();
    for (size_t t = 0; t < terms.size(); ++t) { P.a[t] = terms[t].a; P.b[t] = terms[t].b; P.c[t] = terms[t].c; }
    const size_t n = padded_reals();
    std::vector<const void*> distinct;
    for (const Term& t : terms)
        for (const void* p : {t.a, t.b})
            if (std::find(distinct.begin(), distinct.end(), p) == distinct.end()) distinct.push_back(p);
    ProfScope prof(ctx, "prodsum_" + std::to_string(terms.size()), (double)(distinct.size() + 1) * n * esize());
    if (prec == MFIC_F64)
        hipLaunchKernelGGL(k_prodsum<double>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, (double*)out, P, n);
    else
        hipLaunchKernelGGL(k_prodsum<float>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, (float*)out, P, n);
    MFIC_LAUNCH_CHECK();
}

// convolve2: both operands to padded real space (pruned c2r), product, pruned
// r2c with the unpad rules and the output operator (convolution.hh:495-518)
size_t ConvCore::kbuf_reals() const { return 2 * Np[0] * Np[1] * (N[2] / 2 + 1); }
size_t ConvCore::real_reals() const { return Np[0] * Np[1] * Np[2]; }

void ConvCore::convolve2(const OperandDesc& a, const OperandDesc& b, void* out, int op, double coef, void* kbuf,
                         void* r1, void* r2) const {
    padded_c2r_from_operand(ctx, prec, a, N[0], len[0], kbuf, r1);
    padded_c2r_from_operand(ctx, prec, b, N[0], len[0], kbuf, r2);
    padded_r2c_products(ctx, prec, {{r1, r2, 1.0}}, N[0], len[0], kbuf, out, op, coef, false);
}

void ConvCore::convolve3(const OperandDesc& a, const OperandDesc& b, const OperandDesc& c, void* out, int op,
                         double coef, void* kbuf, void* r1, void* r2, void* fbuf) const {
    convolve2(a, b, fbuf, MFIC_OP_ASSIGN, 0.0, kbuf, r1, r2);
    OperandDesc id{fbuf, MFIC_OPND_IDENT, {0, 0, 0, 0}};
    convolve2(id, c, out, op, coef, kbuf, r1, r2);
}

// ------------------------------------------------- reorganised sources ----

void lpt_sources(mfic_ctx ctx, int order, mfic_grid phi1, mfic_grid phi2, mfic_grid phi3, mfic_grid A3x,
                 mfic_grid A3y, mfic_grid A3z) {
    if (order < 2) return;
    ConvCore C(ctx, phi1->n, phi1->L, phi1->prec);
    MFIC_CHECK(phi1->n[0] == phi1->n[1] && phi1->n[1] == phi1->n[2], MFIC_ERR_SIZE, "LPT sources need a cubic grid");
    const size_t n = phi1->n[0];
    const double L = phi1->L[0];
    const int prec = phi1->prec;
    // compact padded half-spectrum buffers, consecutive: inputs of the fused
    // row pass (phi1 Hessians, phi2 Hessians, Q) and its outputs
    const size_t kb = C.kbuf_reals() * C.esize();
    const FusedKind k1 = order >= 3 ? FusedKind::Src2Q : FusedKind::Src2;
    // inputs of the row passes, consecutive: phi1 Hessians [0, 6), phi2
    // Hessians [6, 12), Q [12, 16).  The 2LPT pass writes its outputs (phi2
    // source, Q) to [11, 16): slot 11 is consumed before the phi2 Hessians
    // fill it.
    char* kin = static_cast<char*>(ctx_scratch(ctx, "fused_in", (size_t)(order >= 3 ? 16 : 7) * kb));
    void* xbuf = ctx_scratch(ctx, "hess_x", 3 * kb);
    padded_hessians_to_kbuf(ctx, prec, phi1->d, n, L, xbuf, kin);
    char* src2 = kin + (order >= 3 ? 11 : 6) * kb;
    // phi2 source (ic_generator.cc:536-541) and, for 3LPT, the inner 3a
    // products Qzz Qyz Qxx Qyy (ic_generator.cc:581-585) from one row pass
    fused_source_rows(ctx, prec, k1, n, kin, src2);
    padded_r2c_from_kbuf(ctx, prec, n, L, src2, phi2->d, MFIC_OP_ASSIGN, 0.0, true);  // -1/k^2: ic_generator.cc:554
    phi2->space = MFIC_KSPACE;
    if (order < 3) return;
    // Q: truncated once (first stage of trunc(trunc(a b) c), convolution.hh:
    // 525-532) and padded again for the outer products, in place
    for (int q = 0; q < 4; ++q) padded_truncate_repad(ctx, prec, n, kin + (12 + q) * kb);
    padded_hessians_to_kbuf(ctx, prec, phi2->d, n, L, xbuf, kin + 6 * kb);
    // phi3 source (3a outer + 3b, ic_generator.cc:581-609) and A3
    // (ic_generator.cc:614-627), then -1/k^2 each
    char* kout = static_cast<char*>(ctx_scratch(ctx, "fused_out", 4 * kb));
    fused_source_rows(ctx, prec, FusedKind::Src3, n, kin, kout);
    mfic_grid outs[4] = {phi3, A3x, A3y, A3z};
    for (int o = 0; o < 4; ++o) {
        padded_r2c_from_kbuf(ctx, prec, n, L, kout + o * kb, outs[o]->d, MFIC_OP_ASSIGN, 0.0, true);
        outs[o]->space = MFIC_KSPACE;
    }
}

template <typename T> __global__ void k_add(T* __restrict__ a, const T* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] += b[i];
}
void add_inplace(mfic_ctx ctx, int prec, void* a, const void* b, size_t n) {
    if (prec == MFIC_F64)
        hipLaunchKernelGGL(k_add<double>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, (double*)a, (const double*)b, n);
    else
        hipLaunchKernelGGL(k_add<float>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, (float*)a, (const float*)b, n);
    MFIC_LAUNCH_CHECK();
}

}  // namespace mfic
