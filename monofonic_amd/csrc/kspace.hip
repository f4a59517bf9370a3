// Pointwise grid kernels: scaling, precision conversion, DC zeroing, inverse
// Laplacian, Fourier shift and the phi1 construction (white noise -> linear
// displacement potential with the E&H amplitude in FP64).
#include <cmath>
#include <mutex>
#include <string>

#include "cosmology.hpp"
#include "fft_core.hpp"
#include "kspace.hpp"

namespace mfic {

static inline unsigned nblocks(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T> __global__ void k_scale(T* __restrict__ d, size_t n, T s) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] *= s;
}
template <typename Ti, typename To> __global__ void k_convert(const Ti* __restrict__ in, To* __restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (To)in[i];
}
template <typename T> __global__ void k_fill(T* __restrict__ d, size_t n, T v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = v;
}

// k-space geometry helper: linear complex index -> (i, j, k), k vector
struct KGeom {
    int n0, n1, n2, nc;
    double kf0, kf1, kf2;
    __device__ __forceinline__ void ijk(size_t idx, int& i, int& j, int& k) const {
        k = (int)(idx % nc);
        size_t r = idx / nc;
        j = (int)(r % n1);
        i = (int)(r / n1);
    }
    // Grid_FFT::get_k (grid_fft.hh:282-296): index n/2 maps to +n/2
    __device__ __forceinline__ double kx(int i) const { return ((double)i - (double)(i > n0 / 2) * n0) * kf0; }
    __device__ __forceinline__ double ky(int j) const { return ((double)j - (double)(j > n1 / 2) * n1) * kf1; }
    __device__ __forceinline__ double kz(int k) const { return ((double)k - (double)(k > n2 / 2) * n2) * kf2; }
};

static KGeom geom_of(const mfic_grid_s* g) {
    KGeom G;
    G.n0 = (int)g->n[0]; G.n1 = (int)g->n[1]; G.n2 = (int)g->n[2]; G.nc = (int)g->nc();
    const double pi2 = 2.0 * 3.14159265358979323846;
    G.kf0 = pi2 / g->L[0]; G.kf1 = pi2 / g->L[1]; G.kf2 = pi2 / g->L[2];
    return G;
}

template <typename T> __global__ void k_inv_laplacian(cx<T>* __restrict__ d, KGeom G, size_t ncplx) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncplx) return;
    int i, j, k;
    G.ijk(idx, i, j, k);
    if (idx == 0) { d[0] = cx<T>{0, 0}; return; }
    const double kx = G.kx(i), ky = G.ky(j), kz = G.kz(k);
    const double k2 = kx * kx + ky * ky + kz * kz;
    cx<T> v = d[idx];
    d[idx] = cx<T>{(T)(-(double)v.x / k2), (T)(-(double)v.y / k2)};
}

template <typename T> __global__ void k_shift(cx<T>* __restrict__ d, KGeom G, size_t ncplx, double a0, double a1, double a2) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncplx) return;
    int i, j, k;
    G.ijk(idx, i, j, k);
    const double sh = a0 * G.kx(i) + a1 * G.ky(j) + a2 * G.kz(k);
    double s, c;
    sincos(sh, &s, &c);
    cx<T> v = d[idx];
    d[idx] = cx<T>{(T)(v.x * c - v.y * s), (T)(v.x * s + v.y * c)};
}

// noise normalisation + phi1 (ic_generator.cc:378-383, 509-516)
template <typename T>
__global__ void k_phi1(cx<T>* __restrict__ phi, const cx<T>* __restrict__ wn, KGeom G, size_t ncplx, EhParams eh,
                       double volfac, int fixing, int inversion) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncplx) return;
    int i, j, k;
    G.ijk(idx, i, j, k);
    cx<T> w = wn[idx];
    double wr = w.x, wi = w.y;
    if (fixing) {
        const double a = hypot(wr, wi);
        if (a != 0.0) { wr /= a; wi /= a; }
    }
    const double sg = inversion ? -1.0 : 1.0;
    wr = sg * wr / volfac;
    wi = sg * wi / volfac;
    const double kx = G.kx(i), ky = G.ky(j), kz = G.kz(k);
    const double k2 = kx * kx + ky * ky + kz * kz;
    double pr = 0.0, pi = 0.0;
    if (idx != 0) {
        const double A = eh_amplitude(eh, sqrt(k2));
        pr = -(wr * A) / k2;
        pi = -(wi * A) / k2;
    }
    phi[idx] = cx<T>{(T)pr, (T)pi};
}

// For a cubic grid with equal box sides |k|^2 = n2 kf^2 with the integer
// n2 = i'^2 + j'^2 + k'^2, so the FP64 transfer-function factor -A(k)/k^2 is
// tabulated once per distinct n2 (3 (N/2)^2 + 1 entries) instead of per mode.
__global__ void k_phi1_table(double* __restrict__ tab, int nmax, double kf, EhParams eh) {
    const int n2 = blockIdx.x * blockDim.x + threadIdx.x;
    if (n2 > nmax) return;
    if (n2 == 0) { tab[0] = 0.0; return; }
    const double k2 = (double)n2 * kf * kf;
    tab[n2] = -eh_amplitude(eh, sqrt(k2)) / k2;
}

template <typename T>
__global__ void k_phi1_tab(cx<T>* __restrict__ phi, const cx<T>* __restrict__ wn, const double* __restrict__ tab,
                           KGeom G, size_t ncplx, double volfac, int fixing, int inversion) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncplx) return;
    int i, j, k;
    G.ijk(idx, i, j, k);
    const cx<T> w = wn[idx];
    double wr = w.x, wi = w.y;
    if (fixing) {
        const double a = hypot(wr, wi);
        if (a != 0.0) { wr /= a; wi /= a; }
    }
    const double sg = inversion ? -1.0 : 1.0;
    const int si = i - (i > G.n0 / 2) * G.n0, sj = j - (j > G.n1 / 2) * G.n1, sk = k - (k > G.n2 / 2) * G.n2;
    const double F = tab[si * si + sj * sj + sk * sk];
    phi[idx] = cx<T>{(T)(sg * wr / volfac * F), (T)(sg * wi / volfac * F)};
}

// ----------------------------------------------------------- launchers ---

template <typename T> static T* dptr(mfic_grid g) { return static_cast<T*>(g->d); }

void grid_scale(mfic_grid g, double s) {
    const int bs = 256;
    if (g->prec == MFIC_F64)
        hipLaunchKernelGGL(k_scale<double>, dim3(nblocks(g->nreal, bs)), dim3(bs), 0, g->ctx->stream, dptr<double>(g), g->nreal, s);
    else
        hipLaunchKernelGGL(k_scale<float>, dim3(nblocks(g->nreal, bs)), dim3(bs), 0, g->ctx->stream, dptr<float>(g), g->nreal, (float)s);
    MFIC_LAUNCH_CHECK();
}

void grid_fill(mfic_grid g, double v) {
    const int bs = 256;
    ProfScope prof(g->ctx, "fill", (double)g->bytes());
    if (g->prec == MFIC_F64)
        hipLaunchKernelGGL(k_fill<double>, dim3(nblocks(g->nreal, bs)), dim3(bs), 0, g->ctx->stream, dptr<double>(g), g->nreal, v);
    else
        hipLaunchKernelGGL(k_fill<float>, dim3(nblocks(g->nreal, bs)), dim3(bs), 0, g->ctx->stream, dptr<float>(g), g->nreal, (float)v);
    MFIC_LAUNCH_CHECK();
}

void convert_buffer(mfic_ctx ctx, const void* in, int in_prec, void* out, int out_prec, size_t n) {
    const int bs = 256;
    if (in_prec == MFIC_F64 && out_prec == MFIC_F32)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(nblocks(n, bs)), dim3(bs), 0, ctx->stream, (const double*)in, (float*)out, n);
    else if (in_prec == MFIC_F32 && out_prec == MFIC_F64)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nblocks(n, bs)), dim3(bs), 0, ctx->stream, (const float*)in, (double*)out, n);
    else
        MFIC_HIP(hipMemcpyAsync(out, in, n * (in_prec == MFIC_F64 ? 8 : 4), hipMemcpyDeviceToDevice, ctx->stream));
    MFIC_LAUNCH_CHECK();
}

void grid_zero_dc(mfic_grid g) {
    // a fill kernel rather than a memset: safe inside captured graphs
    if (g->prec == MFIC_F64)
        hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(64), 0, g->ctx->stream, dptr<double>(g), (size_t)2, 0.0);
    else
        hipLaunchKernelGGL(k_fill<float>, dim3(1), dim3(64), 0, g->ctx->stream, dptr<float>(g), (size_t)2, 0.0f);
    MFIC_LAUNCH_CHECK();
}

void grid_inverse_laplacian(mfic_grid g) {
    const KGeom G = geom_of(g);
    const size_t nc = (size_t)G.n0 * G.n1 * G.nc;
    const int bs = 256;
    ProfScope prof(g->ctx, "inverse_laplacian", 2.0 * g->bytes());
    if (g->prec == MFIC_F64)
        hipLaunchKernelGGL(k_inv_laplacian<double>, dim3(nblocks(nc, bs)), dim3(bs), 0, g->ctx->stream, (cx<double>*)g->d, G, nc);
    else
        hipLaunchKernelGGL(k_inv_laplacian<float>, dim3(nblocks(nc, bs)), dim3(bs), 0, g->ctx->stream, (cx<float>*)g->d, G, nc);
    MFIC_LAUNCH_CHECK();
}

void grid_phase_shift(mfic_grid g, const double s[3]) {
    const KGeom G = geom_of(g);
    const size_t nc = (size_t)G.n0 * G.n1 * G.nc;
    const double a0 = s[0] * (g->L[0] / g->n[0]), a1 = s[1] * (g->L[1] / g->n[1]), a2 = s[2] * (g->L[2] / g->n[2]);
    const int bs = 256;
    if (g->prec == MFIC_F64)
        hipLaunchKernelGGL(k_shift<double>, dim3(nblocks(nc, bs)), dim3(bs), 0, g->ctx->stream, (cx<double>*)g->d, G, nc, a0, a1, a2);
    else
        hipLaunchKernelGGL(k_shift<float>, dim3(nblocks(nc, bs)), dim3(bs), 0, g->ctx->stream, (cx<float>*)g->d, G, nc, a0, a1, a2);
    MFIC_LAUNCH_CHECK();
}

bool phi1_cubic(mfic_grid phi) {
    const KGeom G = geom_of(phi);
    return G.n0 == G.n1 && G.n1 == G.n2 && G.kf0 == G.kf1 && G.kf1 == G.kf2;
}

const double* phi1_table(mfic_grid phi, const EhParams& eh, double* table) {
    MFIC_CHECK(eh.tk_n == 0 || eh.tk != nullptr, MFIC_ERR_STATE, "tabulated transfer function not on the device");
    MFIC_CHECK(phi1_cubic(phi), MFIC_ERR_SIZE, "the -A(k)/k^2 table needs a cubic box");
    const KGeom G = geom_of(phi);
    const int nmax = 3 * (G.n0 / 2) * (G.n0 / 2), bs = 256;
    double* tab = table ? table
                        : static_cast<double*>(ctx_scratch(phi->ctx, "phi1_tab", (size_t)(nmax + 1) * sizeof(double)));
    hipLaunchKernelGGL(k_phi1_table, dim3(nblocks((size_t)nmax + 1, bs)), dim3(bs), 0, phi->ctx->stream, tab, nmax,
                       G.kf0, eh);
    MFIC_LAUNCH_CHECK();
    return tab;
}

void phi1_from_noise(mfic_grid phi, mfic_grid wn, const EhParams& eh, double volfac, int fixing, int inversion,
                     double* table) {
    MFIC_CHECK(eh.tk_n == 0 || eh.tk != nullptr, MFIC_ERR_STATE, "tabulated transfer function not on the device");
    const KGeom G = geom_of(phi);
    const size_t nc = (size_t)G.n0 * G.n1 * G.nc;
    const int bs = 256;
    ProfScope prof(phi->ctx, "phi1", 2.0 * phi->bytes());
    if (phi1_cubic(phi)) {
        const double* tab = phi1_table(phi, eh, table);
        if (phi->prec == MFIC_F64)
            hipLaunchKernelGGL(k_phi1_tab<double>, dim3(nblocks(nc, bs)), dim3(bs), 0, phi->ctx->stream,
                               (cx<double>*)phi->d, (const cx<double>*)wn->d, tab, G, nc, volfac, fixing, inversion);
        else
            hipLaunchKernelGGL(k_phi1_tab<float>, dim3(nblocks(nc, bs)), dim3(bs), 0, phi->ctx->stream,
                               (cx<float>*)phi->d, (const cx<float>*)wn->d, tab, G, nc, volfac, fixing, inversion);
        MFIC_LAUNCH_CHECK();
        return;
    }
    if (phi->prec == MFIC_F64)
        hipLaunchKernelGGL(k_phi1<double>, dim3(nblocks(nc, bs)), dim3(bs), 0, phi->ctx->stream, (cx<double>*)phi->d,
                           (const cx<double>*)wn->d, G, nc, eh, volfac, fixing, inversion);
    else
        hipLaunchKernelGGL(k_phi1<float>, dim3(nblocks(nc, bs)), dim3(bs), 0, phi->ctx->stream, (cx<float>*)phi->d,
                           (const cx<float>*)wn->d, G, nc, eh, volfac, fixing, inversion);
    MFIC_LAUNCH_CHECK();
}

// --------------------------------------------- no-allocation phi1 entry --
// Scratch of mfic_phi1_from_noise_noalloc: the FP64 -A(k)/k^2 table per
// integer |n|^2 of a cubic grid, 3 (N/2)^2 + 1 values, sized for every grid
// the library transforms (N <= 768): 442,369 doubles (3.4 MiB).  Each GPU
// the library runs on holds its own copy until the process ends.  In host
// code `scratch` is only a stand-in for the symbol: its device address comes
// from hipGetSymbolAddress, once per device with that device current (at
// context creation), and only that address reaches kernels.
constexpr size_t kPhi1ScratchCount = 3 * 384 * 384 + 1;
static __device__ __align__(256) double scratch[kPhi1ScratchCount];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices] = {};

void phi1_scratch_setup(int device) {
    MFIC_CHECK(device >= 0 && device < kMaxDevices, MFIC_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(MFIC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<double*>(p);
}

static double* phi1_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    MFIC_CHECK(device >= 0 && device < kMaxDevices && g_scratch_dev[device], MFIC_ERR_STATE,
               "no-allocation scratch not set up on this device (create a context first)");
    return g_scratch_dev[device];
}

}  // namespace mfic

extern "C" int mfic_phi1_from_noise_noalloc(mfic_grid phi, mfic_grid wn, const mfic_cosmo_params* p, int fixing,
                                            int inversion) {
    using namespace mfic;
    return guard_call([&] {
        MFIC_CHECK(phi && phi->d && wn && wn->d, MFIC_ERR_ARG, "null or unallocated grid");
        MFIC_CHECK(p, MFIC_ERR_ARG, "null cosmology");
        MFIC_CHECK(phi->n[0] == wn->n[0] && phi->n[1] == wn->n[1] && phi->n[2] == wn->n[2] && phi->prec == wn->prec,
                   MFIC_ERR_ARG, "phi and noise grids differ in size or precision");
        MFIC_CHECK(wn->space == MFIC_KSPACE, MFIC_ERR_STATE, "white noise must be in k-space");
        MFIC_CHECK(p->tk_n == 0, MFIC_ERR_UNSUPPORTED,
                   "the no-allocation phi1 takes the Eisenstein & Hu transfer function only (a tabulated one needs "
                   "its spline nodes copied to device memory)");
        const KGeom G = geom_of(phi);
        const bool cubic = G.n0 == G.n1 && G.n1 == G.n2 && G.kf0 == G.kf1 && G.kf1 == G.kf2;
        if (cubic) {
            const size_t need = 3 * (size_t)(G.n0 / 2) * (size_t)(G.n0 / 2) + 1;
            MFIC_CHECK(need <= kPhi1ScratchCount, MFIC_ERR_SIZE,
                       "grid of " + std::to_string(G.n0) + "^3 needs a " + std::to_string(need) +
                           "-entry phi1 table; the no-allocation scratch holds " + std::to_string(kPhi1ScratchCount) +
                           " (N <= 768)");
        }
        EhParams eh;
        mfic_cosmo_scalars s;
        cosmo_compute(*p, &s, &eh);
        const double volfac = std::pow(phi->L[0] / (double)phi->n[0] / 2.0 / 3.14159265358979323846, 1.5);
        phi1_from_noise(phi, wn, eh, volfac, fixing, inversion, cubic ? phi1_scratch(phi->ctx->device) : nullptr);
        phi->space = MFIC_KSPACE;
    });
}
