// Complex-to-complex 3D FFT (Grid_FFT<ccomplex_t>, grid_fft.cc:56-63,
// 207-270) and the semiclassical perturbation theory (PPT) fields of the
// field_eulerian output (ic_generator.cc:786-897): psi = sqrt(rho)
// exp(i (phi + phi2) / (hbar D+)), one kinetic drift exp(-i hbar k^2 D+ / 2)
// in k-space, the second-order kick exp(i phi2 / (hbar D+)), then the density
// |psi|^2 - 1 and the velocity hbar Im(psi* grad psi) / (1 + delta).
//
// A complex grid is [n0][n1][n2] complex.  The c2c transform runs the
// contiguous axis as row transforms in LDS and the two strided axes as the
// pencil passes of the r2c/c2r transforms; both directions are unitary.
#include <cmath>
#include <vector>

#include "fft_core.hpp"
#include "mfic_internal.hpp"
#include "ppt.hpp"

namespace mfic {

static inline unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// complex rows of length L along the contiguous axis, RW rows per block
template <typename T, int L, int RW, bool FWD>
__global__ void __launch_bounds__(Plan<L>::P * RW)
k_rows_c2c(cx<T>* __restrict__ data, long nrows, T scale, const cx<T>* __restrict__ tw) {
    constexpr int E = Plan<L>::E, P = Plan<L>::P, RP = row_lds_pitch<L>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int q = threadIdx.x % P, rw = threadIdx.x / P;
    const long row = (long)blockIdx.x * RW + rw;
    const bool active = row < nrows;
    cx<T>* r = data + (active ? row : 0) * (long)L;
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = active ? r[q + P * e] : cx<T>{0, 0};
    fft_regs_row<T, L, FWD>(v, q, lds, rw * RP, tw);
    if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) r[q + P * e] = cscale(v[e], scale);
    }
}

template <typename T, int L, bool FWD>
static void launch_rows_c2c(mfic_ctx ctx, cx<T>* data, long nrows, T scale) {
    constexpr int P = Plan<L>::P;
    int rw = 1;
    while (rw < 64 && P * rw * 2 <= 256) rw *= 2;
    while (rw > 1 && (size_t)row_lds_pitch<L>() * rw * sizeof(cx<T>) > (size_t)ctx->max_lds) rw /= 2;
    const cx<T>* tw = static_cast<const cx<T>*>(twiddles(ctx, L, 0, sizeof(T) == 8 ? MFIC_F64 : MFIC_F32));
    const size_t lds = (size_t)row_lds_pitch<L>() * rw * sizeof(cx<T>);
    const dim3 grid((unsigned)((nrows + rw - 1) / rw));
    ProfScope prof(ctx, std::string("fft_rows_c2c_n") + std::to_string(L) + (sizeof(T) == 8 ? "_f64" : "_f32"),
                   2.0 * (double)nrows * L * sizeof(cx<T>));
    switch (rw) {
#define RWCASE(K)                                                                                                   \
    case K:                                                                                                         \
        if constexpr (P * K <= 1024)                                                                                \
            hipLaunchKernelGGL((k_rows_c2c<T, L, K, FWD>), grid, dim3(P * K), lds, ctx->stream, data, nrows, scale, tw); \
        break;
        RWCASE(1) RWCASE(2) RWCASE(4) RWCASE(8) RWCASE(16) RWCASE(32) RWCASE(64)
#undef RWCASE
    default:
        throw Error(MFIC_ERR_SIZE, "bad row tile");
    }
    MFIC_LAUNCH_CHECK();
}

#define MFIC_C2C_SIZES(X) \
    X(4) X(6) X(8) X(9) X(12) X(16) X(18) X(24) X(32) X(36) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(576) X(768)

template <typename T> void fft_c2c(mfic_ctx ctx, void* data, const size_t n[3], bool forward, double scale) {
    for (int a = 0; a < 3; ++a)
        MFIC_CHECK(fft_length_supported(n[a]), MFIC_ERR_SIZE,
                   "unsupported complex FFT length " + std::to_string(n[a]) + " (n = 2^a 3^b, 4 <= n <= 768)");
    cx<T>* d = static_cast<cx<T>*>(data);
    const long rows = (long)(n[0] * n[1]);
    switch ((int)n[2]) {
#define X(S)                                                                               \
    case S:                                                                                \
        if (forward) launch_rows_c2c<T, S, true>(ctx, d, rows, (T)1);                     \
        else launch_rows_c2c<T, S, false>(ctx, d, rows, (T)1);                            \
        break;
        MFIC_C2C_SIZES(X)
#undef X
    }
    fft_pencils<T>(ctx, data, (int)n[1], (long)n[2], (long)(n[1] * n[2]), (int)n[2], (long)n[0], forward, 1.0,
                   "_c2c_ax1");
    fft_pencils<T>(ctx, data, (int)n[0], (long)(n[1] * n[2]), 0, (int)(n[1] * n[2]), 1, forward, scale, "_c2c_ax0");
}
template void fft_c2c<float>(mfic_ctx, void*, const size_t*, bool, double);
template void fft_c2c<double>(mfic_ctx, void*, const size_t*, bool, double);

// ----------------------------------------------------------- PPT kernels --
struct CGeom {
    int n;
    double kf;
    // Grid_FFT::get_k component of index i (index n/2 maps to +n/2)
    __device__ __forceinline__ double k(int i) const { return ((double)i - (double)(i > n / 2) * n) * kf; }
};

// per-block maxima of |phi(i+1,j,k) - phi|, |phi(i,j+1,k) - phi|,
// |phi(i,j,k+1) - phi| over i < n-1 (periodic in j, k): the loop of
// ic_generator.cc:819-836, which leaves out the last x-plane
template <typename T>
__global__ void k_ppt_maxdphi(const T* __restrict__ phi, int n, double* __restrict__ part) {
    __shared__ double red[256];
    const long pitch = n + 2;
    const long cells = (long)(n - 1) * n * n;
    double m = 0.0;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long)gridDim.x * blockDim.x) {
        const int k = (int)(c % n);
        const long r = c / n;
        const int j = (int)(r % n), i = (int)(r / n);
        const int jr = (j + 1) % n, kr = (k + 1) % n;
        const double p = phi[((long)i * n + j) * pitch + k];
        const double dx = fabs((double)phi[((long)(i + 1) * n + j) * pitch + k] - p);
        const double dy = fabs((double)phi[((long)i * n + jr) * pitch + k] - p);
        const double dz = fabs((double)phi[((long)i * n + j) * pitch + kr] - p);
        m = fmax(m, fmax(dx, fmax(dy, dz)));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// psi = exp(i (phi + phi2) / (hbar D+)) with sqrt(rho) = 1 (no baryon-CDM
// offset: C_species = 0, or delta_bc = 0 for Eisenstein & Hu)
template <typename T>
__global__ void k_ppt_init(cx<T>* __restrict__ psi, const T* __restrict__ phi, const T* __restrict__ phi2, int n,
                           double inv) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)n * n * n) return;
    const long r = c / n;
    const long o = r * (n + 2) + (c - r * n);
    double a = (double)phi[o];
    if (phi2) a += (double)phi2[o];
    double s, co;
    sincos(a * inv, &s, &co);
    psi[c] = cx<T>{(T)co, (T)s};
}

// drift: psi_k * exp(-i hbar |k|^2 D+ / 2)
template <typename T>
__global__ void k_ppt_drift(cx<T>* __restrict__ psi, CGeom G, double half_hbar_d) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = G.n;
    if (c >= n * n * n) return;
    const int k = (int)(c % n), j = (int)((c / n) % n), i = (int)(c / (n * n));
    const double kx = G.k(i), ky = G.k(j), kz = G.k(k);
    double s, co;
    sincos(-half_hbar_d * (kx * kx + ky * ky + kz * kz), &s, &co);
    const cx<T> v = psi[c];
    psi[c] = cx<T>{(T)((double)v.x * co - (double)v.y * s), (T)((double)v.x * s + (double)v.y * co)};
}

// kick: psi * exp(i phi2 / (hbar D+))
template <typename T>
__global__ void k_ppt_kick(cx<T>* __restrict__ psi, const T* __restrict__ phi2, int n, double inv) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)n * n * n) return;
    const long r = c / n;
    double s, co;
    sincos((double)phi2[r * (n + 2) + (c - r * n)] * inv, &s, &co);
    const cx<T> v = psi[c];
    psi[c] = cx<T>{(T)((double)v.x * co - (double)v.y * s), (T)((double)v.x * s + (double)v.y * co)};
}

// delta = |psi|^2 - 1 into a real [n][n][n+2] grid
template <typename T>
__global__ void k_ppt_density(const cx<T>* __restrict__ psi, T* __restrict__ delta, int n) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)n * n * n) return;
    const long r = c / n;
    const cx<T> v = psi[c];
    delta[r * (n + 2) + (c - r * n)] = (T)((double)v.x * v.x + (double)v.y * v.y - 1.0);
}

// g_k = i k_d psi_k (get_k, Nyquist index at +n/2)
template <typename T>
__global__ void k_ppt_grad(cx<T>* g, CGeom G, int d) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = G.n;
    if (c >= n * n * n) return;
    const int idx = d == 0 ? (int)(c / (n * n)) : (d == 1 ? (int)((c / n) % n) : (int)(c % n));
    const double kd = G.k(idx);
    const cx<T> v = g[c];
    g[c] = cx<T>{(T)(-kd * v.y), (T)(kd * v.x)};
}

// v_d = vunit Re[(psi* g - psi g*) / (2 i / hbar)] / (1 + delta)
//     = vunit hbar Im(psi* g) / (1 + delta)     (ic_generator.cc:889-896)
template <typename T>
__global__ void k_ppt_velocity(const cx<T>* __restrict__ psi, const cx<T>* __restrict__ g,
                               const T* __restrict__ delta, T* __restrict__ v, int n, double vh) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)n * n * n) return;
    const long r = c / n;
    const long o = r * (n + 2) + (c - r * n);
    const cx<T> p = psi[c], q = g[c];
    const double im = (double)p.x * q.y - (double)p.y * q.x;
    v[o] = (T)(vh * im / (1.0 + (double)delta[o]));
}

template <typename T>
void ppt_fields_t(mfic_ctx ctx, const PptInput& in, T* delta, T* vel[3], double* hbar_out) {
    const int n = in.n;
    const size_t n3 = (size_t)n * n * n;
    const size_t nn[3] = {(size_t)n, (size_t)n, (size_t)n};
    const int bs = 256;
    const T* phi = static_cast<const T*>(in.phi_r);
    const T* phi2 = static_cast<const T*>(in.phi2_r);
    // hbar from the largest neighbour jump of the growth-scaled phi
    // (ic_generator.cc:815-841)
    const unsigned nb = 1024;
    double* part = static_cast<double*>(ctx_scratch(ctx, "ppt_maxdphi", nb * sizeof(double)));
    hipLaunchKernelGGL(k_ppt_maxdphi<T>, dim3(nb), dim3(256), 0, ctx->stream, phi, n, part);
    MFIC_LAUNCH_CHECK();
    std::vector<double> hp(nb);
    MFIC_HIP(hipMemcpyAsync(hp.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MFIC_HIP(hipStreamSynchronize(ctx->stream));
    double maxdphi = 0.0;
    for (double x : hp) maxdphi = std::fmax(maxdphi, x);
    const double hbar = maxdphi / 3.14159265358979323846 / in.Dplus0 * 1.01;
    MFIC_CHECK(hbar > 0.0, MFIC_ERR_STATE, "PPT: the potential is constant (hbar = 0)");
    if (hbar_out) *hbar_out = hbar;
    const double inv = 1.0 / (hbar * in.Dplus0);
    cx<T>* psi = static_cast<cx<T>*>(in.psi);
    cx<T>* g = static_cast<cx<T>*>(in.grad);
    hipLaunchKernelGGL(k_ppt_init<T>, dim3(nblk(n3, bs)), dim3(bs), 0, ctx->stream, psi,
                       phi, in.order >= 2 ? phi2 : nullptr, n, inv);
    MFIC_LAUNCH_CHECK();
    const double norm = 1.0 / std::sqrt((double)n3);
    const CGeom G{n, 2.0 * 3.14159265358979323846 / in.L};
    fft_c2c<T>(ctx, psi, nn, true, norm);
    hipLaunchKernelGGL(k_ppt_drift<T>, dim3(nblk(n3, bs)), dim3(bs), 0, ctx->stream, psi, G, 0.5 * hbar * in.Dplus0);
    MFIC_LAUNCH_CHECK();
    fft_c2c<T>(ctx, psi, nn, false, norm);
    if (in.order >= 2) {
        hipLaunchKernelGGL(k_ppt_kick<T>, dim3(nblk(n3, bs)), dim3(bs), 0, ctx->stream, psi, phi2, n, inv);
        MFIC_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_ppt_density<T>, dim3(nblk(n3, bs)), dim3(bs), 0, ctx->stream, psi, delta, n);
    MFIC_LAUNCH_CHECK();
    // velocities from the current (ic_generator.cc:878-897): grad psi by
    // i k_d in k-space, one transform pair per component
    const double vh = in.vunit * hbar;
    for (int d = 0; d < 3; ++d) {
        MFIC_HIP(hipMemcpyAsync(g, psi, n3 * sizeof(cx<T>), hipMemcpyDeviceToDevice, ctx->stream));
        fft_c2c<T>(ctx, g, nn, true, norm);
        hipLaunchKernelGGL(k_ppt_grad<T>, dim3(nblk(n3, bs)), dim3(bs), 0, ctx->stream, g, G, d);
        MFIC_LAUNCH_CHECK();
        fft_c2c<T>(ctx, g, nn, false, norm);
        hipLaunchKernelGGL(k_ppt_velocity<T>, dim3(nblk(n3, bs)), dim3(bs), 0, ctx->stream, psi, g, delta, vel[d], n,
                           vh);
        MFIC_LAUNCH_CHECK();
    }
}

void ppt_fields(mfic_ctx ctx, int prec, const PptInput& in, void* delta, void* vel[3], double* hbar) {
    if (prec == MFIC_F64) {
        double* v[3] = {static_cast<double*>(vel[0]), static_cast<double*>(vel[1]), static_cast<double*>(vel[2])};
        ppt_fields_t<double>(ctx, in, static_cast<double*>(delta), v, hbar);
    } else {
        float* v[3] = {static_cast<float*>(vel[0]), static_cast<float*>(vel[1]), static_cast<float*>(vel[2])};
        ppt_fields_t<float>(ctx, in, static_cast<float*>(delta), v, hbar);
    }
}

}  // namespace mfic
