// Differentiable 2LPT: the forward map from real-space white noise to the
// 2LPT displacement fields and its hand-written adjoint (reverse mode).
//
// Forward (ic_generator.cc:378-383, 509-555, 916-970 with the NGenIC
// convention of empty Nyquist planes, random_ngenic.cc:117):
//   w = P_N r2c(delta)                 P_N: zero the i, j or k = N/2 planes
//   phi1 = S w,  S(k) = -A(k) / (k^2 volfac), DC := 0
//   phi2 = lap^-1 mu2(phi1),  mu2 = T[xx yy + xx zz + yy zz - xy^2 - xz^2 - yz^2]
//   Psi_d = c2r( i g_d (g1 phi1 + g2 phi2) ) * unit, DC := 0
// with T the Orszag truncation (convolution.hh:495-651).  Adjoint, for
// cotangents G_d of Psi_d and the full-spectrum inner product:
//   lam = unit * sum_d (-i g_d) r2c(G_d), DC := 0
//   lam1 = g1 lam;  lam_mu = P_N lap^-1 (g2 lam)
//   lam1 += D_xx T(lam_mu (H_yy + H_zz)) + D_yy T(lam_mu (H_xx + H_zz))
//         + D_zz T(lam_mu (H_xx + H_yy)) - 2 D_xy T(lam_mu H_xy)
//         - 2 D_xz T(lam_mu H_xz) - 2 D_yz T(lam_mu H_yz)
//   grad = c2r( P_N S lam1 )
// (the adjoint of a -> T(a b) is l -> T(l b) for real band-limited b, and
// D_ab = -g_a g_b, S and lap^-1 are real diagonal multipliers).  phi1 is
// recomputed from delta in the adjoint instead of being stored.
#include <cmath>
#include <string>

#include "cosmology.hpp"
#include "fft_core.hpp"
#include "kspace.hpp"
#include "mfic_internal.hpp"

namespace mfic {

namespace {

struct Geo {
    int n0, n1, n2, nc;
    double kf0, kf1, kf2;
    __device__ __forceinline__ void ijk(size_t idx, int& i, int& j, int& k) const {
        k = (int)(idx % nc);
        const size_t r = idx / nc;
        j = (int)(r % n1);
        i = (int)(r / n1);
    }
    // Grid_FFT::gradient component (grid_fft.hh:376-385): 0 at index n/2
    __device__ __forceinline__ double g(int d, int i, int j, int k) const {
        const int idx = d == 0 ? i : (d == 1 ? j : k);
        const int n = d == 0 ? n0 : (d == 1 ? n1 : n2);
        const double kf = d == 0 ? kf0 : (d == 1 ? kf1 : kf2);
        return idx == n / 2 ? 0.0 : ((double)idx - (double)(idx > n / 2) * n) * kf;
    }
};

Geo geo_of(mfic_grid g) {
    Geo G;
    G.n0 = (int)g->n[0];
    G.n1 = (int)g->n[1];
    G.n2 = (int)g->n[2];
    G.nc = (int)g->nc();
    const double pi2 = 2.0 * 3.14159265358979323846;
    G.kf0 = pi2 / g->L[0];
    G.kf1 = pi2 / g->L[1];
    G.kf2 = pi2 / g->L[2];
    return G;
}

template <typename T> __global__ void k_zero_nyquist(cx<T>* __restrict__ d, Geo G, size_t nk) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nk) return;
    int i, j, k;
    G.ijk(idx, i, j, k);
    if (i == G.n0 / 2 || j == G.n1 / 2 || k == G.n2 / 2) d[idx] = cx<T>{0, 0};
}

// out = unit * sum_d (-i g_d) G_d, DC := 0
template <typename T>
__global__ void k_grad_adjoint(cx<T>* __restrict__ out, const cx<T>* __restrict__ gx, const cx<T>* __restrict__ gy,
                               const cx<T>* __restrict__ gz, Geo G, size_t nk, double unit) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nk) return;
    int i, j, k;
    G.ijk(idx, i, j, k);
    const cx<T>* in[3] = {gx, gy, gz};
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double gd = G.g(d, i, j, k);
        const cx<T> v = in[d][idx];
        re += gd * (double)v.y;   // (-i g)(x + i y) = g y - i g x
        im -= gd * (double)v.x;
    }
    out[idx] = idx == 0 ? cx<T>{0, 0} : cx<T>{(T)(re * unit), (T)(im * unit)};
}

// out += c * (-g_a g_b) r
template <typename T>
__global__ void k_hess_madd(cx<T>* __restrict__ out, const cx<T>* __restrict__ r, Geo G, size_t nk, int a, int b,
                            double c) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nk) return;
    int i, j, k;
    G.ijk(idx, i, j, k);
    const double m = -c * G.g(a, i, j, k) * G.g(b, i, j, k);
    const cx<T> v = r[idx], o = out[idx];
    out[idx] = cx<T>{(T)((double)o.x + m * (double)v.x), (T)((double)o.y + m * (double)v.y)};
}

unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void kspace_zero_nyquist(mfic_grid g) {
    const Geo G = geo_of(g);
    const size_t nk = (size_t)G.n0 * G.n1 * G.nc;
    if (g->prec == MFIC_F64)
        hipLaunchKernelGGL(k_zero_nyquist<double>, dim3(nblk(nk)), dim3(256), 0, g->ctx->stream, (cx<double>*)g->d, G, nk);
    else
        hipLaunchKernelGGL(k_zero_nyquist<float>, dim3(nblk(nk)), dim3(256), 0, g->ctx->stream, (cx<float>*)g->d, G, nk);
    MFIC_LAUNCH_CHECK();
}

void kspace_grad_adjoint(mfic_grid out, mfic_grid gx, mfic_grid gy, mfic_grid gz, double unit) {
    const Geo G = geo_of(out);
    const size_t nk = (size_t)G.n0 * G.n1 * G.nc;
    ProfScope prof(out->ctx, "adjoint_gradient", 4.0 * out->bytes());
    if (out->prec == MFIC_F64)
        hipLaunchKernelGGL(k_grad_adjoint<double>, dim3(nblk(nk)), dim3(256), 0, out->ctx->stream, (cx<double>*)out->d,
                           (const cx<double>*)gx->d, (const cx<double>*)gy->d, (const cx<double>*)gz->d, G, nk, unit);
    else
        hipLaunchKernelGGL(k_grad_adjoint<float>, dim3(nblk(nk)), dim3(256), 0, out->ctx->stream, (cx<float>*)out->d,
                           (const cx<float>*)gx->d, (const cx<float>*)gy->d, (const cx<float>*)gz->d, G, nk, unit);
    MFIC_LAUNCH_CHECK();
}

void kspace_hess_madd(mfic_grid out, mfic_grid r, int a, int b, double c) {
    const Geo G = geo_of(out);
    const size_t nk = (size_t)G.n0 * G.n1 * G.nc;
    ProfScope prof(out->ctx, "adjoint_hessian_madd", 3.0 * out->bytes());
    if (out->prec == MFIC_F64)
        hipLaunchKernelGGL(k_hess_madd<double>, dim3(nblk(nk)), dim3(256), 0, out->ctx->stream, (cx<double>*)out->d,
                           (const cx<double>*)r->d, G, nk, a, b, c);
    else
        hipLaunchKernelGGL(k_hess_madd<float>, dim3(nblk(nk)), dim3(256), 0, out->ctx->stream, (cx<float>*)out->d,
                           (const cx<float>*)r->d, G, nk, a, b, c);
    MFIC_LAUNCH_CHECK();
}

}  // namespace mfic
