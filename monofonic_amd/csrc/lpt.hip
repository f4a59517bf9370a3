// LPT displacement / velocity assembly in k-space (ic_generator.cc:931-970,
// 987-1047) and the Bravais lattice readout of q + Psi in FP64
// (particle_generator.hh:499-544, 597-639).
#include <cmath>

#include "fft_core.hpp"
#include "lpt.hpp"
#include "mfic_internal.hpp"

namespace mfic {

static inline unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct AsmArgs {
    const void* phi[3];  // phi1, phi2, phi3 (unscaled potentials; may be null)
    const void* A3[3];   // A3x, A3y, A3z
    int n0, n1, n2, nc, order, comp;
    double kf0, kf1, kf2;
    double c1, c2, c3, cA;  // phi_n and A3 coefficients (growth [x vfac])
    double unit;            // position or velocity unit / L
    double sh0, sh1, sh2;   // optional phase shift s_a dx_a (bcc sub-lattice)
    PltDev plt;             // PLT tables (plt.Dr == nullptr: Fourier gradient)
};

// Fourier gradient (operators.hh:50-55): i (idx - [idx>n/2] n) 2pi/L, 0 at n/2
__device__ __forceinline__ double fgrad(int idx, int n, double kf) {
    return idx == n / 2 ? 0.0 : ((double)idx - (double)(idx > n / 2) * n) * kf;
}

// k-space CIC on an ng x ng x (ng/2+1) table (Grid_FFT::get_cic_kspace,
// grid_fft.hh:347-373: periodic in x and y, clamped at the last z plane).
// The eight corners and weights are formed once per mode and applied to
// every table.
struct CicStencil {
    unsigned off[8];
    double w[8];
};
__device__ inline void cic_stencil(int ng, double x, double y, double z, CicStencil& s) {
    const int nc = ng / 2 + 1;
    const int ix = (int)floor(x), iy = (int)floor(y), iz = (int)floor(z);
    const double dx = x - ix, tx = 1.0 - dx, dy = y - iy, ty = 1.0 - dy, dz = z - iz, tz = 1.0 - dz;
    const int ix1 = (ix + 1) % ng, iy1 = (iy + 1) % ng, iz1 = min(iz + 1, nc - 1);
    auto at = [&](int a, int b, int c) { return (unsigned)((a * ng + b) * nc + c); };
    s.off[0] = at(ix, iy, iz);    s.w[0] = tx * ty * tz;
    s.off[1] = at(ix, iy, iz1);   s.w[1] = tx * ty * dz;
    s.off[2] = at(ix, iy1, iz);   s.w[2] = tx * dy * tz;
    s.off[3] = at(ix, iy1, iz1);  s.w[3] = tx * dy * dz;
    s.off[4] = at(ix1, iy, iz);   s.w[4] = dx * ty * tz;
    s.off[5] = at(ix1, iy, iz1);  s.w[5] = dx * ty * dz;
    s.off[6] = at(ix1, iy1, iz);  s.w[6] = dx * dy * tz;
    s.off[7] = at(ix1, iy1, iz1); s.w[7] = dx * dy * dz;
}
__device__ inline double cic_apply(const double* __restrict__ T, const CicStencil& s) {
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) v += T[s.off[c]] * s.w[c];
    return v;
}

// lattice_gradient::gradient for all three axes (particle_plt.hh:551-585)
// from the interpolated eigenvalue tables Dr, Dtheta, Dphi at the mode's
// table position (ix, iy, iz) = (i, j, k) * ng / n; c0 = 1 / (mr L)
__device__ inline void plt_combine(int ng, double c0, double ix, double iy, double iz, double Dr, double Dt,
                                   double Dp, double g[3]) {
    const double kf = 2.0 * 3.14159265358979323846;
    const double hn = (double)(ng / 2);
    const double kx = (ix - (double)(ix > hn) * ng) * kf, ky = (iy - (double)(iy > hn) * ng) * kf,
                 kz = (iz - (double)(iz > hn) * ng) * kf;
    const double kp2 = kx * kx + ky * ky;
    const double kr = sqrt(kp2 + kz * kz);
    const double kmod = kr * c0;
    // spherical basis of k without trigonometry: theta = acos(kz/|k|),
    // phi = atan2(ky, kx) (phi = 0 on the kz axis, as atan2(0, 0))
    double st = 0.0, ct = 1.0, sp = 0.0, cp = 1.0;
    if (kr > 0.0) {
        const double kp = sqrt(kp2), ir = 1.0 / kr;
        ct = kz * ir;
        st = kp * ir;
        if (kp > 0.0) {
            const double ip = 1.0 / kp;
            cp = kx * ip;
            sp = ky * ip;
        }
    }
    g[0] = kmod * (Dr * st * cp + Dt * ct * cp - Dp * sp);
    g[1] = kmod * (Dr * st * sp + Dt * ct * sp + Dp * cp);
    g[2] = kmod * (Dr * ct - Dt * st);
}

// lattice_gradient::gradient + vfac_corr (particle_plt.hh:551-585) for
// N-grid mode (i, j, k) with the eight-corner CIC lookups of the tables;
// mr = ng / n and c0 = 1 / (mr L) come from the host
__device__ inline void plt_grad(const PltDev& P, double mr, double c0, int i, int j, int k, double g[3],
                                double* vcorr) {
    const double ix = i * mr, iy = j * mr, iz = k * mr;
    CicStencil cs;
    cic_stencil(P.ng, ix, iy, iz, cs);
    plt_combine(P.ng, c0, ix, iy, iz, cic_apply(P.Dr, cs), cic_apply(P.Dtheta, cs), cic_apply(P.Dphi, cs), g);
    if (vcorr) *vcorr = cic_apply(P.vfac, cs);
}

// The CIC weights of a lookup factor into x-y (fixed along a row (i, j) of
// modes) and z.  A block stages, for each of its rows, the x-y bilinear sums
// of the four tables at every table z (ng/2 + 1 values per table) in LDS;
// each mode then interpolates in z: two LDS reads per table instead of
// eight table gathers.  Rows row0 .. row0 + nrow - 1 (rows past n^2 unused).
__device__ inline void plt_stage_rows(const PltDev& P, double mr, int n, size_t row0, int nrow, double* sxy) {
    const int ng = P.ng, ncg = ng / 2 + 1;
    for (int v = threadIdx.x; v < nrow * 4 * ncg; v += blockDim.x) {
        const int z = v % ncg, tab = (v / ncg) % 4, rr = v / (4 * ncg);
        const size_t row = row0 + rr;
        if (row >= (size_t)n * n) continue;
        const int i = (int)(row / n), j = (int)(row % n);
        const double x = i * mr, y = j * mr;
        const int ix = (int)floor(x), iy = (int)floor(y);
        const double dx = x - ix, tx = 1.0 - dx, dy = y - iy, ty = 1.0 - dy;
        const int ix1 = (ix + 1) % ng, iy1 = (iy + 1) % ng;
        const double* T = tab == 0 ? P.Dr : tab == 1 ? P.Dtheta : tab == 2 ? P.Dphi : P.vfac;
        auto at = [&](int a, int b) { return T[((size_t)a * ng + b) * ncg + z]; };
        sxy[v] = at(ix, iy) * (tx * ty) + at(ix, iy1) * (tx * dy) + at(ix1, iy) * (dx * ty) + at(ix1, iy1) * (dx * dy);
    }
}
// plt_grad from the staged rows: sxy points at the mode's row
__device__ inline void plt_grad_staged(const double* __restrict__ sxy, int ng, double mr, double c0, int i, int j,
                                       int k, double g[3], double* vcorr) {
    const int ncg = ng / 2 + 1;
    const double iz = k * mr;
    const int z0 = (int)floor(iz), z1 = min(z0 + 1, ncg - 1);
    const double dz = iz - z0, tz = 1.0 - dz;
    auto lerp = [&](int tab) { return sxy[tab * ncg + z0] * tz + sxy[tab * ncg + z1] * dz; };
    plt_combine(ng, c0, i * mr, j * mr, iz, lerp(0), lerp(1), lerp(2), g);
    *vcorr = lerp(3);
}

template <typename T>
__global__ void k_assemble(cx<T>* __restrict__ out, AsmArgs a, double L, size_t ntot) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ntot) return;
    const int k = (int)(idx % a.nc);
    const size_t r = idx / a.nc;
    const int j = (int)(r % a.n1), i = (int)(r / a.n1);
    const bool vel = a.comp >= 3;
    const int d = a.comp % 3, dp = (d + 1) % 3, dpp = (d + 2) % 3;
    double g[3], vc = 1.0;
    if (a.plt.Dr) {
        const double mr = (double)a.plt.ng / (double)a.n0;
        plt_grad(a.plt, mr, 1.0 / (mr * L), i, j, k, g, vel ? &vc : nullptr);
    } else {
        g[0] = fgrad(i, a.n0, a.kf0);
        g[1] = fgrad(j, a.n1, a.kf1);
        g[2] = fgrad(k, a.n2, a.kf2);
    }
    // potential sum in FP64 with a compensated (two-sum) accumulation
    double sr = 0.0, si = 0.0, er = 0.0, ei = 0.0;
    const double cf[3] = {a.c1, a.c2, a.c3};
    for (int o = 0; o < a.order; ++o) {
        const cx<T> v = static_cast<const cx<T>*>(a.phi[o])[idx];
        const double xr = cf[o] * (double)v.x, xi = cf[o] * (double)v.y;
        double t = sr + xr, bb = t - sr;
        er += (sr - (t - bb)) + (xr - bb);
        sr = t;
        t = si + xi; bb = t - si;
        ei += (si - (t - bb)) + (xi - bb);
        si = t;
    }
    sr += er;
    si += ei;
    // (i g_d) * phitot
    double vr = -g[d] * si, vi = g[d] * sr;
    if (a.order > 2) {
        const cx<T> a1 = static_cast<const cx<T>*>(a.A3[dpp])[idx], a2 = static_cast<const cx<T>*>(a.A3[dp])[idx];
        // (i g_d') A3_d'' - (i g_d'') A3_d'
        const double cr = (-g[dp] * (double)a1.y) - (-g[dpp] * (double)a2.y);
        const double ci = (g[dp] * (double)a1.x) - (g[dpp] * (double)a2.x);
        vr += a.cA * cr;
        vi += a.cA * ci;
    }
    if (vel) { vr *= vc; vi *= vc; }
    vr *= a.unit;
    vi *= a.unit;
    if (a.sh0 != 0.0 || a.sh1 != 0.0 || a.sh2 != 0.0) {
        const double kx = ((double)i - (double)(i > a.n0 / 2) * a.n0) * a.kf0;
        const double ky = ((double)j - (double)(j > a.n1 / 2) * a.n1) * a.kf1;
        const double kz = ((double)k - (double)(k > a.n2 / 2) * a.n2) * a.kf2;
        double s, c;
        sincos(a.sh0 * kx + a.sh1 * ky + a.sh2 * kz, &s, &c);
        const double tr = vr * c - vi * s, ti = vr * s + vi * c;
        vr = tr;
        vi = ti;
    }
    if (idx == 0) { vr = 0.0; vi = 0.0; }
    out[idx] = cx<T>{(T)vr, (T)vi};
}

// All six components (and, for bcc, their (1/2,1/2,1/2)-shifted copies) in
// one pass: the potentials are read once, the gradient (Fourier or PLT) is
// evaluated once per mode.  out[comp + 6*sub] (null pointers are skipped).
struct AsmAllArgs {
    const void* phi[3];
    const void* A3[3];
    void* out[12];
    int n, nc, order, nsub;
    double kf, L;
    double c[4], cv[4];       // displacement / velocity coefficients
    double ulen, uvel;        // lunit / L, vunit / L
    double sh;                // sub-lattice shift (1/2) * dx
    PltDev plt;
    double plt_mr, plt_c0;    // ng / n and n / (ng L) of the PLT lookup
};

// ORDER, PLT and NSUB are compile-time so that every field load of the mode
// is issued before the gradient (a CIC gather for PLT) is evaluated.  STAGE:
// PLT lookups from per-row x-y sums in LDS (plt_stage_rows), blocks of
// ASM_BLOCK modes spanning at most asm_stage_rows(nc) rows.
constexpr int ASM_BLOCK = 256;
__host__ __device__ constexpr int asm_stage_rows(int nc) { return (ASM_BLOCK + nc - 1) / nc + 1; }
template <typename T, int ORDER, bool PLT, int NSUB, bool STAGE = false>
__global__ void __launch_bounds__(ASM_BLOCK) k_assemble_all(AsmAllArgs a, size_t ntot) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* sxy = reinterpret_cast<double*>(smem_raw);
    const size_t row0 = (size_t)blockIdx.x * blockDim.x / a.nc;
    if constexpr (STAGE) {
        plt_stage_rows(a.plt, a.plt_mr, a.n, row0, asm_stage_rows(a.nc), sxy);
        __syncthreads();
    }
    if (idx >= ntot) return;
    cx<T> pv[ORDER], av[3];
#pragma unroll
    for (int o = 0; o < ORDER; ++o) pv[o] = static_cast<const cx<T>*>(a.phi[o])[idx];
    if constexpr (ORDER > 2) {
#pragma unroll
        for (int d = 0; d < 3; ++d) av[d] = static_cast<const cx<T>*>(a.A3[d])[idx];
    }
    const int k = (int)(idx % a.nc);
    const size_t r = idx / a.nc;
    const int j = (int)(r % a.n), i = (int)(r / a.n);
    double g[3], vc = 1.0;
    if constexpr (STAGE) {
        plt_grad_staged(sxy + (r - row0) * 4 * (a.plt.ng / 2 + 1), a.plt.ng, a.plt_mr, a.plt_c0, i, j, k, g, &vc);
    } else if constexpr (PLT) {
        plt_grad(a.plt, a.plt_mr, a.plt_c0, i, j, k, g, &vc);
    } else {
        g[0] = fgrad(i, a.n, a.kf);
        g[1] = fgrad(j, a.n, a.kf);
        g[2] = fgrad(k, a.n, a.kf);
    }
    // compensated sums of c_o phi_o for displacement and velocity
    auto csum = [&](const double* cf, double& sr, double& si) {
        sr = si = 0.0;
        double er = 0.0, ei = 0.0;
#pragma unroll
        for (int o = 0; o < ORDER; ++o) {
            const double xr = cf[o] * (double)pv[o].x, xi = cf[o] * (double)pv[o].y;
            double t = sr + xr, bb = t - sr;
            er += (sr - (t - bb)) + (xr - bb);
            sr = t;
            t = si + xi; bb = t - si;
            ei += (si - (t - bb)) + (xi - bb);
            si = t;
        }
        sr += er;
        si += ei;
    };
    double dr, di, vr, vi;
    csum(a.c, dr, di);
    csum(a.cv, vr, vi);
    double ph_c = 1.0, ph_s = 0.0;
    if constexpr (NSUB > 1) {
        const double kx = ((double)i - (double)(i > a.n / 2) * a.n) * a.kf;
        const double ky = ((double)j - (double)(j > a.n / 2) * a.n) * a.kf;
        const double kz = ((double)k - (double)(k > a.n / 2) * a.n) * a.kf;
        sincos(a.sh * kx + a.sh * ky + a.sh * kz, &ph_s, &ph_c);
    }
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) {
        const bool vel = comp >= 3;
        const int d = comp % 3, dp = (d + 1) % 3, dpp = (d + 2) % 3;
        const double sr = vel ? vr : dr, si = vel ? vi : di;
        double xr = -g[d] * si, xi = g[d] * sr;
        if constexpr (ORDER > 2) {
            const double cr = (-g[dp] * (double)av[dpp].y) - (-g[dpp] * (double)av[dp].y);
            const double ci = (g[dp] * (double)av[dpp].x) - (g[dpp] * (double)av[dp].x);
            const double cA = vel ? a.cv[3] : a.c[3];
            xr += cA * cr;
            xi += cA * ci;
        }
        if (vel) { xr *= vc; xi *= vc; }
        const double u = vel ? a.uvel : a.ulen;
        xr *= u;
        xi *= u;
        if (idx == 0) { xr = 0.0; xi = 0.0; }
        if (a.out[comp]) static_cast<cx<T>*>(a.out[comp])[idx] = cx<T>{(T)xr, (T)xi};
        if (NSUB > 1 && a.out[comp + 6])
            static_cast<cx<T>*>(a.out[comp + 6])[idx] = cx<T>{(T)(xr * ph_c - xi * ph_s), (T)(xr * ph_s + xi * ph_c)};
    }
}

// three real-space components -> interleaved [ipart][3] FP64 (positions add
// the lattice coordinate ((ijk + shift)/N) * lunit)
template <typename T>
__global__ void k_readout3(const T* __restrict__ f0, const T* __restrict__ f1, const T* __restrict__ f2,
                           double* __restrict__ dst, int n, int is_pos, double lunit, double shift, size_t pbase) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t np = (size_t)n * n * n;
    if (p >= np) return;
    const int k = (int)(p % n);
    const size_t r = p / n;
    const int j = (int)(r % n), i = (int)(r / n);
    const size_t src = r * (n + 2) + k;
    double v0 = (double)f0[src], v1 = (double)f1[src], v2 = (double)f2[src];
    if (is_pos) {
        v0 = (((double)i + shift) / (double)n) * lunit + v0;
        v1 = (((double)j + shift) / (double)n) * lunit + v1;
        v2 = (((double)k + shift) / (double)n) * lunit + v2;
    }
    double* o = dst + 3 * (pbase + p);
    o[0] = v0;
    o[1] = v1;
    o[2] = v2;
}

// particle value = q + field (positions, FP64) or field (velocities)
template <typename T>
__global__ void k_readout(const T* __restrict__ f, double* __restrict__ dst, int n, int idim, int is_pos, double lunit,
                          double shift, size_t pbase) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t np = (size_t)n * n * n;
    if (p >= np) return;
    const int k = (int)(p % n);
    const size_t r = p / n;
    const int j = (int)(r % n), i = (int)(r / n);
    double v = (double)f[r * (n + 2) + k];
    if (is_pos) {
        const int q = idim == 0 ? i : idim == 1 ? j : k;
        v = (((double)q + shift) / (double)n) * lunit + v;
    }
    dst[3 * (pbase + p) + idim] = v;
}

// ------------------------------------------------------------- host -----

void assemble_component(mfic_ctx ctx, int prec, const size_t n[3], double L, int order, int comp, const double cf[4],
                        double unit, const PltDev& plt, const void* const phi[3], const void* const A3[3],
                        void* out, const double shift_dx[3]) {
    AsmArgs a{};
    for (int o = 0; o < 3; ++o) { a.phi[o] = phi[o]; a.A3[o] = A3[o]; }
    a.n0 = (int)n[0]; a.n1 = (int)n[1]; a.n2 = (int)n[2]; a.nc = (int)(n[2] / 2 + 1);
    a.order = order;
    a.comp = comp;
    const double pi2 = 2.0 * 3.14159265358979323846;
    a.kf0 = pi2 / L; a.kf1 = pi2 / L; a.kf2 = pi2 / L;
    a.c1 = cf[0]; a.c2 = cf[1]; a.c3 = cf[2]; a.cA = cf[3];
    a.unit = unit;
    a.sh0 = shift_dx ? shift_dx[0] : 0.0;
    a.sh1 = shift_dx ? shift_dx[1] : 0.0;
    a.sh2 = shift_dx ? shift_dx[2] : 0.0;
    a.plt = plt;
    const size_t ntot = n[0] * n[1] * (n[2] / 2 + 1);
    ProfScope prof(ctx, plt.Dr ? "assemble_plt" : "assemble",
                   (double)ntot * (prec == MFIC_F64 ? 16 : 8) * (order + (order > 2 ? 2 : 0) + 1));
    if (prec == MFIC_F64)
        hipLaunchKernelGGL(k_assemble<double>, dim3(nblk(ntot, 256)), dim3(256), 0, ctx->stream, (cx<double>*)out, a, L, ntot);
    else
        hipLaunchKernelGGL(k_assemble<float>, dim3(nblk(ntot, 256)), dim3(256), 0, ctx->stream, (cx<float>*)out, a, L, ntot);
    MFIC_LAUNCH_CHECK();
}

void assemble_all(mfic_ctx ctx, int prec, size_t n, double L, int order, const double c[4], const double cv[4],
                  double ulen, double uvel, const PltDev& plt, const void* const phi[3], const void* const A3[3],
                  void* const out[12], int nsub) {
    AsmAllArgs a{};
    for (int o = 0; o < 3; ++o) { a.phi[o] = phi[o]; a.A3[o] = A3[o]; }
    int nout = 0;
    for (int q = 0; q < 12; ++q) { a.out[q] = out[q]; nout += out[q] != nullptr; }
    a.n = (int)n; a.nc = (int)(n / 2 + 1); a.order = order; a.nsub = nsub;
    a.kf = 2.0 * 3.14159265358979323846 / L;
    a.L = L;
    for (int q = 0; q < 4; ++q) { a.c[q] = c[q]; a.cv[q] = cv[q]; }
    a.ulen = ulen; a.uvel = uvel;
    a.sh = 0.5 * L / (double)n;
    a.plt = plt;
    a.plt_mr = (double)plt.ng / (double)n;
    a.plt_c0 = plt.ng ? 1.0 / (a.plt_mr * L) : 0.0;
    const size_t ntot = n * n * (n / 2 + 1);
    ProfScope prof(ctx, plt.Dr ? "assemble_all_plt" : "assemble_all",
                   (double)ntot * (prec == MFIC_F64 ? 16 : 8) * (order + (order > 2 ? 3 : 0) + nout));
    MFIC_CHECK(order >= 1 && order <= 3 && (nsub == 1 || nsub == 2), MFIC_ERR_ARG, "assembly: bad order or sub-lattices");
    const int key = (order - 1) * 4 + (plt.Dr ? 2 : 0) + (nsub - 1);
    // staged PLT lookups where a block's rows are few (long rows)
    const size_t stage_lds = (size_t)asm_stage_rows(a.nc) * 4 * (plt.ng / 2 + 1) * sizeof(double);
    const bool stage = plt.Dr && a.nc >= 128 && stage_lds <= 32768;
    const dim3 grid(nblk(ntot, ASM_BLOCK)), block(ASM_BLOCK);
    switch (key) {
#define ASM(O, P, S)                                                                                              \
    case (O - 1) * 4 + (P ? 2 : 0) + (S - 1):                                                                     \
        if (P && stage) {                                                                                         \
            if (prec == MFIC_F64)                                                                                 \
                hipLaunchKernelGGL((k_assemble_all<double, O, P, S, P>), grid, block, stage_lds, ctx->stream, a, ntot); \
            else                                                                                                  \
                hipLaunchKernelGGL((k_assemble_all<float, O, P, S, P>), grid, block, stage_lds, ctx->stream, a, ntot); \
        } else if (prec == MFIC_F64) {                                                                            \
            hipLaunchKernelGGL((k_assemble_all<double, O, P, S>), grid, block, 0, ctx->stream, a, ntot);           \
        } else {                                                                                                  \
            hipLaunchKernelGGL((k_assemble_all<float, O, P, S>), grid, block, 0, ctx->stream, a, ntot);            \
        }                                                                                                         \
        break;
        ASM(1, false, 1) ASM(1, false, 2) ASM(1, true, 1) ASM(1, true, 2)
        ASM(2, false, 1) ASM(2, false, 2) ASM(2, true, 1) ASM(2, true, 2)
        ASM(3, false, 1) ASM(3, false, 2) ASM(3, true, 1) ASM(3, true, 2)
#undef ASM
    }
    MFIC_LAUNCH_CHECK();
}

void lattice_readout3(mfic_ctx ctx, int prec, const void* const f[3], int n, int is_pos, double lunit, double shift,
                      size_t pbase, double* dst) {
    const size_t np = (size_t)n * n * n;
    ProfScope prof(ctx, "readout3", (double)np * 3 * ((prec == MFIC_F64 ? 8 : 4) + 8));
    if (prec == MFIC_F64)
        hipLaunchKernelGGL(k_readout3<double>, dim3(nblk(np, 256)), dim3(256), 0, ctx->stream, (const double*)f[0],
                           (const double*)f[1], (const double*)f[2], dst, n, is_pos, lunit, shift, pbase);
    else
        hipLaunchKernelGGL(k_readout3<float>, dim3(nblk(np, 256)), dim3(256), 0, ctx->stream, (const float*)f[0],
                           (const float*)f[1], (const float*)f[2], dst, n, is_pos, lunit, shift, pbase);
    MFIC_LAUNCH_CHECK();
}

void lattice_readout_raw(mfic_ctx ctx, int prec, const void* field, int n, int idim, int is_pos, double lunit,
                         double shift, size_t pbase, double* dst) {
    const size_t np = (size_t)n * n * n;
    ProfScope prof(ctx, "readout", (double)np * ((prec == MFIC_F64 ? 8 : 4) + 8));
    if (prec == MFIC_F64)
        hipLaunchKernelGGL(k_readout<double>, dim3(nblk(np, 256)), dim3(256), 0, ctx->stream, (const double*)field, dst, n,
                           idim, is_pos, lunit, shift, pbase);
    else
        hipLaunchKernelGGL(k_readout<float>, dim3(nblk(np, 256)), dim3(256), 0, ctx->stream, (const float*)field, dst, n,
                           idim, is_pos, lunit, shift, pbase);
    MFIC_LAUNCH_CHECK();
}

}  // namespace mfic
