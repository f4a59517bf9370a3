// PLT lattice corrections (particle_plt.hh:50-587), tables built once per
// lattice type on an ng^3 unit-box mesh:
//   device: Ewald-split dynamical matrix per site (FP64, one thread per site,
//           particle_plt.hh:160-273) and the unitary r2c of its six
//           components with this library's FFT (particle_plt.hh:276-288);
//   host:   per-mode symmetric eigen-decomposition, folding into the first
//           Brillouin zone (first key insertion wins, i.e. the one-thread
//           order of the reference's critical-section map insert,
//           particle_plt.hh:333-395), neighbour averaging of off-lattice
//           modes (:406-459) and projection on the spherical basis of k
//           (:461-499).
// The per-mode lookup used by the assembly kernel lives in lpt.hip.
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "fft_core.hpp"
#include "lpt.hpp"
#include "mfic_internal.hpp"

namespace mfic {

namespace {
constexpr double kPi = 3.14159265358979323846;

struct Lat {
    double B[9], R[9];
    int IR[9];
    int charge_fac;
    std::vector<std::array<double, 3>> normals;
};

Lat lattice_def(int lat) {
    const double tp = 2.0 * kPi, fp = 4.0 * kPi, p = kPi;
    Lat L{};
    if (lat == 0) {
        const double B[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, R[9] = {tp, 0, 0, 0, tp, 0, 0, 0, tp};
        const int IR[9] = {2, 0, 0, 0, 2, 0, 0, 0, 2};
        std::memcpy(L.B, B, sizeof B); std::memcpy(L.R, R, sizeof R); std::memcpy(L.IR, IR, sizeof IR);
        L.charge_fac = 1;
        L.normals = {{p, 0, 0}, {-p, 0, 0}, {0, p, 0}, {0, -p, 0}, {0, 0, p}, {0, 0, -p}};
    } else if (lat == 1) {
        const double B[9] = {1, 0, 0.5, 0, 1, 0.5, 0, 0, 0.5}, R[9] = {tp, 0, 0, 0, tp, 0, -tp, -tp, fp};
        const int IR[9] = {2, 0, 0, 0, 2, 0, 1, 1, 1};
        std::memcpy(L.B, B, sizeof B); std::memcpy(L.R, R, sizeof R); std::memcpy(L.IR, IR, sizeof IR);
        L.charge_fac = 2;
        L.normals = {{0, p, p}, {0, -p, p}, {0, p, -p}, {0, -p, -p}, {p, 0, p}, {-p, 0, p},
                     {p, 0, -p}, {-p, 0, -p}, {p, p, 0}, {-p, p, 0}, {p, -p, 0}, {-p, -p, 0}};
    } else {
        const double B[9] = {0, 0.5, 0, 0.5, 0, 1, 0.5, 0.5, 0}, R[9] = {-fp, fp, tp, 0, 0, tp, fp, 0, -tp};
        const int IR[9] = {0, 1, 1, 1, 0, 1, 0, 2, 0};
        std::memcpy(L.B, B, sizeof B); std::memcpy(L.R, R, sizeof R); std::memcpy(L.IR, IR, sizeof IR);
        L.charge_fac = 4;
        L.normals = {{tp, 0, 0}, {-tp, 0, 0}, {0, tp, 0}, {0, -tp, 0}, {0, 0, tp}, {0, 0, -tp},
                     {p, p, p}, {p, p, -p}, {p, -p, p}, {p, -p, -p}, {-p, p, p}, {-p, p, -p}, {-p, -p, p}, {-p, -p, -p}};
    }
    return L;
}

struct EwaldArgs {
    double B[9], R[9];
    int ng;
    double alpha, charge, fft_norm;
};

__device__ inline void matvec(const double* A, const double* v, double* o) {
    for (int m = 0; m < 3; ++m) o[m] = A[3 * m] * v[0] + A[3 * m + 1] * v[1] + A[3 * m + 2] * v[2];
}

// one lattice site: real-space images (7^3) + reciprocal vectors (7^3 - 1);
// writes the 6 independent components into 6 real-space ng^3 grids
__global__ void k_ewald(double* __restrict__ D, size_t comp_stride, EwaldArgs a) {
    const int site = blockIdx.x * blockDim.x + threadIdx.x;
    const int ng = a.ng;
    if (site >= ng * ng * ng) return;
    const int k = site % ng, j = (site / ng) % ng, i = site / (ng * ng);
    const double dx = 1.0 / ng;
    double x[3] = {dx * i, dx * j, dx * k}, ar[3];
    matvec(a.B, x, ar);
    for (int c = 0; c < 3; ++c) ar[c] = fmod(2 * 1.0 + ar[c], 1.0);
    const double alpha = a.alpha, alpha2 = alpha * alpha, alpha3 = alpha2 * alpha;
    const double pi32 = pow(kPi, 1.5), sqrtpi = sqrt(kPi);
    double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int ix = -3; ix <= 3; ix++)
        for (int iy = -3; iy <= 3; iy++)
            for (int iz = -3; iz <= 3; iz++) {
                const double nn[3] = {(double)ix, (double)iy, (double)iz};
                double bn[3], d[3];
                matvec(a.B, nn, bn);
                for (int c = 0; c < 3; ++c) d[c] = ar[c] - bn[c];
                const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (r < 1e-14) continue;
                const double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
                const double ex = exp(-alpha2 * r2);
                const double K1 = -alpha3 / pi32 * ex / r2;
                const double K2 = (erfc(alpha * r) + 2.0 * alpha / sqrtpi * ex * r) / (4.0 * kPi);
                for (int mu = 0; mu < 3; ++mu)
                    for (int nu = mu; nu < 3; ++nu) {
                        const double dd = d[mu] * d[nu] * K1 + ((mu == nu ? 1.0 : 0.0) / r3 - 3.0 * (d[mu] * d[nu]) / r5) * K2;
                        M[3 * mu + nu] += dd;
                        if (mu != nu) M[3 * nu + mu] += dd;
                    }
            }
    for (int ix = -3; ix <= 3; ix++)
        for (int iy = -3; iy <= 3; iy++)
            for (int iz = -3; iz <= 3; iz++) {
                if (abs(ix) + abs(iy) + abs(iz) == 0) continue;
                const double kn[3] = {(double)ix / ng, (double)iy / ng, (double)iz / ng};
                double ak[3];
                matvec(a.R, kn, ak);
                const double k2 = ak[0] * ak[0] + ak[1] * ak[1] + ak[2] * ak[2];
                const double term = exp(-k2 / (4 * alpha2)) * cos(ak[0] * ar[0] + ak[1] * ar[1] + ak[2] * ar[2]) / k2 * a.fft_norm;
                for (int mu = 0; mu < 3; ++mu)
                    for (int nu = mu; nu < 3; ++nu) {
                        const double dd = ak[mu] * ak[nu] * term;
                        M[3 * mu + nu] += dd;
                        if (mu != nu) M[3 * nu + mu] += dd;
                    }
            }
    const size_t id = ((size_t)i * ng + j) * (ng + 2) + k;
    const int comps[6] = {0, 1, 2, 4, 5, 8};  // xx xy xz yy yz zz
    for (int c = 0; c < 6; ++c) {
        double v = M[comps[c]] * a.charge;
        if (site == 0) v = (c == 0 || c == 3 || c == 5) ? 1.0 / 3.0 : 0.0;  // r = 0 background
        D[c * comp_stride + id] = v;
    }
}

// symmetric 3x3 eigen-decomposition (cyclic Jacobi), ascending eigenvalues,
// eigenvectors as columns
void jacobi3(const double Ain[9], double ev[3], double V[9]) {
    double A[9];
    std::memcpy(A, Ain, sizeof A);
    for (int q = 0; q < 9; ++q) V[q] = (q % 4 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        if (std::fabs(A[1]) + std::fabs(A[2]) + std::fabs(A[5]) == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double apq = A[3 * p + q];
                if (apq == 0.0) continue;
                const double th = (A[3 * q + q] - A[3 * p + p]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double x = A[3 * k + p], y = A[3 * k + q];
                    A[3 * k + p] = c * x - s * y;
                    A[3 * k + q] = s * x + c * y;
                }
                for (int k = 0; k < 3; ++k) {
                    const double x = A[3 * p + k], y = A[3 * q + k];
                    A[3 * p + k] = c * x - s * y;
                    A[3 * q + k] = s * x + c * y;
                }
                for (int k = 0; k < 3; ++k) {
                    const double x = V[3 * k + p], y = V[3 * k + q];
                    V[3 * k + p] = c * x - s * y;
                    V[3 * k + q] = s * x + c * y;
                }
            }
    }
    int o[3] = {0, 1, 2};
    const double e[3] = {A[0], A[4], A[8]};
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (e[o[b]] < e[o[a]]) std::swap(o[a], o[b]);
    double W[9];
    for (int c = 0; c < 3; ++c) {
        ev[c] = e[o[c]];
        for (int r = 0; r < 3; ++r) W[3 * r + c] = V[3 * r + o[c]];
    }
    std::memcpy(V, W, sizeof W);
}
}  // namespace

PltTables::~PltTables() {
    if (dev) (void)hipFree(dev);
}

void plt_build(mfic_ctx ctx, int lat, int ng, PltTables& out) {
    MFIC_CHECK(lat >= 0 && lat <= 2, MFIC_ERR_CONFIG, "PLT lattice must be sc, bcc or fcc");
    MFIC_CHECK(ng % 2 == 0 && fft_length_supported(ng) && fft_length_supported(ng / 2), MFIC_ERR_SIZE,
               "unsupported PLT mesh size");
    const Lat L = lattice_def(lat);
    const int nc = ng / 2 + 1;
    const size_t nr = (size_t)ng * ng * (ng + 2), nk = (size_t)ng * ng * nc;
    double* dD = nullptr;
    MFIC_HIP(hipMalloc(&dD, 6 * nr * sizeof(double)));
    MFIC_HIP(hipMemsetAsync(dD, 0, 6 * nr * sizeof(double), ctx->stream));
    EwaldArgs a{};
    std::memcpy(a.B, L.B, sizeof a.B);
    std::memcpy(a.R, L.R, sizeof a.R);
    a.ng = ng;
    a.alpha = 1.0 / std::sqrt(2) / 4.0;
    a.charge = 1.0 / std::pow((double)ng, 3) / L.charge_fac;
    a.fft_norm = 1.0 / std::pow((double)ng, 3.0);
    const int sites = ng * ng * ng;
    hipLaunchKernelGGL(k_ewald, dim3((sites + 127) / 128), dim3(128), 0, ctx->stream, dD, nr, a);
    MFIC_LAUNCH_CHECK();
    const size_t nn[3] = {(size_t)ng, (size_t)ng, (size_t)ng};
    const double norm = 1.0 / std::sqrt((double)ng * ng * ng);
    for (int c = 0; c < 6; ++c) fft_r2c<double>(ctx, dD + c * nr, nn, norm);
    std::vector<double> hD(6 * nr);
    MFIC_HIP(hipMemcpyAsync(hD.data(), dD, 6 * nr * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MFIC_HIP(hipStreamSynchronize(ctx->stream));
    MFIC_HIP(hipFree(dD));

    const double fft_norm12 = 1.0 / std::pow((double)ng, 1.5);
    auto sidx = [ng](int q) { return (double)q - (double)(q > ng / 2) * ng; };
    std::vector<std::array<double, 6>> t6(nk);  // l0 l1 l2 v0 v1 v2
    const int KR = ng + 4, KS = 2 * KR + 1;  // FBZ keys satisfy |key| <= ng
    std::vector<int> map((size_t)KS * KS * KS, -1);
    auto key = [KR, KS](int x, int y, int z) { return ((size_t)(x + KR) * KS + (y + KR)) * KS + (z + KR); };
    for (int i = 0; i < ng; i++)
        for (int j = 0; j < ng; j++)
            for (int k = 0; k < nc; k++) {
                const size_t q = ((size_t)i * ng + j) * nc + k;
                auto re = [&](int c) { return hD[c * nr + 2 * q] / fft_norm12; };
                const double A[9] = {re(0), re(1), re(2), re(1), re(3), re(4), re(2), re(4), re(5)};
                double ev[3], V[9];
                if (q == 0) {
                    // k = 0: D is isotropic in exact arithmetic and any basis
                    // diagonalises it; use the Cartesian one (largest-
                    // eigenvalue vector e_z) instead of rounding noise
                    for (int c = 0; c < 9; ++c) V[c] = (c % 4 == 0) ? 1.0 : 0.0;
                    ev[0] = ev[1] = ev[2] = (A[0] + A[4] + A[8]) / 3.0;
                } else {
                    jacobi3(A, ev, V);
                }
                const double s = 2.0 * kPi * ng;
                const double e3[3] = {V[2] / s, V[5] / s, V[8] / s};
                const double arv[3] = {sidx(i) * 2.0 * kPi / s, sidx(j) * 2.0 * kPi / s, sidx(k) * 2.0 * kPi / s};
                double av[3];
                for (int m = 0; m < 3; ++m) av[m] = L.R[3 * m] * arv[0] + L.R[3 * m + 1] * arv[1] + L.R[3 * m + 2] * arv[2];
                for (int l1 = -1; l1 <= 1; ++l1)
                    for (int l2 = -1; l2 <= 1; ++l2)
                        for (int l3 = -1; l3 <= 1; ++l3)
                            for (int is = 0; is <= 1; ++is) {
                                const double sg = 2.0 * is - 1.0;
                                const double ls[3] = {(double)l1, (double)l2, (double)l3};
                                double vk[3];
                                for (int m = 0; m < 3; ++m)
                                    vk[m] = sg * av[m] + (L.R[3 * m] * ls[0] + L.R[3 * m + 1] * ls[1] + L.R[3 * m + 2] * ls[2]);
                                bool inside = true;
                                for (const auto& nv : L.normals) {
                                    const double nd = nv[0] * vk[0] + nv[1] * vk[1] + nv[2] * vk[2];
                                    const double n2 = nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2];
                                    if (nd > 1.0001 * n2) { inside = false; break; }
                                }
                                if (!inside) continue;
                                const int ix = (int)std::round(vk[0] * ng / (2.0 * kPi));
                                const int iy = (int)std::round(vk[1] * ng / (2.0 * kPi));
                                const int iz = (int)std::round(vk[2] * ng / (2.0 * kPi));
                                if (std::abs(ix) <= KR && std::abs(iy) <= KR && std::abs(iz) <= KR) {
                                    int& slot = map[key(ix, iy, iz)];
                                    if (slot < 0) slot = (int)q;
                                }
                                t6[q] = {ev[2], ev[1], ev[0], e3[0], e3[1], e3[2]};
                            }
            }
    std::vector<double> tab(4 * nk);
    double *Dr = tab.data(), *Dt = Dr + nk, *Dp = Dt + nk, *Vf = Dp + nk;
    for (int i = 0; i < ng; i++)
        for (int j = 0; j < ng; j++)
            for (int k = 0; k < nc; k++) {
                const size_t q = ((size_t)i * ng + j) * nc + k;
                const int ii = i > ng / 2 ? i - ng : i, jj = j > ng / 2 ? j - ng : j, kk = k > ng / 2 ? k - ng : k;
                const double kvi[3] = {(double)ii, (double)jj, (double)kk};
                const int v0 = L.IR[0] * i + L.IR[1] * j + L.IR[2] * k, v1 = L.IR[3] * i + L.IR[4] * j + L.IR[5] * k,
                          v2 = L.IR[6] * i + L.IR[7] * j + L.IR[8] * k;
                const bool on = (v0 % 2 == 0) && (v1 % 2 == 0) && (v2 % 2 == 0);
                const int nb[6][3] = {{ii - 1, jj, kk}, {ii + 1, jj, kk}, {ii, jj - 1, kk},
                                      {ii, jj + 1, kk}, {ii, jj, kk - 1}, {ii, jj, kk + 1}};
                double l[3] = {0, 0, 0}, v[3] = {0, 0, 0};
                int cnt = 0;
                for (int m = 0; m < (on ? 1 : 6); ++m) {
                    const int x = on ? ii : nb[m][0], y = on ? jj : nb[m][1], z = on ? kk : nb[m][2];
                    if (std::abs(x) > KR || std::abs(y) > KR || std::abs(z) > KR) continue;
                    const int r = map[key(x, y, z)];
                    if (r < 0) continue;
                    ++cnt;
                    const auto& e = t6[r];
                    const double dot = e[3] * kvi[0] + e[4] * kvi[1] + e[5] * kvi[2];
                    const double sg = dot < 0.0 ? -1.0 : 1.0;
                    for (int c = 0; c < 3; ++c) { l[c] += e[c]; v[c] += e[3 + c] * sg; }
                }
                if (!on)
                    for (int c = 0; c < 3; ++c) { l[c] /= (double)cnt; v[c] /= (double)cnt; }
                const double kv[3] = {sidx(i) * 2.0 * kPi, sidx(j) * 2.0 * kPi, sidx(k) * 2.0 * kPi};
                const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                const double e1[3] = {v[0] / vn, v[1] / vn, v[2] / vn};
                const double kr = std::sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
                const double kphi = kr > 0.0 ? std::atan2(kv[1], kv[0]) : 0.0;
                const double kth = kr > 0.0 ? std::acos(kv[2] / kr) : 0.0;
                const double st = std::sin(kth), ct = std::cos(kth), sp = std::sin(kphi), cp = std::cos(kphi);
                const double er[3] = {st * cp, st * sp, ct}, et[3] = {ct * cp, ct * sp, -st}, ep[3] = {-sp, cp, 0.0};
                double ren = e1[0] * er[0] + e1[1] * er[1] + e1[2] * er[2];
                if (ren < 0.01) ren = 1.0;
                Dr[q] = 1.0;
                Dt[q] = (e1[0] * et[0] + e1[1] * et[1] + e1[2] * et[2]) / ren;
                Dp[q] = (e1[0] * ep[0] + e1[1] * ep[1] + e1[2] * ep[2]) / ren;
                Vf[q] = 1.0 / (0.25 * (std::sqrt(1. + 24 * l[0]) - 1.));
            }
    Vf[0] = 1.0; Dr[0] = 1.0; Dt[0] = 0.0; Dp[0] = 0.0;
    if (out.dev) MFIC_HIP(hipFree(out.dev));
    MFIC_HIP(hipMalloc(&out.dev, 4 * nk * sizeof(double)));
    MFIC_HIP(hipMemcpy(out.dev, tab.data(), 4 * nk * sizeof(double), hipMemcpyHostToDevice));
    out.host = std::move(tab);
    out.ng = ng;
    out.lattice = lat;
    out.view.ng = ng;
    out.view.Dr = out.dev;
    out.view.Dtheta = out.dev + nk;
    out.view.Dphi = out.dev + 2 * nk;
    out.view.vfac = out.dev + 3 * nk;
}

}  // namespace mfic
