// Cosmology scalars and the Eisenstein & Hu transfer function, shared by the
// host (normalisation integral) and the device (per-mode phi1 amplitude).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/mfic.h"

namespace mfic {

// Eisenstein & Hu (1998) fit scalars (transfer_eisenstein.cc:58-122) plus the
// primordial slope / running and the amplitude normalisation.
struct EhParams {
    double keq, s, ksilk, ac, bc, ab, bnode, bb;  // fit scalars, Mpc units
    double h, fb, fc;                             // k[h/Mpc] -> k[1/Mpc], fractions
    double alpha_s, beta_s, k_p, n_s, sqrtpnorm;
    // tabulated transfer function (tk_n > 0): tk_n nodes of 5 doubles
    // (ln k_i, then the cubic a b c d in t = ln k - ln k_i on the interval
    // starting there), in the memory of the side that evaluates it
    const double* tk;
    int tk_n;
    double tk_lkmin;  // ln k[1]: smaller k are clamped to it
};

// ln T at x = ln k on the tabulated spline (the interval holding x, found by
// bisection; x at or beyond the last node uses the last interval)
__host__ __device__ inline double tk_spline_eval(const double* tk, int n, double x) {
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tk[5 * mid] > x) hi = mid;
        else lo = mid;
    }
    const double* c = tk + 5 * lo;
    const double t = x - c[0];
    return c[1] + t * (c[2] + t * (c[3] + t * c[4]));
}

// running of the spectral index (cosmology_parameters.cc:154-169), k in the
// units the caller's plugin passes
__host__ __device__ inline double running_factor(const EhParams& e, double k) {
    if (e.alpha_s == 0.0 && e.beta_s == 0.0) return 1.0;
    const double lnk = log(k / e.k_p);
    return pow(k / e.k_p, 0.5 * (0.5 * (e.alpha_s * lnk + e.beta_s * lnk * lnk / 3.0)));
}

// tabulated T(k), k in h/Mpc (transfer_CAMB_file.cc:195-199, 244-245)
__host__ __device__ inline double tk_transfer(const EhParams& e, double kh) {
    const double x = fmax(log(kh), e.tk_lkmin);
    return exp(tk_spline_eval(e.tk, e.tk_n, x)) * running_factor(e, exp(x) * e.h);
}

// T(k) for k in h/Mpc: fb T_b + fc T_c times the running factor
// (transfer_eisenstein.cc:125-201, 238-242, cosmology_parameters.cc:154-169),
// or the tabulated transfer function when one is set
__host__ __device__ inline double eh_transfer(const EhParams& e, double kh) {
    if (e.tk_n > 0) return tk_transfer(e, kh);
    const double k = fabs(kh * e.h);
    double tb = 1.0, tc = 1.0;
    if (k != 0.0) {
        const double q = k / 13.41 / e.keq, xx = k * e.s;
        const double lnb = log(2.718282 + 1.8 * e.bc * q), lnnb = log(2.718282 + 1.8 * q);
        const double pq = pow(q, 1.08);
        const double Ca = 14.2 / e.ac + 386.0 / (1 + 69.9 * pq), Cna = 14.2 + 386.0 / (1 + 69.9 * pq);
        const double x54 = xx / 5.4, f = 1.0 / (1.0 + (x54 * x54) * (x54 * x54));
        tc = f * lnb / (lnb + Cna * (q * q)) + (1 - f) * lnb / (lnb + Ca * (q * q));
        const double bx = e.bnode / xx;
        const double xt = k * (e.s * pow(1. + bx * bx * bx, -1. / 3.));
        const double T0 = lnnb / (lnnb + Cna * (q * q));
        const double x52 = xx / 5.2, bbx = e.bb / xx;
        tb = sin(xt) / xt * (T0 / (1. + x52 * x52) + e.ab / (1. + bbx * bbx * bbx) * exp(-pow(k / e.ksilk, 1.4)));
    }
    return (e.fb * tb + e.fc * tc) * running_factor(e, kh);
}

// get_amplitude (cosmology_calculator.hh:429-432): k^{n_s/2} T(k) sqrt(pnorm)
__host__ __device__ inline double eh_amplitude(const EhParams& e, double kh) {
    return pow(kh, 0.5 * e.n_s) * eh_transfer(e, kh) * e.sqrtpnorm;
}

EhParams eh_setup(const mfic_cosmo_params& p);
// spline nodes of a tabulated transfer function (5 doubles per node, the
// layout tk_spline_eval reads); throws std::invalid_argument on a table that
// is too short, not increasing in k, or not positive
std::vector<double> tk_spline_build(const double* k, const double* T, size_t n);
// largest |k| (h/Mpc) the table can evaluate
inline double tk_kmax(const mfic_cosmo_params& p) { return p.tk_n > 0 ? p.tk_k[p.tk_n - 1] : INFINITY; }
// device copy of a table's spline nodes, owned by the caller (hipFree);
// sets eh->tk / tk_n / tk_lkmin.  No allocation without a table.
void* tk_upload(const mfic_cosmo_params& p, EhParams* eh);
void cosmo_defaults(mfic_cosmo_params* p);
void cosmo_compute(const mfic_cosmo_params& p, mfic_cosmo_scalars* s, EhParams* eh);
void cosmo_eval(const mfic_cosmo_params& p, const std::string& what, const double* x, double* y, size_t n);

}  // namespace mfic
