// Cosmology scalars and the Eisenstein & Hu transfer function, shared by the
// host (normalisation integral) and the device (per-mode phi1 amplitude).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/mfic.h"

namespace mfic {

// Eisenstein & Hu (1998) fit scalars (transfer_eisenstein.cc:58-122) plus the
// primordial slope / running and the amplitude normalisation.
struct EhParams {
    double keq, s, ksilk, ac, bc, ab, bnode, bb;  // fit scalars, Mpc units
    double h, fb, fc;                             // k[h/Mpc] -> k[1/Mpc], fractions
    double alpha_s, beta_s, k_p, n_s, sqrtpnorm;
};

// T(k) for k in h/Mpc: fb T_b + fc T_c times the running factor
// (transfer_eisenstein.cc:125-201, cosmology_parameters.cc:154-169)
__host__ __device__ inline double eh_transfer(const EhParams& e, double kh) {
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
    double run = 1.0;
    if (!(e.alpha_s == 0.0 && e.beta_s == 0.0)) {
        const double lnk = log(kh / e.k_p);
        run = pow(kh / e.k_p, 0.5 * (0.5 * (e.alpha_s * lnk + e.beta_s * lnk * lnk / 3.0)));
    }
    return (e.fb * tb + e.fc * tc) * run;
}

// get_amplitude (cosmology_calculator.hh:429-432): k^{n_s/2} T(k) sqrt(pnorm)
__host__ __device__ inline double eh_amplitude(const EhParams& e, double kh) {
    return pow(kh, 0.5 * e.n_s) * eh_transfer(e, kh) * e.sqrtpnorm;
}

EhParams eh_setup(const mfic_cosmo_params& p);
void cosmo_defaults(mfic_cosmo_params* p);
void cosmo_compute(const mfic_cosmo_params& p, mfic_cosmo_scalars* s, EhParams* eh);
void cosmo_eval(const mfic_cosmo_params& p, const std::string& what, const double* x, double* y, size_t n);

}  // namespace mfic
