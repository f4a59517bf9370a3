/*
 * compat_check.cpp -- test program for include/mfic_compat.hpp: the LPT
 * source block of ic_generator::run written against the reference's class
 * interfaces (Grid_FFT, RNG_ngenic, cosmology calculator, OrszagConvolver,
 * op::), built as a shared library whose entry points the GPU tests call
 * and compare with the oracle's potentials.
 */
#include <cmath>
#include <string>

#include "mfic_compat.hpp"

using namespace mfic_compat;

namespace {

std::string g_err;

template <typename real_t>
void potentials(size_t ngrid, double boxlen, long seed, int LPTorder, const mfic_cosmo_params& cp, double* out) {
    using ccomplex_t = std::complex<real_t>;
    calculator the_cosmo_calc(cp);

    // ic_generator.cc:245-266
    const real_t astart = 1.0 / (1.0 + cp.zstart);
    const real_t volfac(std::pow(boxlen / ngrid / 2.0 / M_PI, 1.5));
    const real_t Dplus0 = the_cosmo_calc.get_growth_factor(astart);
    const real_t g1 = -Dplus0;
    const real_t g2 = ((LPTorder > 1) ? -3.0 / 7.0 * Dplus0 * Dplus0 : 0.0);
    const real_t g3 = ((LPTorder > 2) ? 1.0 / 3.0 * Dplus0 * Dplus0 * Dplus0 : 0.0);
    const real_t g3c = ((LPTorder > 2) ? 1.0 / 7.0 * Dplus0 * Dplus0 * Dplus0 : 0.0);

    // ic_generator.cc:290-310
    Grid_FFT<real_t> wnoise({ngrid, ngrid, ngrid}, {(real_t)boxlen, (real_t)boxlen, (real_t)boxlen});
    RNG_ngenic rng(seed);
    rng.Fill_Grid(wnoise);
    wnoise.FourierTransformForward();

    Grid_FFT<real_t> phi({ngrid, ngrid, ngrid}, {(real_t)boxlen, (real_t)boxlen, (real_t)boxlen});
    Grid_FFT<real_t> phi2({ngrid, ngrid, ngrid}, {(real_t)boxlen, (real_t)boxlen, (real_t)boxlen}, false);
    Grid_FFT<real_t> phi3({ngrid, ngrid, ngrid}, {(real_t)boxlen, (real_t)boxlen, (real_t)boxlen}, false);
    Grid_FFT<real_t> A3x({ngrid, ngrid, ngrid}, {(real_t)boxlen, (real_t)boxlen, (real_t)boxlen}, false);
    Grid_FFT<real_t> A3y({ngrid, ngrid, ngrid}, {(real_t)boxlen, (real_t)boxlen, (real_t)boxlen}, false);
    Grid_FFT<real_t> A3z({ngrid, ngrid, ngrid}, {(real_t)boxlen, (real_t)boxlen, (real_t)boxlen}, false);
    Grid_FFT<real_t>* A3[3] = {&A3x, &A3y, &A3z};
    OrszagConvolver<real_t> Conv({ngrid, ngrid, ngrid}, {(real_t)boxlen, (real_t)boxlen, (real_t)boxlen});

    // ic_generator.cc:378-383 (no fixing, no inversion)
    wnoise.apply_function_k([&](auto wn) { return wn / volfac; });

    // ic_generator.cc:509-516
    phi.FourierTransformForward(false);
    phi.assign_function_of_grids_kdep(
        [&](auto k, auto wn) {
            real_t kmod = k.norm();
            ccomplex_t delta = wn * (real_t)the_cosmo_calc.get_amplitude(kmod, delta_matter);
            return -delta / (kmod * kmod);
        },
        wnoise);
    phi.zero_DC_mode();

    // ic_generator.cc:523-537
    if (LPTorder > 1) {
        phi2.allocate();
        phi2.FourierTransformForward(false);
        Conv.convolve_SumOfHessians(phi, {0, 0}, phi, {1, 1}, {2, 2}, op::assign_to(phi2));
        Conv.convolve_Hessians(phi, {1, 1}, phi, {2, 2}, op::add_to(phi2));
        Conv.convolve_Hessians(phi, {0, 1}, phi, {0, 1}, op::subtract_from(phi2));
        Conv.convolve_Hessians(phi, {0, 2}, phi, {0, 2}, op::subtract_from(phi2));
        Conv.convolve_Hessians(phi, {1, 2}, phi, {1, 2}, op::subtract_from(phi2));
        phi2.apply_InverseLaplacian();
    }

    // ic_generator.cc:570-629
    if (LPTorder > 2) {
        phi3.allocate();
        phi3.FourierTransformForward(false);
        Conv.convolve_Hessians(phi, {0, 0}, phi, {1, 1}, phi, {2, 2}, op::assign_to(phi3));
        Conv.convolve_Hessians(phi, {0, 1}, phi, {0, 2}, phi, {1, 2}, op::multiply_add_to(phi3, 2.0));
        Conv.convolve_Hessians(phi, {1, 2}, phi, {1, 2}, phi, {0, 0}, op::subtract_from(phi3));
        Conv.convolve_Hessians(phi, {0, 2}, phi, {0, 2}, phi, {1, 1}, op::subtract_from(phi3));
        Conv.convolve_Hessians(phi, {0, 1}, phi, {0, 1}, phi, {2, 2}, op::subtract_from(phi3));

        Conv.convolve_SumOfHessians(phi, {0, 0}, phi2, {1, 1}, {2, 2}, op::multiply_add_to(phi3, -5.0 / 7.0));
        Conv.convolve_SumOfHessians(phi, {1, 1}, phi2, {2, 2}, {0, 0}, op::multiply_add_to(phi3, -5.0 / 7.0));
        Conv.convolve_SumOfHessians(phi, {2, 2}, phi2, {0, 0}, {1, 1}, op::multiply_add_to(phi3, -5.0 / 7.0));
        Conv.convolve_Hessians(phi, {0, 1}, phi2, {0, 1}, op::multiply_add_to(phi3, +10.0 / 7.0));
        Conv.convolve_Hessians(phi, {0, 2}, phi2, {0, 2}, op::multiply_add_to(phi3, +10.0 / 7.0));
        Conv.convolve_Hessians(phi, {1, 2}, phi2, {1, 2}, op::multiply_add_to(phi3, +10.0 / 7.0));
        phi3.apply_InverseLaplacian();

        for (int idim = 0; idim < 3; ++idim) {
            int idimp = (idim + 1) % 3, idimpp = (idim + 2) % 3;
            A3[idim]->allocate();
            A3[idim]->FourierTransformForward(false);
            Conv.convolve_Hessians(phi2, {idim, idimp}, phi, {idim, idimpp}, op::assign_to(*A3[idim]));
            Conv.convolve_Hessians(phi2, {idim, idimpp}, phi, {idim, idimp}, op::subtract_from(*A3[idim]));
            Conv.convolve_DifferenceOfHessians(phi, {idimp, idimpp}, phi2, {idimp, idimp}, {idimpp, idimpp},
                                               op::add_to(*A3[idim]));
            Conv.convolve_DifferenceOfHessians(phi2, {idimp, idimpp}, phi, {idimp, idimp}, {idimpp, idimpp},
                                               op::subtract_from(*A3[idim]));
            A3[idim]->apply_InverseLaplacian();
        }
    }

    // ic_generator.cc:645-659
    phi *= g1;
    if (LPTorder > 1) phi2 *= g2;
    if (LPTorder > 2) {
        phi3 *= g3;
        (*A3[0]) *= g3c;
        (*A3[1]) *= g3c;
        (*A3[2]) *= g3c;
    }

    // k-space buffers, FP64, one after another: phi, phi2, phi3, A3x, A3y, A3z
    Grid_FFT<real_t>* all[6] = {&phi, &phi2, &phi3, &A3x, &A3y, &A3z};
    const int count = LPTorder == 1 ? 1 : (LPTorder == 2 ? 2 : 6);
    const size_t stride = ngrid * ngrid * (ngrid + 2);
    for (int i = 0; i < count; ++i) check(mfic_grid_download(all[i]->handle(), out + i * stride, MFIC_F64));
}

template <typename F> int guarded(F f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // namespace

extern "C" {

const char* compat_last_error(void) { return g_err.c_str(); }

int compat_set_device(int device) {
    set_device(device);
    return 0;
}

int compat_release(void) {
    return guarded([] { release_context(); });
}

/* cosmology calculator queries at a = 1/(1+zstart) and wave number k [h/Mpc]:
 * out = {D+(a), vfact(a), f(a), T(k), amplitude(k)} (host only) */
int compat_calculator(const mfic_cosmo_params* cp, double k, double* out) {
    return guarded([&] {
        calculator c(*cp);
        TransferFunction_eisenstein tf(*cp);
        const double a = 1.0 / (1.0 + cp->zstart);
        out[0] = c.get_growth_factor(a);
        out[1] = c.get_vfact(a);
        out[2] = c.get_f(a);
        out[3] = tf.compute(k, delta_matter);
        out[4] = c.get_amplitude(k, delta_matter);
    });
}

/* out: 1, 2 or 6 k-space grids of n*n*(n+2) doubles (phi, phi2, phi3, A3) */
int compat_potentials(size_t n, double box, long seed, int order, int prec, const mfic_cosmo_params* cp, double* out) {
    return guarded([&] {
        if (prec == MFIC_F64) potentials<double>(n, box, seed, order, *cp, out);
        else potentials<float>(n, box, seed, order, *cp, out);
    });
}

/* Grid_FFT round trip and element access: fills a real grid from host values
 * through relem, transforms forward and back, reads every element back */
int compat_grid_roundtrip(size_t n0, size_t n1, size_t n2, const double* in, double* fwd, double* back) {
    return guarded([&] {
        Grid_FFT<double> g({n0, n1, n2}, {1.0, 1.0, 1.0});
        for (size_t i = 0; i < n0; ++i)
            for (size_t j = 0; j < n1; ++j)
                for (size_t k = 0; k < n2; ++k) g.relem(i, j, k) = in[(i * n1 + j) * n2 + k];
        g.FourierTransformForward();
        const size_t npc = n2 / 2 + 1;
        for (size_t i = 0; i < n0; ++i)
            for (size_t j = 0; j < n1; ++j)
                for (size_t k = 0; k < npc; ++k) {
                    const auto v = g.kelem(i, j, k);
                    fwd[2 * ((i * n1 + j) * npc + k)] = v.real();
                    fwd[2 * ((i * n1 + j) * npc + k) + 1] = v.imag();
                }
        g.FourierTransformBackward();
        for (size_t i = 0; i < n0; ++i)
            for (size_t j = 0; j < n1; ++j)
                for (size_t k = 0; k < n2; ++k) back[(i * n1 + j) * n2 + k] = g.relem(i, j, k);
    });
}

}  // extern "C"
