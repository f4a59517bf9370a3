// Complex FFT and semiclassical PPT fields (ppt.hip).
#pragma once
#include <cstddef>

#include "mfic_internal.hpp"

namespace mfic {

// unitary in-place c2c transform of a device [n0][n1][n2] complex array
template <typename T> void fft_c2c(mfic_ctx ctx, void* data, const size_t n[3], bool forward, double scale);

struct PptInput {
    int n = 0, order = 1;
    double L = 0, Dplus0 = 0;
    double vunit = 0;                          // D+ vfac / L * velocity_unit
    const void *phi_r = nullptr, *phi2_r = nullptr;  // growth-scaled, real [n][n][n+2]
    void *psi = nullptr, *grad = nullptr;      // [n][n][n] complex work arrays
};
// delta and vel[0..2]: real [n][n][n+2] grids of the input precision
void ppt_fields(mfic_ctx ctx, int prec, const PptInput& in, void* delta, void* vel[3], double* hbar);

}  // namespace mfic
