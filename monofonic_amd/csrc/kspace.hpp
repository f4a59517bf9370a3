// Host launchers of the pointwise grid kernels (kspace.hip).
#pragma once
#include "cosmology.hpp"
#include "mfic_internal.hpp"

namespace mfic {
void grid_scale(mfic_grid g, double s);
void grid_fill(mfic_grid g, double v);
void grid_zero_dc(mfic_grid g);
void grid_inverse_laplacian(mfic_grid g);
void grid_phase_shift(mfic_grid g, const double s[3]);
void convert_buffer(mfic_ctx ctx, const void* in, int in_prec, void* out, int out_prec, size_t n);
// table: device buffer of 3 (N/2)^2 + 1 doubles for the cubic-grid path, or
// null to use the context's scratch
void phi1_from_noise(mfic_grid phi, mfic_grid wn, const EhParams& eh, double volfac, int fixing, int inversion,
                     double* table = nullptr);
// whether phi's box is cubic (equal sizes and sides): the case of the
// tabulated -A(k)/k^2 factor
bool phi1_cubic(mfic_grid phi);
// the FP64 -A(k)/k^2 table per integer |n|^2 of phi's cubic box (3 (N/2)^2 + 1
// values) in `table`, or in the context's scratch when null
const double* phi1_table(mfic_grid phi, const EhParams& eh, double* table = nullptr);
// resolves the no-allocation scratch's address on `device` (current)
void phi1_scratch_setup(int device);
}  // namespace mfic
