// NGenIC white-noise generator (random_ngenic.cc:39-154).
#pragma once
#include <cstdint>

#include "mfic_internal.hpp"

namespace mfic {
// seed table of n*n 32-bit seeds in NGenIC's shell order (host)
void ngenic_seedtable(long seed, size_t n, uint32_t* table);
// fill the k-space grid g from a device copy of the seed table
void ngenic_fill(mfic_grid g, const uint32_t* dev_table);
// the noise fill with phi1_from_noise folded into its stores (cubic box):
// phitab is the -A(k)/k^2 table of phi1_table, volfac / fixing / inversion
// as for phi1_from_noise
void ngenic_phi1_fill(mfic_grid phi, const uint32_t* dev_table, const double* phitab, double volfac, int fixing,
                      int inversion);
}  // namespace mfic
