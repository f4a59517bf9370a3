// NGenIC white-noise generator (random_ngenic.cc:39-154).
#pragma once
#include <cstdint>

#include "mfic_internal.hpp"

namespace mfic {
// seed table of n*n 32-bit seeds in NGenIC's shell order (host)
void ngenic_seedtable(long seed, size_t n, uint32_t* table);
// fill the k-space grid g from a device copy of the seed table
void ngenic_fill(mfic_grid g, const uint32_t* dev_table);
}  // namespace mfic
