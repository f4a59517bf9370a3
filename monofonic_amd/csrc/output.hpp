// Particle output writers (output.hip).
#pragma once
#include <cstddef>
#include <cstdint>

#include "mfic_internal.hpp"

namespace mfic {

// where the particle values come from: the device arrays of a run, staged
// chunk by chunk on ctx's stream, or host arrays (ctx = null); FP64 [p][3]
// positions and velocities of npart = overload * ncell particles
struct ParticleSource {
    mfic_ctx ctx = nullptr;
    const double *pos = nullptr, *vel = nullptr;
    size_t npart = 0, ncell = 0;
    unsigned overload = 1;
    uint64_t id_offset = 0;  // IDoffset * overload * ncell (particle_generator.hh:231)
};

// output_plugin units of the Gadget-HDF5 writer (output_gadget_hdf5.cc:79-85)
void gadget_units(double boxlength, double zstart, double* lunit, double* vunit, double* munit);
// one Gadget-HDF5 file of a run's particles; baryons: the baryon species
// (secondary lattice) or null; tabulated: the run used a tabulated transfer
// function (recorded in ICs_parameters)
void gadget_write(const char* fname, const mfic_ic_config& cfg, const mfic_cosmo_params& cosmo,
                  const mfic_gadget_opts& opts, const ParticleSource& src, const ParticleSource* baryons,
                  bool tabulated);
// output_plugin units of the SWIFT writer (output_swift.cc:63-69)
void swift_units(double boxlength, const mfic_cosmo_params& cosmo, double* lunit, double* vunit, double* munit);
// one SWIFT initial-condition file of a run's particles; baryons: the
// baryon species (secondary lattice) or null
void swift_write(const char* fname, const mfic_ic_config& cfg, const mfic_cosmo_params& cosmo,
                 const mfic_swift_opts& opts, const ParticleSource& src, const ParticleSource* baryons,
                 bool tabulated);

}  // namespace mfic
