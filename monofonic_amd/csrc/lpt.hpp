// LPT assembly, lattice readout and PLT tables.
#pragma once
#include <vector>

#include "mfic_internal.hpp"

namespace mfic {

// device view of the PLT tables (each ng x ng x (ng/2+1) doubles)
struct PltDev {
    int ng = 0;
    const double* Dr = nullptr;
    const double* Dtheta = nullptr;
    const double* Dphi = nullptr;
    const double* vfac = nullptr;
};

struct PltTables {
    int ng = 0, lattice = 0;
    double* dev = nullptr;
    std::vector<double> host;  // Dr, Dtheta, Dphi, vfac concatenated
    PltDev view;
    PltTables() = default;
    PltTables(const PltTables&) = delete;
    PltTables& operator=(const PltTables&) = delete;
    ~PltTables();
};

void plt_build(mfic_ctx ctx, int lattice, int ng, PltTables& out);

// cf = {c1, c2, c3, cA}: coefficients of phi1, phi2, phi3 and of the A3 curl
void assemble_component(mfic_ctx ctx, int prec, const size_t n[3], double L, int order, int comp, const double cf[4],
                        double unit, const PltDev& plt, const void* const phi[3], const void* const A3[3], void* out,
                        const double shift_dx[3]);

// all 6 components (x nsub sub-lattices) in one pass; out[comp + 6 sub]
void assemble_all(mfic_ctx ctx, int prec, size_t n, double L, int order, const double c[4], const double cv[4],
                  double ulen, double uvel, const PltDev& plt, const void* const phi[3], const void* const A3[3],
                  void* const out[12], int nsub);
// three real-space component fields -> interleaved FP64 particle values
void lattice_readout3(mfic_ctx ctx, int prec, const void* const f[3], int n, int is_pos, double lunit, double shift,
                      size_t pbase, double* dst);

// offset of the secondary lattice of a Bravais load in cells
// (particle_generator.hh:58-63): sc (1/2,1/2,1/2), bcc (1/2,1/2,0),
// fcc (1/2,1/2,1/2)
inline void second_lattice_shift(int lattice, double g[3]) {
    g[0] = g[1] = 0.5;
    g[2] = lattice == MFIC_LATTICE_BCC ? 0.0 : 0.5;
}
void lattice_readout_raw(mfic_ctx ctx, int prec, const void* field, int n, int idim, int is_pos, double lunit,
                         double shift, size_t pbase, double* dst);

}  // namespace mfic
