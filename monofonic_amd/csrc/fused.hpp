// Fused padded row pass of the LPT sources (fused.hip).
#pragma once
#include <cstddef>

#include "mfic_internal.hpp"

namespace mfic {

enum class FusedKind {
    Src2,   // in: 6 phi1 Hessians; out: 2LPT source
    Src2Q,  // in: 6 phi1 Hessians; out: 2LPT source, Qzz, Qyz, Qxx, Qyy
    Src3,   // in: 6 phi1 + 6 phi2 Hessians + 4 Q; out: phi3 source, A3x, A3y, A3z
};

// in / out: consecutive compact padded half-spectrum buffers ([Np][Np][N/2+1]
// complex each) after the axis-0 and axis-1 c2r passes / before the axis-1
// and axis-0 r2c passes
void fused_source_rows(mfic_ctx ctx, int prec, FusedKind kind, size_t n, const void* in, void* out);
int fused_inputs(FusedKind kind);
// last (contiguous-axis) c2r pass of three component grids (k-space after the
// axis-0 and axis-1 passes, [n][n][n/2+1] complex) fused with the lattice
// readout into FP64 [p][3] at particle offset pbase; positions sit at
// (ijk + shift) cells per axis
void rows_c2r_particles(mfic_ctx ctx, int prec, const void* const f[3], int n, int is_pos, double lunit,
                        const double shift[3], size_t pbase, double* dst);
int fused_outputs(FusedKind kind);

}  // namespace mfic
