// Pruned / fused 3/2-padded transforms (pruned.hip).
#pragma once
#include <vector>

#include "mfic_internal.hpp"
#include "orszag.hpp"

namespace mfic {

struct ProdTerm {
    const void* a;  // padded real field [Np][Np][Np]
    const void* b;
    double c;
};

// operand of an N-grid k-space field -> padded real field rout ([Np]^3),
// through the compact padded k buffer kbuf ([Np][Np][N/2+1] complex)
void padded_c2r_from_operand(mfic_ctx ctx, int prec, const OperandDesc& o, size_t n, double L, void* kbuf, void* rout);
// sum_t c_t a_t b_t (padded real) -> truncated N-grid k-space with the unpad
// rules, output operator op / coef and optional -1/k^2 (DC := 0)
void padded_r2c_products(mfic_ctx ctx, int prec, const std::vector<ProdTerm>& terms, size_t n, double L, void* kbuf,
                         void* out, int op, double coef, bool invlap);

// operand -> compact padded buffer after the axis-0 and axis-1 c2r passes
// (half-spectrum rows ready for a row pass)
void padded_c2r_to_kbuf(mfic_ctx ctx, int prec, const OperandDesc& o, size_t n, double L, void* kbuf);
// the six Hessians xx yy zz xy xz yz of an N-grid k-space field -> six
// consecutive compact padded buffers (out) after the axis-0 and axis-1 c2r
// passes; xbuf is scratch for three compact buffers
void padded_hessians_to_kbuf(mfic_ctx ctx, int prec, const void* f, size_t n, double L, void* xbuf, void* out);
// a compact padded buffer of r2c'd rows of a product (real space along axes
// 0 and 1) -> the same buffer truncated to the source modes with Nyquist
// planes zeroed and padded again, ready as the input of a row pass: the unpad
// / pad_insert pair of a two-stage product, as two in-place filter passes
void padded_truncate_repad(mfic_ctx ctx, int prec, size_t n, void* kbuf);
// compact padded buffer of r2c'd rows -> axis-1 and axis-0 r2c passes with the
// unpad rules, output operator and optional -1/k^2 into the N-grid; kbuf is
// overwritten
void padded_r2c_from_kbuf(mfic_ctx ctx, int prec, size_t n, double L, void* kbuf, void* out, int op, double coef,
                          bool invlap);

}  // namespace mfic
