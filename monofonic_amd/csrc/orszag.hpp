// Orszag 3/2 convolver core and the reorganised LPT source schedule.
#pragma once
#include <vector>

#include "mfic_internal.hpp"

namespace mfic {

struct OperandDesc {
    const void* f;  // k-space N-grid field (device)
    int kind;       // mfic_opkind
    int d[4];
};

// Padded-grid operations for one N-grid size (OrszagConvolver, convolution.hh:412-651)
struct ConvCore {
    mfic_ctx ctx;
    int prec;
    size_t N[3], Np[3];
    double len[3];
    ConvCore(mfic_ctx c, const size_t n[3], const double L[3], int prec);
    size_t padded_reals() const;
    size_t grid_reals() const;
    size_t esize() const;
    struct Term { const void* a; const void* b; double c; };
    void pad_insert(const OperandDesc& o, void* fp) const;
    void unpad(const void* fp, void* out, int op, double coef) const;
    void c2r_padded(void* fp) const;
    void r2c_padded(void* fp) const;
    void mul(void* a, const void* b) const;
    void prodsum(void* out, const std::vector<Term>& terms) const;
    // buffer sizes (reals) of the pruned path: compact padded k buffer and a
    // padded real field
    size_t kbuf_reals() const;
    size_t real_reals() const;
    void convolve2(const OperandDesc& a, const OperandDesc& b, void* out, int op, double coef, void* kbuf, void* r1,
                   void* r2) const;
    void convolve3(const OperandDesc& a, const OperandDesc& b, const OperandDesc& c, void* out, int op, double coef,
                   void* kbuf, void* r1, void* r2, void* fbuf) const;
};

void add_inplace(mfic_ctx ctx, int prec, void* a, const void* b, size_t n);

// phi2 (order >= 2), phi3 / A3 (order 3) from phi1, inverse Laplacian applied,
// growth factors not applied
void lpt_sources(mfic_ctx ctx, int order, mfic_grid phi1, mfic_grid phi2, mfic_grid phi3, mfic_grid A3x,
                 mfic_grid A3y, mfic_grid A3z);

}  // namespace mfic
