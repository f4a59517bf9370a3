// C ABI, differentiable 2LPT: forward map white noise -> 2LPT displacement
// and its adjoint (the maths is in adjoint.hip); plus wrapping of external
// device buffers as grids for framework bindings.  No reference
// counterpart: monofonIC has no derivative of its initial-condition map.
#include <cmath>
#include <memory>
#include <string>

#include "cosmology.hpp"
#include "kspace.hpp"
#include "mfic_internal.hpp"
#include "orszag.hpp"

namespace mfic {
void kspace_zero_nyquist(mfic_grid g);
void kspace_grad_adjoint(mfic_grid out, mfic_grid gx, mfic_grid gy, mfic_grid gz, double unit);
void kspace_hess_madd(mfic_grid out, mfic_grid r, int a, int b, double c);
}  // namespace mfic

using mfic::Error;

namespace {
void ck(int rc) {
    if (rc != MFIC_OK) throw Error(rc, mfic_last_error());
}
struct GridPtr {
    mfic_grid g = nullptr;
    GridPtr() = default;
    GridPtr(const GridPtr&) = delete;
    ~GridPtr() {
        if (g) (void)mfic_grid_destroy(g);
    }
};
}  // namespace

struct mfic_lpt2_s {
    mfic_ctx ctx = nullptr;
    size_t n = 0;
    double L = 0.0, unit = 1.0;
    int prec = MFIC_F32;
    mfic::EhParams eh{};
    mfic_lpt_coef coef{};
    GridPtr w, phi1, phi2, lam, lmu, r, t[3];
    mfic_conv conv = nullptr;
    void* d_tk = nullptr;  // spline nodes of a tabulated transfer function
    ~mfic_lpt2_s() {
        if (conv) (void)mfic_conv_destroy(conv);
        if (d_tk) (void)hipFree(d_tk);
    }
};

namespace {
void make(mfic_lpt2_s* p, GridPtr& g, int space) {
    const size_t n3[3] = {p->n, p->n, p->n};
    const double L3[3] = {p->L, p->L, p->L};
    ck(mfic_grid_create(p->ctx, n3, L3, p->prec, space, &g.g));
}
void check_io(const mfic_lpt2_s* p, mfic_grid g, const char* what) {
    MFIC_CHECK(g && g->d, MFIC_ERR_ARG, std::string("null grid: ") + what);
    MFIC_CHECK(g->n[0] == p->n && g->n[1] == p->n && g->n[2] == p->n && g->prec == p->prec, MFIC_ERR_ARG,
               std::string("grid size / precision differs from the plan: ") + what);
}
// phi1 = S P_N r2c(noise) (the forward map's first linear stage)
void phi1_of(mfic_lpt2_s* p, mfic_grid noise) {
    ck(mfic_grid_copy(p->w.g, noise));
    ck(mfic_grid_set_space(p->w.g, MFIC_RSPACE));
    ck(mfic_fft_forward(p->w.g));
    mfic::kspace_zero_nyquist(p->w.g);
    mfic::phi1_from_noise(p->phi1.g, p->w.g, p->eh, p->coef.volfac, 0, 0);
    p->phi1.g->space = MFIC_KSPACE;
}
}  // namespace

extern "C" {

int mfic_grid_wrap(mfic_ctx ctx, const size_t n[3], const double L[3], int prec, int space, void* device_ptr,
                   mfic_grid* out) {
    return mfic::guard_call([&] {
        MFIC_CHECK(ctx && n && L && device_ptr && out, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(prec == MFIC_F32 || prec == MFIC_F64, MFIC_ERR_ARG, "precision must be MFIC_F32 or MFIC_F64");
        MFIC_CHECK(space == MFIC_KSPACE || space == MFIC_RSPACE, MFIC_ERR_ARG, "bad space");
        MFIC_CHECK(n[0] > 0 && n[1] > 0 && n[2] > 0 && n[2] % 2 == 0, MFIC_ERR_SIZE, "bad grid size");
        MFIC_CHECK(((uintptr_t)device_ptr % (prec == MFIC_F64 ? 16 : 8)) == 0, MFIC_ERR_ARG,
                   "device buffer must be aligned to one complex element");
        MFIC_HIP(hipSetDevice(ctx->device));
        mfic::require_device_memory(device_ptr, "the wrapped buffer");
        auto* g = new mfic_grid_s();
        g->ctx = ctx;
        for (int i = 0; i < 3; ++i) { g->n[i] = n[i]; g->L[i] = L[i]; }
        g->prec = prec;
        g->space = space;
        g->nreal = n[0] * n[1] * (n[2] + 2);
        g->d = device_ptr;
        g->owns = false;
        ++ctx->children;
        *out = g;
    });
}

int mfic_lpt2_create(mfic_ctx ctx, size_t n, double L, int prec, const mfic_cosmo_params* cosmo, double unit,
                     mfic_lpt2* out) {
    return mfic::guard_call([&] {
        MFIC_CHECK(ctx && cosmo && out, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(prec == MFIC_F32 || prec == MFIC_F64, MFIC_ERR_ARG, "bad precision");
        MFIC_CHECK(L > 0.0, MFIC_ERR_CONFIG, "BoxLength must be positive");
        auto p = std::make_unique<mfic_lpt2_s>();
        p->ctx = ctx;
        p->n = n;
        p->L = L;
        p->prec = prec;
        p->unit = unit;
        mfic_cosmo_scalars sc;
        mfic::cosmo_compute(*cosmo, &sc, &p->eh);
        {
            const size_t n3[3] = {n, n, n};
            const double L3[3] = {L, L, L};
            mfic::check_tk_range(*cosmo, n3, L3);
        }
        p->d_tk = mfic::tk_upload(*cosmo, &p->eh);
        ck(mfic_lpt_coefficients(cosmo, n, L, 2, &p->coef));
        make(p.get(), p->w, MFIC_RSPACE);
        make(p.get(), p->phi1, MFIC_KSPACE);
        make(p.get(), p->phi2, MFIC_KSPACE);
        make(p.get(), p->lam, MFIC_KSPACE);
        make(p.get(), p->lmu, MFIC_KSPACE);
        make(p.get(), p->r, MFIC_KSPACE);
        for (auto& t : p->t) make(p.get(), t, MFIC_RSPACE);
        const size_t n3[3] = {n, n, n};
        const double L3[3] = {L, L, L};
        ck(mfic_conv_create(ctx, n3, L3, prec, &p->conv));
        ++ctx->children;
        *out = p.release();
    });
}

int mfic_lpt2_destroy(mfic_lpt2 p) {
    return mfic::guard_call([&] {
        if (!p) return;
        (void)hipStreamSynchronize(p->ctx->stream);
        mfic_ctx ctx = p->ctx;
        delete p;
        --ctx->children;
    });
}

int mfic_lpt2_forward(mfic_lpt2 p, mfic_grid noise, mfic_grid psi_x, mfic_grid psi_y, mfic_grid psi_z) {
    return mfic::guard_call([&] {
        MFIC_CHECK(p, MFIC_ERR_ARG, "null plan");
        check_io(p, noise, "noise");
        mfic_grid psi[3] = {psi_x, psi_y, psi_z};
        for (auto g : psi) check_io(p, g, "displacement");
        phi1_of(p, noise);
        mfic::lpt_sources(p->ctx, 2, p->phi1.g, p->phi2.g, nullptr, nullptr, nullptr, nullptr);
        for (int d = 0; d < 3; ++d) {
            ck(mfic_lpt_component(p->ctx, 2, d, &p->coef, nullptr, p->unit, p->phi1.g, p->phi2.g, nullptr, nullptr,
                                  nullptr, nullptr, psi[d]));
            ck(mfic_fft_backward(psi[d]));
        }
    });
}

int mfic_lpt2_adjoint(mfic_lpt2 p, mfic_grid noise, mfic_grid cot_x, mfic_grid cot_y, mfic_grid cot_z,
                      mfic_grid grad) {
    return mfic::guard_call([&] {
        MFIC_CHECK(p, MFIC_ERR_ARG, "null plan");
        check_io(p, noise, "noise");
        check_io(p, grad, "gradient");
        mfic_grid cot[3] = {cot_x, cot_y, cot_z};
        for (auto g : cot) check_io(p, g, "cotangent");
        phi1_of(p, noise);  // recomputed, not stored by the forward call
        for (int d = 0; d < 3; ++d) {
            ck(mfic_grid_copy(p->t[d].g, cot[d]));
            ck(mfic_grid_set_space(p->t[d].g, MFIC_RSPACE));
            ck(mfic_fft_forward(p->t[d].g));
        }
        // lam = unit sum_d (-i g_d) G_d (adjoint of the gradient, unit and DC := 0)
        mfic::kspace_grad_adjoint(p->lam.g, p->t[0].g, p->t[1].g, p->t[2].g, p->unit);
        // phi2 branch: lam_mu = P_N lap^-1 (g2 lam)
        ck(mfic_grid_copy(p->lmu.g, p->lam.g));
        ck(mfic_grid_scale(p->lmu.g, p->coef.g2));
        ck(mfic_grid_inverse_laplacian(p->lmu.g));
        mfic::kspace_zero_nyquist(p->lmu.g);
        // phi1 branch
        ck(mfic_grid_scale(p->lam.g, p->coef.g1));
        // adjoint of the six Orszag products of the 2LPT source
        mfic_operand L1{p->lmu.g, MFIC_OPND_IDENT, {0, 0, 0, 0}};
        struct Term { int kind, d[4], a, b; double c; };
        static const Term terms[6] = {
            {MFIC_OPND_SUM_HESS, {1, 1, 2, 2}, 0, 0, 1.0}, {MFIC_OPND_SUM_HESS, {0, 0, 2, 2}, 1, 1, 1.0},
            {MFIC_OPND_SUM_HESS, {0, 0, 1, 1}, 2, 2, 1.0}, {MFIC_OPND_HESS, {0, 1, 0, 0}, 0, 1, -2.0},
            {MFIC_OPND_HESS, {0, 2, 0, 0}, 0, 2, -2.0},    {MFIC_OPND_HESS, {1, 2, 0, 0}, 1, 2, -2.0}};
        for (const auto& t : terms) {
            mfic_operand h{p->phi1.g, t.kind, {t.d[0], t.d[1], t.d[2], t.d[3]}};
            ck(mfic_conv_convolve2(p->conv, &L1, &h, p->r.g, MFIC_OP_ASSIGN, 0.0));
            mfic::kspace_hess_madd(p->lam.g, p->r.g, t.a, t.b, t.c);
        }
        // adjoint of phi1 = S P_N r2c(noise)
        mfic::phi1_from_noise(p->w.g, p->lam.g, p->eh, p->coef.volfac, 0, 0);
        p->w.g->space = MFIC_KSPACE;
        mfic::kspace_zero_nyquist(p->w.g);
        ck(mfic_fft_backward(p->w.g));
        ck(mfic_grid_copy(grad, p->w.g));
    });
}

}  // extern "C"
