// C ABI, second half: white noise, convolver, LPT sources and assembly, PLT
// tables, lattice readout and the whole-run object.
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "cosmology.hpp"
#include "fused.hpp"
#include "kspace.hpp"
#include "lpt.hpp"
#include "mfic_internal.hpp"
#include "output.hpp"
#include "ppt.hpp"
#include "ngenic.hpp"
#include "orszag.hpp"


struct mfic_conv_s {
    std::unique_ptr<mfic::ConvCore> core;
    void *kbuf = nullptr, *r1 = nullptr, *r2 = nullptr, *fbuf = nullptr;
    ~mfic_conv_s() {
        for (void* p : {kbuf, r1, r2, fbuf})
            if (p) (void)hipFree(p);
    }
};

struct mfic_plt_s {
    mfic_ctx ctx = nullptr;
    mfic::PltTables t;
};

struct mfic_run_s {
    mfic_ctx ctx = nullptr;
    mfic_ic_config cfg{};
    mfic_cosmo_params cosmo{};
    mfic_cosmo_scalars sc{};
    mfic::EhParams eh{};
    mfic_lpt_coef coef{};
    mfic_grid pot[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // phi1 phi2 phi3 A3x A3y A3z
    mfic_grid comp[24] = {};  // particle mode: component fields per sub-lattice (fcc: 4)
    mfic_grid fields[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    uint32_t* d_table = nullptr;
    void* d_tk = nullptr;  // spline nodes of a tabulated transfer function
    double *d_pos = nullptr, *d_vel = nullptr;
    double *d_pos_b = nullptr, *d_vel_b = nullptr;  // baryon species (secondary lattice)
    size_t npart = 0;
    mfic::PltTables plt;
    bool has_plt = false;
    ~mfic_run_s();
};

namespace mfic {


static void need(mfic_grid g, const char* what) {
    MFIC_CHECK(g && g->d, MFIC_ERR_ARG, std::string("null or unallocated grid: ") + what);
}

static OperandDesc to_desc(const mfic_operand* o) {
    MFIC_CHECK(o && o->f && o->f->d, MFIC_ERR_ARG, "null operand");
    MFIC_CHECK(o->kind >= MFIC_OPND_HESS && o->kind <= MFIC_OPND_IDENT, MFIC_ERR_ARG, "bad operand kind");
    for (int i = 0; i < 4; ++i) MFIC_CHECK(o->d[i] >= 0 && o->d[i] < 3, MFIC_ERR_ARG, "operand direction out of range");
    MFIC_CHECK(o->f->space == MFIC_KSPACE, MFIC_ERR_STATE, "operand fields must be in k-space");
    return OperandDesc{o->f->d, o->kind, {o->d[0], o->d[1], o->d[2], o->d[3]}};
}

static void* dmalloc(size_t bytes, const char* what) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) throw Error(MFIC_ERR_ALLOC, std::string("device allocation failed: ") + what);
    return p;
}

static mfic_grid new_grid(mfic_ctx ctx, size_t n, double L, int prec, int space) {
    const size_t nn[3] = {n, n, n};
    const double LL[3] = {L, L, L};
    mfic_grid g = nullptr;
    int rc = mfic_grid_create(ctx, nn, LL, prec, space, &g);
    if (rc != MFIC_OK) throw Error(rc, mfic_last_error());
    return g;
}

static void compute_coefs(const mfic_cosmo_params& p, size_t n, double L, int order, const mfic_cosmo_scalars& s,
                          mfic_lpt_coef* k) {
    k->volfac = std::pow(L / (double)n / 2.0 / 3.14159265358979323846, 1.5);
    k->Dplus0 = s.Dplus_start;
    k->vfac = s.vfac_start;
    const double D = k->Dplus0;
    k->g1 = -D;
    k->g2 = order > 1 ? -3.0 / 7.0 * D * D : 0.0;
    k->g3 = order > 2 ? 1.0 / 3.0 * D * D * D : 0.0;
    k->g3c = order > 2 ? 1.0 / 7.0 * D * D * D : 0.0;
    k->vfac1 = k->vfac;
    k->vfac2 = 2 * k->vfac;
    k->vfac3 = 3 * k->vfac;
}

static void component_coefs(const mfic_lpt_coef& k, int comp, double cf[4]) {
    if (comp < 3) {
        cf[0] = k.g1; cf[1] = k.g2; cf[2] = k.g3; cf[3] = k.g3c;
    } else {
        cf[0] = k.vfac1 * k.g1; cf[1] = k.vfac2 * k.g2; cf[2] = k.vfac3 * k.g3; cf[3] = k.vfac3 * k.g3c;
    }
}

}  // namespace mfic

using namespace mfic;

mfic_run_s::~mfic_run_s() {
    for (auto& g : pot) if (g) mfic_grid_destroy(g);
    for (auto& g : fields) if (g) mfic_grid_destroy(g);
    for (auto& g : comp) if (g) mfic_grid_destroy(g);
    if (d_table) (void)hipFree(d_table);
    if (d_tk) (void)hipFree(d_tk);
    if (d_pos) (void)hipFree(d_pos);
    if (d_vel) (void)hipFree(d_vel);
    if (d_pos_b) (void)hipFree(d_pos_b);
    if (d_vel_b) (void)hipFree(d_vel_b);
}

extern "C" {

int mfic_ngenic_seedtable(long seed, size_t n, uint32_t* table) {
    return guard_call([&] {
        MFIC_CHECK(table && n > 0, MFIC_ERR_ARG, "null table or empty grid");
        ngenic_seedtable(seed, n, table);
    });
}

int mfic_ngenic_fill(mfic_grid g, long seed) {
    return guard_call([&] {
        need(g, "noise");
        const size_t n = g->n[0];
        std::vector<uint32_t> tab(n * n);
        ngenic_seedtable(seed, n, tab.data());
        uint32_t* dt = static_cast<uint32_t*>(ctx_scratch(g->ctx, "seedtable", n * n * sizeof(uint32_t)));
        MFIC_HIP(hipMemcpyAsync(dt, tab.data(), n * n * sizeof(uint32_t), hipMemcpyHostToDevice, g->ctx->stream));
        grid_fill(g, 0.0);
        ngenic_fill(g, dt);
        g->space = MFIC_KSPACE;
        MFIC_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int mfic_conv_create(mfic_ctx ctx, const size_t n[3], const double L[3], int prec, mfic_conv* out) {
    return guard_call([&] {
        MFIC_CHECK(ctx && n && L && out, MFIC_ERR_ARG, "null argument");
        auto c = std::make_unique<mfic_conv_s>();
        c->core = std::make_unique<ConvCore>(ctx, n, L, prec);
        const size_t es = c->core->esize();
        c->kbuf = dmalloc(c->core->kbuf_reals() * es, "convolver padded k buffer");
        c->r1 = dmalloc(c->core->real_reals() * es, "convolver padded real buffer");
        c->r2 = dmalloc(c->core->real_reals() * es, "convolver padded real buffer");
        c->fbuf = dmalloc(c->core->grid_reals() * es, "convolver N-grid buffer");
        ++ctx->children;
        *out = c.release();
    });
}

int mfic_conv_destroy(mfic_conv c) {
    return guard_call([&] {
        if (!c) return;
        mfic_ctx ctx = c->core->ctx;
        (void)hipStreamSynchronize(ctx->stream);
        delete c;
        --ctx->children;
    });
}

static void conv_check_out(mfic_conv c, mfic_grid out) {
    MFIC_CHECK(c, MFIC_ERR_ARG, "null convolver");
    need(out, "output");
    MFIC_CHECK(out->prec == c->core->prec && out->n[0] == c->core->N[0] && out->n[1] == c->core->N[1] &&
                   out->n[2] == c->core->N[2],
               MFIC_ERR_ARG, "output grid does not match the convolver");
}

int mfic_conv_convolve2(mfic_conv c, const mfic_operand* a, const mfic_operand* b, mfic_grid out, int op, double coef) {
    return guard_call([&] {
        conv_check_out(c, out);
        MFIC_CHECK(op >= MFIC_OP_ASSIGN && op <= MFIC_OP_MADD, MFIC_ERR_ARG, "bad output operator");
        const OperandDesc da = to_desc(a), db = to_desc(b);
        MFIC_CHECK(a->f->prec == c->core->prec && b->f->prec == c->core->prec, MFIC_ERR_ARG, "precision mismatch");
        c->core->convolve2(da, db, out->d, op, coef, c->kbuf, c->r1, c->r2);
        out->space = MFIC_KSPACE;
    });
}

int mfic_conv_convolve3(mfic_conv c, const mfic_operand* a, const mfic_operand* b, const mfic_operand* cc,
                        mfic_grid out, int op, double coef) {
    return guard_call([&] {
        conv_check_out(c, out);
        MFIC_CHECK(op >= MFIC_OP_ASSIGN && op <= MFIC_OP_MADD, MFIC_ERR_ARG, "bad output operator");
        const OperandDesc da = to_desc(a), db = to_desc(b), dc = to_desc(cc);
        c->core->convolve3(da, db, dc, out->d, op, coef, c->kbuf, c->r1, c->r2, c->fbuf);
        out->space = MFIC_KSPACE;
    });
}

int mfic_lpt_sources(mfic_ctx ctx, int order, mfic_grid phi1, mfic_grid phi2, mfic_grid phi3, mfic_grid A3x,
                     mfic_grid A3y, mfic_grid A3z) {
    return guard_call([&] {
        MFIC_CHECK(ctx, MFIC_ERR_ARG, "null context");
        MFIC_CHECK(order >= 1 && order <= 3, MFIC_ERR_CONFIG, "LPT order must be 1, 2 or 3");
        need(phi1, "phi1");
        MFIC_CHECK(phi1->space == MFIC_KSPACE, MFIC_ERR_STATE, "phi1 must be in k-space");
        if (order > 1) need(phi2, "phi2");
        if (order > 2) { need(phi3, "phi3"); need(A3x, "A3x"); need(A3y, "A3y"); need(A3z, "A3z"); }
        lpt_sources(ctx, order, phi1, phi2, phi3, A3x, A3y, A3z);
    });
}

int mfic_lpt_coefficients(const mfic_cosmo_params* p, size_t n, double L, int order, mfic_lpt_coef* out) {
    return guard_call([&] {
        MFIC_CHECK(p && out && n > 0 && L > 0, MFIC_ERR_ARG, "bad argument");
        mfic_cosmo_scalars s;
        cosmo_compute(*p, &s, nullptr);
        compute_coefs(*p, n, L, order, s, out);
    });
}

int mfic_plt_create(mfic_ctx ctx, int lattice, int ng, mfic_plt* out) {
    return guard_call([&] {
        MFIC_CHECK(ctx && out, MFIC_ERR_ARG, "null argument");
        auto p = std::make_unique<mfic_plt_s>();
        p->ctx = ctx;
        plt_build(ctx, lattice, ng, p->t);
        ++ctx->children;
        *out = p.release();
    });
}

int mfic_plt_destroy(mfic_plt p) {
    return guard_call([&] {
        if (!p) return;
        mfic_ctx ctx = p->ctx;
        delete p;
        --ctx->children;
    });
}

int mfic_plt_download(mfic_plt p, double* Dr, double* Dtheta, double* Dphi, double* vfac) {
    return guard_call([&] {
        MFIC_CHECK(p, MFIC_ERR_ARG, "null tables");
        const size_t nk = (size_t)p->t.ng * p->t.ng * (p->t.ng / 2 + 1);
        double* dst[4] = {Dr, Dtheta, Dphi, vfac};
        for (int c = 0; c < 4; ++c)
            if (dst[c]) std::memcpy(dst[c], p->t.host.data() + c * nk, nk * sizeof(double));
    });
}

int mfic_lpt_component(mfic_ctx ctx, int order, int comp, const mfic_lpt_coef* k, mfic_plt plt, double unit,
                       mfic_grid phi1, mfic_grid phi2, mfic_grid phi3, mfic_grid A3x, mfic_grid A3y, mfic_grid A3z,
                       mfic_grid out) {
    return guard_call([&] {
        MFIC_CHECK(ctx && k, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(comp >= 0 && comp < 6, MFIC_ERR_ARG, "component must be 0..5");
        MFIC_CHECK(order >= 1 && order <= 3, MFIC_ERR_CONFIG, "LPT order must be 1, 2 or 3");
        need(phi1, "phi1");
        need(out, "out");
        const void* phi[3] = {phi1->d, order > 1 ? (need(phi2, "phi2"), phi2->d) : nullptr,
                              order > 2 ? (need(phi3, "phi3"), phi3->d) : nullptr};
        const void* A3[3] = {nullptr, nullptr, nullptr};
        if (order > 2) {
            need(A3x, "A3x"); need(A3y, "A3y"); need(A3z, "A3z");
            A3[0] = A3x->d; A3[1] = A3y->d; A3[2] = A3z->d;
        }
        double cf[4];
        component_coefs(*k, comp, cf);
        PltDev pv;
        if (plt) pv = plt->t.view;
        MFIC_CHECK(phi1->n[0] == phi1->n[1] && phi1->n[1] == phi1->n[2], MFIC_ERR_SIZE, "assembly needs a cubic grid");
        assemble_component(ctx, out->prec, out->n, out->L[0], order, comp, cf, unit, pv, phi, A3, out->d, nullptr);
        out->space = MFIC_KSPACE;
    });
}

int mfic_lattice_readout_index(mfic_grid field, int lattice, int lattice_index, int idim, int is_pos, double lunit,
                               double* dst) {
    return guard_call([&] {
        need(field, "field");
        MFIC_CHECK(dst, MFIC_ERR_ARG, "null destination");
        MFIC_CHECK(lattice >= MFIC_LATTICE_SC && lattice <= MFIC_LATTICE_FCC, MFIC_ERR_UNSUPPORTED,
                   "readout supports the sc, bcc and fcc lattices");
        MFIC_CHECK(lattice_index == 0 || lattice_index == 1, MFIC_ERR_ARG, "lattice index must be 0 or 1");
        MFIC_HIP(hipSetDevice(field->ctx->device));
        require_device_memory(dst, "the readout destination");
        MFIC_CHECK(idim >= 0 && idim < 3, MFIC_ERR_ARG, "idim must be 0..2");
        MFIC_CHECK(field->space == MFIC_RSPACE, MFIC_ERR_STATE, "field must be in real space");
        const int n = (int)field->n[0];
        double g[3] = {0.0, 0.0, 0.0};
        if (lattice_index == 1) {
            // the secondary lattice: a global shift first (particle_generator.hh:58-63, 512-516)
            second_lattice_shift(lattice, g);
            int rc = mfic_grid_shift(field, g, 1);
            if (rc != MFIC_OK) throw Error(rc, mfic_last_error());
        }
        // lattice_shifts (particle_generator.hh:47-53); each sub-lattice is
        // read after a shift relative to the previous one, as the reference does
        static const double SUB[3][4][3] = {{{0, 0, 0}},
                                            {{0, 0, 0}, {0.5, 0.5, 0.5}},
                                            {{0, 0, 0}, {0, 0.5, 0.5}, {0.5, 0, 0.5}, {0.5, 0.5, 0}}};
        for (int sub = 0; sub < (1 << lattice); ++sub) {
            if (sub > 0) {
                double rel[3];
                for (int a = 0; a < 3; ++a) rel[a] = SUB[lattice][sub][a] - SUB[lattice][sub - 1][a];
                int rc = mfic_grid_shift(field, rel, 1);
                if (rc != MFIC_OK) throw Error(rc, mfic_last_error());
            }
            lattice_readout_raw(field->ctx, field->prec, field->d, n, idim, is_pos, lunit,
                                SUB[lattice][sub][idim] + g[idim], (size_t)sub * n * n * n, dst);
        }
    });
}

int mfic_lattice_readout(mfic_grid field, int lattice, int idim, int is_pos, double lunit, double* dst) {
    return mfic_lattice_readout_index(field, lattice, 0, idim, is_pos, lunit, dst);
}

int mfic_run_create(mfic_ctx ctx, const mfic_ic_config* cfg, const mfic_cosmo_params* cosmo, mfic_run* out) {
    return guard_call([&] {
        MFIC_CHECK(ctx && cfg && cosmo && out, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(cfg->lpt_order >= 1 && cfg->lpt_order <= 3, MFIC_ERR_CONFIG, "LPTorder must be 1, 2 or 3");
        MFIC_CHECK(cfg->lattice >= MFIC_LATTICE_SC && cfg->lattice <= MFIC_LATTICE_FCC, MFIC_ERR_UNSUPPORTED,
                   "ParticleLoad must be sc, bcc or fcc");
        MFIC_CHECK(cfg->prec == MFIC_F32 || cfg->prec == MFIC_F64, MFIC_ERR_CONFIG, "bad precision");
        MFIC_CHECK(cfg->n % 2 == 0 || cfg->lpt_order == 1, MFIC_ERR_SIZE,
                   "Orszag convolver for LPTorder>1 requires even grid size!");
        MFIC_CHECK(cfg->L > 0, MFIC_ERR_CONFIG, "BoxLength must be positive");
        MFIC_CHECK(!cfg->baryons || cfg->particles, MFIC_ERR_UNSUPPORTED,
                   "baryon species are built for particle output only");
        MFIC_CHECK(!cfg->baryons || cosmo->tk_n == 0, MFIC_ERR_UNSUPPORTED,
                   "baryon species need the Eisenstein & Hu transfer function (delta_bc = 0)");
        auto r = std::make_unique<mfic_run_s>();
        r->ctx = ctx;
        r->cfg = *cfg;
        r->cosmo = *cosmo;
        cosmo_compute(*cosmo, &r->sc, &r->eh);
        {
            const size_t n3[3] = {cfg->n, cfg->n, cfg->n};
            const double L3[3] = {cfg->L, cfg->L, cfg->L};
            check_tk_range(*cosmo, n3, L3);
        }
        r->d_tk = tk_upload(*cosmo, &r->eh);
        // the caller's table arrays are borrowed for this call only
        r->cosmo.tk_k = r->cosmo.tk_T = nullptr;
        r->cosmo.tk_n = 0;
        compute_coefs(*cosmo, cfg->n, cfg->L, cfg->lpt_order, r->sc, &r->coef);
        const size_t n = cfg->n;
        for (int p = 0; p < (cfg->lpt_order == 1 ? 1 : cfg->lpt_order == 2 ? 2 : 6); ++p)
            r->pot[p] = new_grid(ctx, n, cfg->L, cfg->prec, MFIC_KSPACE);
        if (cfg->particles) {
            // component grids: 6 per sub-lattice, and a second set for the
            // shifted output of the sc baryon pass
            const int ncomp = cfg->lattice == MFIC_LATTICE_FCC ? 24 : (cfg->lattice == MFIC_LATTICE_BCC || cfg->baryons) ? 12 : 6;
            for (int q = 0; q < ncomp; ++q) r->comp[q] = new_grid(ctx, n, cfg->L, cfg->prec, MFIC_KSPACE);
            r->npart = (size_t(1) << cfg->lattice) * n * n * n;
            r->d_pos = static_cast<double*>(dmalloc(3 * r->npart * sizeof(double), "particle positions"));
            r->d_vel = static_cast<double*>(dmalloc(3 * r->npart * sizeof(double), "particle velocities"));
            if (cfg->baryons) {
                r->d_pos_b = static_cast<double*>(dmalloc(3 * r->npart * sizeof(double), "baryon positions"));
                r->d_vel_b = static_cast<double*>(dmalloc(3 * r->npart * sizeof(double), "baryon velocities"));
            }
        } else {
            for (int c = 0; c < 6; ++c) r->fields[c] = new_grid(ctx, n, cfg->L, cfg->prec, MFIC_RSPACE);
        }
        // RNG_ngenic constructor work (seed table) happens here, outside the run
        std::vector<uint32_t> tab(n * n);
        ngenic_seedtable(cfg->seed, n, tab.data());
        r->d_table = static_cast<uint32_t*>(dmalloc(n * n * sizeof(uint32_t), "seed table"));
        MFIC_HIP(hipMemcpy(r->d_table, tab.data(), n * n * sizeof(uint32_t), hipMemcpyHostToDevice));
        if (cfg->plt) {
            plt_build(ctx, cfg->lattice, 64, r->plt);
            r->has_plt = true;
        }
        MFIC_HIP(hipStreamSynchronize(ctx->stream));
        ++ctx->children;
        *out = r.release();
    });
}

int mfic_run_destroy(mfic_run r) {
    return guard_call([&] {
        if (!r) return;
        mfic_ctx ctx = r->ctx;
        (void)hipStreamSynchronize(ctx->stream);
        delete r;
        --ctx->children;
    });
}

int mfic_run_execute(mfic_run r, int blocking) {
    return guard_call([&] {
        MFIC_CHECK(r, MFIC_ERR_ARG, "null run");
        const mfic_ic_config& c = r->cfg;
        mfic_grid phi1 = r->pot[0];
        phi1->space = MFIC_KSPACE;
        if (phi1_cubic(phi1)) {  // phi1 formed in the noise kernel's stores
            ngenic_phi1_fill(phi1, r->d_table, phi1_table(phi1, r->eh), r->coef.volfac, c.do_fixing, c.do_inversion);
        } else {
            grid_fill(phi1, 0.0);
            ngenic_fill(phi1, r->d_table);
            phi1_from_noise(phi1, phi1, r->eh, r->coef.volfac, c.do_fixing, c.do_inversion);
        }
        if (c.lpt_order > 1) lpt_sources(r->ctx, c.lpt_order, phi1, r->pot[1], r->pot[2], r->pot[3], r->pot[4], r->pot[5]);
        const void* phi[3] = {phi1->d, r->pot[1] ? r->pot[1]->d : nullptr, r->pot[2] ? r->pot[2]->d : nullptr};
        const void* A3[3] = {r->pot[3] ? r->pot[3]->d : nullptr, r->pot[4] ? r->pot[4]->d : nullptr,
                             r->pot[5] ? r->pot[5]->d : nullptr};
        PltDev pv;
        if (r->has_plt) pv = r->plt.view;
        const size_t nn[3] = {c.n, c.n, c.n};
        const double norm = 1.0 / std::sqrt((double)c.n * (double)c.n * (double)c.n);
        double cd[4], cv[4];
        component_coefs(r->coef, 0, cd);
        component_coefs(r->coef, 3, cv);
        const int nsub = c.particles ? (1 << c.lattice) : 1;
        void* outs[12] = {};
        mfic_grid* og = c.particles ? r->comp : r->fields;
        for (int q = 0; q < 6; ++q) outs[q] = og[q]->d;
        // the second bcc sub-lattice is not assembled in k-space: its
        // components come out of the first sub-lattice's axis-0 pass
        assemble_all(r->ctx, c.prec, c.n, c.L, c.lpt_order, cd, cv, c.lunit / c.L, c.vunit / c.L, pv, phi, A3, outs, 1);
        if (c.particles) {
            // sub-lattice offsets in half cells (lattice_shifts,
            // particle_generator.hh:47-53)
            static const int SUB[3][4][3] = {{{0, 0, 0}},
                                             {{0, 0, 0}, {1, 1, 1}},
                                             {{0, 0, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}}};
            const size_t cbytes = og[0]->bytes();
            // strided passes per component, every sub-lattice's phase formed
            // in the axis-0 pass from one load per output pair, then the
            // contiguous pass fused with the readout of the three components
            // of each sub-lattice; gw = global offset of the species' lattice
            auto species = [&](const int gw[3], double* dpos, double* dvel) {
                const bool global = gw[0] || gw[1] || gw[2];
                for (int q = 0; q < 6; ++q) {
                    if (nsub == 1 && !global) {
                        if (c.prec == MFIC_F64) fft_c2r_cols<double>(r->ctx, (double*)og[q]->d, nn);
                        else fft_c2r_cols<float>(r->ctx, (float*)og[q]->d, nn);
                        continue;
                    }
                    int w[4][3];
                    for (int s = 0; s < 4; ++s)
                        for (int a = 0; a < 3; ++a) w[s][a] = gw[a] + SUB[c.lattice][s < nsub ? s : 0][a];
                    if (nsub == 4)  // fcc: the second pair of sub-lattices from a copy of the component
                        MFIC_HIP(hipMemcpyAsync(og[q + 12]->d, og[q]->d, cbytes, hipMemcpyDeviceToDevice, r->ctx->stream));
                    for (int pr = 0; pr < (nsub == 4 ? 2 : 1); ++pr) {
                        void* a = og[q + 12 * pr]->d;
                        void* b = og[q + 12 * pr + 6]->d;
                        const int* wa = (pr == 0 && !global && nsub == 2) ? nullptr : w[2 * pr];
                        const int* wb = (pr == 0 && !global && nsub == 2) ? nullptr : w[2 * pr + 1];
                        if (c.prec == MFIC_F64) fft_c2r_cols_shift<double>(r->ctx, (double*)a, (double*)b, nn, wa, wb);
                        else fft_c2r_cols_shift<float>(r->ctx, (float*)a, (float*)b, nn, wa, wb);
                    }
                }
                for (int s = 0; s < nsub; ++s) {
                    const void* fp[3] = {og[6 * s]->d, og[6 * s + 1]->d, og[6 * s + 2]->d};
                    const void* fv[3] = {og[6 * s + 3]->d, og[6 * s + 4]->d, og[6 * s + 5]->d};
                    const size_t pbase = (size_t)s * c.n * c.n * c.n;
                    double sh[3];
                    for (int a = 0; a < 3; ++a) sh[a] = 0.5 * (gw[a] + SUB[c.lattice][s][a]);
                    rows_c2r_particles(r->ctx, c.prec, fp, (int)c.n, 1, c.lunit, sh, pbase, dpos);
                    rows_c2r_particles(r->ctx, c.prec, fv, (int)c.n, 0, 0.0, sh, pbase, dvel);
                }
            };
            const int zero[3] = {0, 0, 0};
            species(zero, r->d_pos, r->d_vel);
            if (c.baryons) {
                // the baryon species on the secondary lattice (ic_generator.cc:
                // 716-745, particle_generator.hh:512-531): with delta_bc = 0
                // (Eisenstein & Hu) its displacements and velocities are the
                // same fields read out on the lattice offset by
                // second_lattice_shift.  The components are assembled again:
                // the dark matter pass transformed them in place.
                assemble_all(r->ctx, c.prec, c.n, c.L, c.lpt_order, cd, cv, c.lunit / c.L, c.vunit / c.L, pv, phi, A3,
                             outs, 1);
                double g[3];
                second_lattice_shift(c.lattice, g);
                const int gw[3] = {(int)(2 * g[0]), (int)(2 * g[1]), (int)(2 * g[2])};
                species(gw, r->d_pos_b, r->d_vel_b);
            }
        } else {
            for (int q = 0; q < 6; ++q) {
                if (c.prec == MFIC_F64) fft_c2r<double>(r->ctx, (double*)og[q]->d, nn, norm);
                else fft_c2r<float>(r->ctx, (float*)og[q]->d, nn, norm);
                og[q]->space = MFIC_RSPACE;
            }
        }
        if (blocking) MFIC_HIP(hipStreamSynchronize(r->ctx->stream));
    });
}

int mfic_run_field(mfic_run r, int comp, mfic_grid* out) {
    return guard_call([&] {
        MFIC_CHECK(r && out && comp >= 0 && comp < 6, MFIC_ERR_ARG, "bad argument");
        MFIC_CHECK(!r->cfg.particles, MFIC_ERR_STATE, "run was configured for particle output");
        *out = r->fields[comp];
    });
}

int mfic_run_particles(mfic_run r, double** pos, double** vel, size_t* npart) {
    return guard_call([&] {
        MFIC_CHECK(r, MFIC_ERR_ARG, "null run");
        MFIC_CHECK(r->cfg.particles, MFIC_ERR_STATE, "run was configured for field output");
        if (pos) *pos = r->d_pos;
        if (vel) *vel = r->d_vel;
        if (npart) *npart = r->npart;
    });
}

int mfic_run_baryon_particles(mfic_run r, double** pos, double** vel, size_t* npart) {
    return guard_call([&] {
        MFIC_CHECK(r, MFIC_ERR_ARG, "null run");
        MFIC_CHECK(r->cfg.particles && r->cfg.baryons, MFIC_ERR_STATE, "run has no baryon species");
        if (pos) *pos = r->d_pos_b;
        if (vel) *vel = r->d_vel_b;
        if (npart) *npart = r->npart;
    });
}

int mfic_run_potential(mfic_run r, int which, mfic_grid* out) {
    return guard_call([&] {
        MFIC_CHECK(r && out && which >= 0 && which < 6, MFIC_ERR_ARG, "bad argument");
        MFIC_CHECK(r->pot[which] != nullptr, MFIC_ERR_STATE, "potential not computed for this LPT order");
        *out = r->pot[which];
    });
}

// the baryon species of a run as a particle source: the secondary lattice's
// particles with IDs after the dark matter's (IDoffset 1,
// ic_generator.cc:722-724, particle_generator.hh:231)
static ParticleSource baryon_source(mfic_run r, const ParticleSource& dm) {
    ParticleSource b = dm;
    b.pos = r->d_pos_b;
    b.vel = r->d_vel_b;
    b.id_offset = (uint64_t)dm.overload * dm.ncell;
    return b;
}

int mfic_gadget_units(double boxlength, double zstart, double* lunit, double* vunit, double* munit) {
    return guard_call([&] {
        MFIC_CHECK(lunit && vunit && munit, MFIC_ERR_ARG, "null output");
        MFIC_CHECK(boxlength > 0 && zstart > -1.0, MFIC_ERR_CONFIG, "bad box length or redshift");
        gadget_units(boxlength, zstart, lunit, vunit, munit);
    });
}

static void check_gadget_units(const mfic_ic_config& c, double zstart) {
    double lu, vu, mu;
    gadget_units(c.L, zstart, &lu, &vu, &mu);
    MFIC_CHECK(std::fabs(c.lunit - lu) <= 1e-12 * lu && std::fabs(c.vunit - vu) <= 1e-12 * vu, MFIC_ERR_CONFIG,
               "particles must be in the Gadget-HDF5 units: lunit = " + std::to_string(lu) +
                   ", vunit = " + std::to_string(vu) + " (mfic_gadget_units)");
}

int mfic_gadget_write(mfic_run r, const char* fname, const mfic_gadget_opts* opts) {
    return guard_call([&] {
        MFIC_CHECK(r && fname && opts, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(r->cfg.particles && r->d_pos && r->d_vel, MFIC_ERR_STATE, "run was configured for field output");
        check_gadget_units(r->cfg, r->cosmo.zstart);
        ParticleSource src;
        src.ctx = r->ctx;
        src.pos = r->d_pos;
        src.vel = r->d_vel;
        src.npart = r->npart;
        src.ncell = r->cfg.n * r->cfg.n * r->cfg.n;
        src.overload = 1u << r->cfg.lattice;
        ParticleSource bsrc = baryon_source(r, src);
        MFIC_HIP(hipSetDevice(r->ctx->device));
        gadget_write(fname, r->cfg, r->cosmo, *opts, src, r->cfg.baryons ? &bsrc : nullptr, r->d_tk != nullptr);
    });
}

int mfic_gadget_write_host(const char* fname, const mfic_ic_config* cfg, const mfic_cosmo_params* cosmo,
                           const mfic_gadget_opts* opts, const double* pos, const double* vel) {
    return guard_call([&] {
        MFIC_CHECK(fname && cfg && cosmo && opts && pos && vel, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(cfg->lattice >= MFIC_LATTICE_SC && cfg->lattice <= MFIC_LATTICE_FCC, MFIC_ERR_UNSUPPORTED,
                   "ParticleLoad must be sc, bcc or fcc");
        MFIC_CHECK(cfg->n > 0, MFIC_ERR_SIZE, "empty lattice");
        check_gadget_units(*cfg, cosmo->zstart);
        ParticleSource src;
        src.pos = pos;
        src.vel = vel;
        src.ncell = cfg->n * cfg->n * cfg->n;
        src.overload = 1u << cfg->lattice;
        src.npart = src.ncell * src.overload;
        MFIC_CHECK(!cfg->baryons, MFIC_ERR_UNSUPPORTED, "host arrays carry one species: write baryon runs from the run");
        gadget_write(fname, *cfg, *cosmo, *opts, src, nullptr, cosmo->tk_n > 0);
    });
}

int mfic_swift_units(double boxlength, const mfic_cosmo_params* cosmo, double* lunit, double* vunit, double* munit) {
    return guard_call([&] {
        MFIC_CHECK(cosmo && lunit && vunit && munit, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(boxlength > 0 && cosmo->H0 > 0, MFIC_ERR_CONFIG, "bad box length or H0");
        swift_units(boxlength, *cosmo, lunit, vunit, munit);
    });
}

static void check_swift_units(const mfic_ic_config& c, const mfic_cosmo_params& cosmo) {
    double lu, vu, mu;
    swift_units(c.L, cosmo, &lu, &vu, &mu);
    MFIC_CHECK(std::fabs(c.lunit - lu) <= 1e-12 * lu && std::fabs(c.vunit - vu) <= 1e-12 * vu, MFIC_ERR_CONFIG,
               "particles must be in the SWIFT units: lunit = " + std::to_string(lu) + ", vunit = " +
                   std::to_string(vu) + " (mfic_swift_units)");
}

int mfic_swift_write(mfic_run r, const char* fname, const mfic_swift_opts* opts) {
    return guard_call([&] {
        MFIC_CHECK(r && fname && opts, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(r->cfg.particles && r->d_pos && r->d_vel, MFIC_ERR_STATE, "run was configured for field output");
        check_swift_units(r->cfg, r->cosmo);
        ParticleSource src;
        src.ctx = r->ctx;
        src.pos = r->d_pos;
        src.vel = r->d_vel;
        src.npart = r->npart;
        src.ncell = r->cfg.n * r->cfg.n * r->cfg.n;
        src.overload = 1u << r->cfg.lattice;
        ParticleSource bsrc = baryon_source(r, src);
        MFIC_HIP(hipSetDevice(r->ctx->device));
        swift_write(fname, r->cfg, r->cosmo, *opts, src, r->cfg.baryons ? &bsrc : nullptr, r->d_tk != nullptr);
    });
}

int mfic_swift_write_host(const char* fname, const mfic_ic_config* cfg, const mfic_cosmo_params* cosmo,
                          const mfic_swift_opts* opts, const double* pos, const double* vel) {
    return guard_call([&] {
        MFIC_CHECK(fname && cfg && cosmo && opts && pos && vel, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(cfg->lattice >= MFIC_LATTICE_SC && cfg->lattice <= MFIC_LATTICE_FCC, MFIC_ERR_UNSUPPORTED,
                   "ParticleLoad must be sc, bcc or fcc");
        MFIC_CHECK(cfg->n > 0, MFIC_ERR_SIZE, "empty lattice");
        check_swift_units(*cfg, *cosmo);
        ParticleSource src;
        src.pos = pos;
        src.vel = vel;
        src.ncell = cfg->n * cfg->n * cfg->n;
        src.overload = 1u << cfg->lattice;
        src.npart = src.ncell * src.overload;
        MFIC_CHECK(!cfg->baryons, MFIC_ERR_UNSUPPORTED, "host arrays carry one species: write baryon runs from the run");
        swift_write(fname, *cfg, *cosmo, *opts, src, nullptr, cosmo->tk_n > 0);
    });
}

int mfic_fft_c2c(mfic_ctx ctx, void* data, const size_t n[3], int prec, int forward) {
    return guard_call([&] {
        MFIC_CHECK(ctx && data && n, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(prec == MFIC_F32 || prec == MFIC_F64, MFIC_ERR_ARG, "bad precision");
        MFIC_HIP(hipSetDevice(ctx->device));
        require_device_memory(data, "the complex grid");
        const double norm = 1.0 / std::sqrt((double)n[0] * (double)n[1] * (double)n[2]);
        if (prec == MFIC_F64) fft_c2c<double>(ctx, data, n, forward != 0, norm);
        else fft_c2c<float>(ctx, data, n, forward != 0, norm);
    });
}

int mfic_ppt_fields(mfic_run r, double velocity_unit, mfic_grid delta, mfic_grid vx, mfic_grid vy, mfic_grid vz,
                    double* hbar) {
    return guard_call([&] {
        MFIC_CHECK(r, MFIC_ERR_ARG, "null run");
        const mfic_ic_config& c = r->cfg;
        mfic_grid outs[4] = {delta, vx, vy, vz};
        for (mfic_grid g : outs) {
            need(g, "output");
            MFIC_CHECK(g->prec == c.prec && g->n[0] == c.n && g->n[1] == c.n && g->n[2] == c.n, MFIC_ERR_ARG,
                       "PPT outputs must be real grids of the run's size and precision");
        }
        MFIC_CHECK(r->pot[0] && r->pot[0]->space == MFIC_KSPACE, MFIC_ERR_STATE, "execute the run first");
        const size_t n = c.n, es = c.prec == MFIC_F64 ? 8 : 4;
        const size_t nn[3] = {n, n, n};
        const double norm = 1.0 / std::sqrt((double)n * n * n);
        // growth-scaled real-space potentials (ic_generator.cc:644-658, 809, 847)
        struct Buf {
            void* p = nullptr;
            ~Buf() {
                if (p) (void)hipFree(p);
            }
        } ph1, ph2, psi, grad;
        const size_t greal = n * n * (n + 2) * es;
        ph1.p = dmalloc(greal, "PPT phi");
        if (c.lpt_order >= 2) ph2.p = dmalloc(greal, "PPT phi2");
        psi.p = dmalloc(n * n * n * 2 * es, "PPT psi");
        grad.p = dmalloc(n * n * n * 2 * es, "PPT grad psi");
        MFIC_HIP(hipMemcpyAsync(ph1.p, r->pot[0]->d, greal, hipMemcpyDeviceToDevice, r->ctx->stream));
        if (ph2.p) MFIC_HIP(hipMemcpyAsync(ph2.p, r->pot[1]->d, greal, hipMemcpyDeviceToDevice, r->ctx->stream));
        if (c.prec == MFIC_F64) {
            fft_c2r<double>(r->ctx, (double*)ph1.p, nn, norm * r->coef.g1);
            if (ph2.p) fft_c2r<double>(r->ctx, (double*)ph2.p, nn, norm * r->coef.g2);
        } else {
            fft_c2r<float>(r->ctx, (float*)ph1.p, nn, norm * r->coef.g1);
            if (ph2.p) fft_c2r<float>(r->ctx, (float*)ph2.p, nn, norm * r->coef.g2);
        }
        PptInput in;
        in.n = (int)n;
        in.order = c.lpt_order;
        in.L = c.L;
        in.Dplus0 = r->coef.Dplus0;
        in.vunit = r->coef.Dplus0 * r->coef.vfac / c.L * velocity_unit;
        in.phi_r = ph1.p;
        in.phi2_r = ph2.p;
        in.psi = psi.p;
        in.grad = grad.p;
        void* vel[3] = {vx->d, vy->d, vz->d};
        ppt_fields(r->ctx, c.prec, in, delta->d, vel, hbar);
        MFIC_HIP(hipStreamSynchronize(r->ctx->stream));
        for (mfic_grid g : outs) g->space = MFIC_RSPACE;
    });
}

}  // extern "C"
