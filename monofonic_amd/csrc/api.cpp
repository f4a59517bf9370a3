// C ABI of the mfic library: contexts, grids, FFTs, cosmology and phi1.
// Every entry point converts C++ exceptions into negative status codes.
#include <cmath>
#include <cstring>
#include <string>

#include "cosmology.hpp"
#include "kspace.hpp"
#include "mfic_internal.hpp"

namespace mfic {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

template <typename F> static int guard(F&& f) {
    try {
        f();
        return MFIC_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return MFIC_ERR_CONFIG;
    } catch (const std::bad_alloc& e) {
        set_last_error("host allocation failed");
        return MFIC_ERR_ALLOC;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MFIC_ERR_ARG;
    }
}

void* ctx_scratch(mfic_ctx ctx, const std::string& key, size_t bytes) {
    auto& s = ctx->scratch[key];
    if (s.second < bytes) {
        MFIC_CHECK(!ctx->capturing, MFIC_ERR_STATE,
                   "scratch '" + key + "' would grow during graph capture: run the sequence once before capturing");
        if (s.first) MFIC_HIP(hipFree(s.first));
        s.first = nullptr;
        s.second = 0;
        hipError_t e = hipMalloc(&s.first, bytes);
        if (e != hipSuccess) throw Error(MFIC_ERR_ALLOC, "scratch allocation of " + std::to_string(bytes) + " bytes failed");
        s.second = bytes;
    }
    return s.first;
}

static hipEvent_t pool_event(mfic_ctx ctx) {
    if (!ctx->prof_pool.empty()) {
        hipEvent_t e = ctx->prof_pool.back();
        ctx->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    MFIC_HIP(hipEventCreate(&e));
    return e;
}

ProfScope::ProfScope(mfic_ctx ctx, const std::string& name, double bytes) : ctx_(ctx) {
    if (!ctx_ || !ctx_->prof || ctx_->capturing) return;
    mfic_prof_rec r{name, bytes, pool_event(ctx_), pool_event(ctx_)};
    MFIC_HIP(hipEventRecord(r.a, ctx_->stream));
    idx_ = (long)ctx_->prof_recs.size();
    ctx_->prof_recs.push_back(r);
}

ProfScope::~ProfScope() {
    if (idx_ >= 0) (void)hipEventRecord(ctx_->prof_recs[idx_].b, ctx_->stream);
}

static void check_grid(mfic_grid g) {
    MFIC_CHECK(g != nullptr && g->d != nullptr, MFIC_ERR_ARG, "null or unallocated grid");
}

}  // namespace mfic

using namespace mfic;

namespace mfic {

void require_device_memory(const void* p, const char* what) {
    hipPointerAttribute_t at{};
    const hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // unknown to the runtime: not device memory
        at.type = hipMemoryTypeUnregistered;
    }
    MFIC_CHECK(at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged, MFIC_ERR_ARG,
               std::string(what) + " must be device memory (host memory is refused)");
}

}  // namespace mfic

extern "C" {

const char* mfic_last_error(void) { return g_last_error.c_str(); }
const char* mfic_version(void) { return "mfic 0.1 (gfx950)"; }

int mfic_ctx_create(int device, mfic_ctx* out) {
    return guard([&] {
        MFIC_CHECK(out != nullptr, MFIC_ERR_ARG, "null output");
        int ndev = 0;
        MFIC_HIP(hipGetDeviceCount(&ndev));
        MFIC_CHECK(device >= 0 && device < ndev, MFIC_ERR_ARG, "no such HIP device");
        MFIC_HIP(hipSetDevice(device));
        auto* c = new mfic_ctx_s();
        c->device = device;
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || lds <= 0)
            lds = 65536;
        c->max_lds = (size_t)lds;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw Error(MFIC_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        try {
            phi1_scratch_setup(device);  // device is current
        } catch (...) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            throw;
        }
        *out = c;
    });
}

int mfic_ctx_destroy(mfic_ctx ctx) {
    return guard([&] {
        if (!ctx) return;
        const long live = ctx->children.load();
        MFIC_CHECK(live == 0, MFIC_ERR_STATE,
                   "context still has " + std::to_string(live) +
                       " live handle(s) (grids, convolvers, PLT tables, runs, graphs, 2LPT plans): destroy them first");
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& kv : ctx->tables) (void)hipFree(kv.second.dev);
        for (auto& kv : ctx->scratch) (void)hipFree(kv.second.first);
        for (auto& r : ctx->prof_recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (auto& e : ctx->prof_pool) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int mfic_ctx_sync(mfic_ctx ctx) {
    return guard([&] {
        MFIC_CHECK(ctx, MFIC_ERR_ARG, "null context");
        MFIC_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int mfic_ctx_stream(mfic_ctx ctx, void** stream) {
    return guard([&] {
        MFIC_CHECK(ctx && stream, MFIC_ERR_ARG, "null argument");
        *stream = (void*)ctx->stream;
    });
}

int mfic_prof_enable(mfic_ctx ctx, int on) {
    return guard([&] {
        MFIC_CHECK(ctx, MFIC_ERR_ARG, "null context");
        ctx->prof = on != 0;
    });
}

int mfic_prof_report(mfic_ctx ctx, char* buf, size_t len) {
    return guard([&] {
        MFIC_CHECK(ctx && buf && len > 0, MFIC_ERR_ARG, "bad argument");
        MFIC_HIP(hipStreamSynchronize(ctx->stream));
        struct Agg { long count = 0; double ms = 0, bytes = 0; };
        std::map<std::string, Agg> agg;
        for (auto& r : ctx->prof_recs) {
            float ms = 0.f;
            MFIC_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            Agg& a = agg[r.name];
            a.count += 1;
            a.ms += ms;
            a.bytes += r.bytes;
            ctx->prof_pool.push_back(r.a);
            ctx->prof_pool.push_back(r.b);
        }
        ctx->prof_recs.clear();
        std::string out = "{";
        bool first = true;
        char tmp[512];
        for (auto& kv : agg) {
            snprintf(tmp, sizeof tmp, "%s\"%s\": {\"count\": %ld, \"ms_total\": %.6f, \"bytes_total\": %.0f}",
                     first ? "" : ", ", kv.first.c_str(), kv.second.count, kv.second.ms, kv.second.bytes);
            out += tmp;
            first = false;
        }
        out += "}";
        MFIC_CHECK(out.size() + 1 <= len, MFIC_ERR_ARG, "report buffer too small (need " + std::to_string(out.size() + 1) + ")");
        std::memcpy(buf, out.c_str(), out.size() + 1);
    });
}

int mfic_grid_create(mfic_ctx ctx, const size_t n[3], const double L[3], int prec, int space, mfic_grid* out) {
    return guard([&] {
        MFIC_CHECK(ctx && n && L && out, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(prec == MFIC_F32 || prec == MFIC_F64, MFIC_ERR_ARG, "precision must be MFIC_F32 or MFIC_F64");
        MFIC_CHECK(space == MFIC_KSPACE || space == MFIC_RSPACE, MFIC_ERR_ARG, "bad space");
        MFIC_CHECK(n[0] > 0 && n[1] > 0 && n[2] > 0, MFIC_ERR_SIZE, "grid sizes must be positive");
        auto* g = new mfic_grid_s();
        g->ctx = ctx;
        for (int i = 0; i < 3; ++i) { g->n[i] = n[i]; g->L[i] = L[i]; }
        g->prec = prec;
        g->space = space;
        g->nreal = n[0] * n[1] * (n[2] + 2);
        hipError_t e = hipMalloc(&g->d, g->bytes());
        if (e != hipSuccess) {
            delete g;
            throw Error(MFIC_ERR_ALLOC, "device allocation of " + std::to_string(n[0] * n[1] * (n[2] + 2) * (prec ? 8 : 4)) + " bytes failed");
        }
        ++ctx->children;
        grid_fill(g, 0.0);
        *out = g;
    });
}

int mfic_grid_destroy(mfic_grid g) {
    return guard([&] {
        if (!g) return;
        (void)hipStreamSynchronize(g->ctx->stream);
        if (g->owns && g->d) (void)hipFree(g->d);
        --g->ctx->children;
        delete g;
    });
}

int mfic_grid_set_space(mfic_grid g, int space) {
    return guard([&] {
        check_grid(g);
        MFIC_CHECK(space == MFIC_KSPACE || space == MFIC_RSPACE, MFIC_ERR_ARG, "bad space");
        g->space = space;
    });
}

int mfic_grid_get_space(mfic_grid g, int* space) {
    return guard([&] {
        check_grid(g);
        MFIC_CHECK(space, MFIC_ERR_ARG, "null output");
        *space = g->space;
    });
}

int mfic_grid_device_ptr(mfic_grid g, void** ptr, size_t* nreal) {
    return guard([&] {
        check_grid(g);
        if (ptr) *ptr = g->d;
        if (nreal) *nreal = g->nreal;
    });
}

int mfic_grid_upload(mfic_grid g, const void* host, int host_prec) {
    return guard([&] {
        check_grid(g);
        MFIC_CHECK(host, MFIC_ERR_ARG, "null host pointer");
        const size_t hb = g->nreal * (host_prec == MFIC_F64 ? 8 : 4);
        if (host_prec == g->prec) {
            MFIC_HIP(hipMemcpyAsync(g->d, host, hb, hipMemcpyHostToDevice, g->ctx->stream));
        } else {
            void* st = ctx_scratch(g->ctx, "stage", hb);
            MFIC_HIP(hipMemcpyAsync(st, host, hb, hipMemcpyHostToDevice, g->ctx->stream));
            convert_buffer(g->ctx, st, host_prec, g->d, g->prec, g->nreal);
        }
        MFIC_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int mfic_grid_download(mfic_grid g, void* host, int host_prec) {
    return guard([&] {
        check_grid(g);
        MFIC_CHECK(host, MFIC_ERR_ARG, "null host pointer");
        const size_t hb = g->nreal * (host_prec == MFIC_F64 ? 8 : 4);
        if (host_prec == g->prec) {
            MFIC_HIP(hipMemcpyAsync(host, g->d, hb, hipMemcpyDeviceToHost, g->ctx->stream));
        } else {
            void* st = ctx_scratch(g->ctx, "stage", hb);
            convert_buffer(g->ctx, g->d, g->prec, st, host_prec, g->nreal);
            MFIC_HIP(hipMemcpyAsync(host, st, hb, hipMemcpyDeviceToHost, g->ctx->stream));
        }
        MFIC_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int mfic_grid_copy(mfic_grid dst, mfic_grid src) {
    return guard([&] {
        check_grid(dst);
        check_grid(src);
        MFIC_CHECK(dst->n[0] == src->n[0] && dst->n[1] == src->n[1] && dst->n[2] == src->n[2], MFIC_ERR_ARG,
                   "grid sizes differ");
        convert_buffer(dst->ctx, src->d, src->prec, dst->d, dst->prec, dst->nreal);
        dst->space = src->space;
    });
}

int mfic_grid_zero(mfic_grid g) {
    return guard([&] {
        check_grid(g);
        grid_fill(g, 0.0);
    });
}

int mfic_grid_scale(mfic_grid g, double c) {
    return guard([&] {
        check_grid(g);
        grid_scale(g, c);
    });
}

int mfic_grid_zero_dc(mfic_grid g) {
    return guard([&] {
        check_grid(g);
        MFIC_CHECK(g->space == MFIC_KSPACE, MFIC_ERR_STATE,
                   "zero_dc on a real-space grid is the mean subtraction, outside the hot path");
        grid_zero_dc(g);
    });
}

static void fft_forward_impl(mfic_grid g) {
    if (g->space == MFIC_KSPACE) return;
    const double norm = 1.0 / std::sqrt((double)g->n[0] * (double)g->n[1] * (double)g->n[2]);
    if (g->prec == MFIC_F64) fft_r2c<double>(g->ctx, (double*)g->d, g->n, norm);
    else fft_r2c<float>(g->ctx, (float*)g->d, g->n, norm);
    g->space = MFIC_KSPACE;
}
static void fft_backward_impl(mfic_grid g) {
    if (g->space == MFIC_RSPACE) return;
    const double norm = 1.0 / std::sqrt((double)g->n[0] * (double)g->n[1] * (double)g->n[2]);
    if (g->prec == MFIC_F64) fft_c2r<double>(g->ctx, (double*)g->d, g->n, norm);
    else fft_c2r<float>(g->ctx, (float*)g->d, g->n, norm);
    g->space = MFIC_RSPACE;
}

int mfic_grid_inverse_laplacian(mfic_grid g) {
    return guard([&] {
        check_grid(g);
        fft_forward_impl(g);
        grid_inverse_laplacian(g);
    });
}

int mfic_grid_shift(mfic_grid g, const double s[3], int transform_back) {
    return guard([&] {
        check_grid(g);
        MFIC_CHECK(s, MFIC_ERR_ARG, "null shift");
        fft_forward_impl(g);
        grid_phase_shift(g, s);
        if (transform_back) fft_backward_impl(g);
    });
}

int mfic_fft_forward(mfic_grid g) {
    return guard([&] {
        check_grid(g);
        fft_forward_impl(g);
    });
}

int mfic_fft_backward(mfic_grid g) {
    return guard([&] {
        check_grid(g);
        fft_backward_impl(g);
    });
}

int mfic_fft_size_supported(size_t n) { return fft_length_supported(n) ? 1 : 0; }

int mfic_fft_set_tile(mfic_ctx ctx, int length, int kind, int prec, int tile, int lds_pad) {
    return guard([&] {
        MFIC_CHECK(ctx, MFIC_ERR_ARG, "null context");
        MFIC_CHECK(fft_length_supported(length), MFIC_ERR_SIZE, "unsupported FFT length");
        MFIC_CHECK(kind >= 0 && kind <= MFIC_PASS_TRUNC_AX1, MFIC_ERR_ARG,
                   "kind must be 0 (strided pencils), 1 (rows), 2 (axis-0 pass with the shifted sub-lattice) or a "
                   "padded pass kind 3..8");
        MFIC_CHECK(prec == MFIC_F32 || prec == MFIC_F64, MFIC_ERR_ARG, "bad precision");
        const auto key = std::make_tuple(length, kind, prec);
        if (tile <= 0) { ctx->tiles.erase(key); return; }
        MFIC_CHECK((tile & (tile - 1)) == 0 && tile <= 64, MFIC_ERR_ARG, "tile must be a power of two <= 64");
        MFIC_CHECK(lds_pad >= 0 && lds_pad <= 8, MFIC_ERR_ARG, "lds_pad must be in [0, 8]");
        ctx->tiles[key] = {tile, lds_pad};
    });
}

int mfic_fft_get_tile(mfic_ctx ctx, int length, int kind, int prec, int* tile, int* lds_pad) {
    return guard([&] {
        MFIC_CHECK(ctx && tile && lds_pad, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(fft_length_supported(length), MFIC_ERR_SIZE, "unsupported FFT length");
        MFIC_CHECK(kind >= 0 && kind <= MFIC_PASS_TRUNC_AX1, MFIC_ERR_ARG, "bad tile kind");
        MFIC_CHECK(prec == MFIC_F32 || prec == MFIC_F64, MFIC_ERR_ARG, "bad precision");
        const auto it = ctx->tiles.find(std::make_tuple(length, kind, prec));
        *tile = it == ctx->tiles.end() ? 0 : it->second.first;
        *lds_pad = it == ctx->tiles.end() ? 0 : it->second.second;
    });
}

int mfic_fft_plan_count(int length) { return length > 0 ? fft_plan_count(length) : 0; }

int mfic_fft_plan_info(int length, int variant, int* elems_per_thread, int radices[4]) {
    return guard([&] {
        MFIC_CHECK(elems_per_thread && radices, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(length > 0 && fft_length_supported(length), MFIC_ERR_SIZE, "unsupported FFT length");
        fft_plan_info(length, variant, elems_per_thread, radices);
    });
}

static void check_plan_key(mfic_ctx ctx, int length, int pass, int prec) {
    MFIC_CHECK(ctx, MFIC_ERR_ARG, "null context");
    MFIC_CHECK(length > 0 && fft_length_supported(length), MFIC_ERR_SIZE, "unsupported FFT length");
    MFIC_CHECK(pass == MFIC_PASS_PENCILS || (pass >= MFIC_PASS_AX0_SHIFT && pass <= MFIC_PASS_FUSED_ROWS), MFIC_ERR_ARG,
               "pass must be one of mfic_fft_pass (row passes have one plan)");
    MFIC_CHECK(prec == MFIC_F32 || prec == MFIC_F64, MFIC_ERR_ARG, "bad precision");
}

int mfic_fft_set_plan(mfic_ctx ctx, int length, int pass, int prec, int variant) {
    return guard([&] {
        check_plan_key(ctx, length, pass, prec);
        const auto key = std::make_tuple(length, pass, prec);
        if (variant <= 0) { ctx->plans.erase(key); return; }
        if (pass == MFIC_PASS_FUSED_ROWS) {  // two forms at every padded length
            MFIC_CHECK(variant < 2, MFIC_ERR_ARG, "the fused row passes have two forms");
            ctx->plans[key] = variant;
            return;
        }
        MFIC_CHECK(variant < fft_plan_count(length), MFIC_ERR_ARG,
                   "length " + std::to_string(length) + " has " + std::to_string(fft_plan_count(length)) + " radix plan(s)");
        ctx->plans[key] = variant;
    });
}

int mfic_fft_get_plan(mfic_ctx ctx, int length, int pass, int prec, int* variant) {
    return guard([&] {
        check_plan_key(ctx, length, pass, prec);
        MFIC_CHECK(variant, MFIC_ERR_ARG, "null argument");
        *variant = plan_variant(ctx, length, pass, prec);
    });
}

void mfic_cosmo_defaults(mfic_cosmo_params* p) {
    if (p) cosmo_defaults(p);
}

int mfic_cosmo_compute(const mfic_cosmo_params* p, mfic_cosmo_scalars* out) {
    return guard([&] {
        MFIC_CHECK(p && out, MFIC_ERR_ARG, "null argument");
        cosmo_compute(*p, out, nullptr);
    });
}

int mfic_cosmo_eval(const mfic_cosmo_params* p, const char* what, const double* x, double* y, size_t count) {
    return guard([&] {
        MFIC_CHECK(p && what && x && y, MFIC_ERR_ARG, "null argument");
        cosmo_eval(*p, what, x, y, count);
    });
}

int mfic_phi1_from_noise(mfic_grid phi, mfic_grid wn, const mfic_cosmo_params* p, int fixing, int inversion) {
    return guard([&] {
        check_grid(phi);
        check_grid(wn);
        MFIC_CHECK(p, MFIC_ERR_ARG, "null cosmology");
        MFIC_CHECK(phi->n[0] == wn->n[0] && phi->n[1] == wn->n[1] && phi->n[2] == wn->n[2] && phi->prec == wn->prec,
                   MFIC_ERR_ARG, "phi and noise grids differ in size or precision");
        MFIC_CHECK(wn->space == MFIC_KSPACE, MFIC_ERR_STATE, "white noise must be in k-space");
        EhParams eh;
        mfic_cosmo_scalars s;
        cosmo_compute(*p, &s, &eh);
        check_tk_range(*p, phi->n, phi->L);
        const double volfac = std::pow(phi->L[0] / (double)phi->n[0] / 2.0 / 3.14159265358979323846, 1.5);
        // a table is copied to the device for this call and released after
        // the stream has consumed it
        void* dtk = tk_upload(*p, &eh);
        try {
            phi1_from_noise(phi, wn, eh, volfac, fixing, inversion);
        } catch (...) {
            if (dtk) { (void)hipStreamSynchronize(phi->ctx->stream); (void)hipFree(dtk); }
            throw;
        }
        if (dtk) {
            MFIC_HIP(hipStreamSynchronize(phi->ctx->stream));
            MFIC_HIP(hipFree(dtk));
        }
        phi->space = MFIC_KSPACE;
    });
}

}  // extern "C"
