// C ABI, batching: HIP-graph capture of any sequence of library calls on a
// context stream (many boxes / seeds replayed by one graph launch) and the
// tile and radix plan autotuner.  No reference counterpart: monofonIC runs one box per
// process and leaves transform planning to FFTW (grid_fft.cc:53-54).
#include <algorithm>
#include <cmath>
#include <vector>

#include "mfic_internal.hpp"

struct mfic_graph_s {
    mfic_ctx ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};


extern "C" {

int mfic_graph_begin(mfic_ctx ctx) {
    return mfic::guard_call([&] {
        MFIC_CHECK(ctx, MFIC_ERR_ARG, "null context");
        MFIC_CHECK(!ctx->capturing, MFIC_ERR_STATE, "a capture is already in progress on this context");
        MFIC_HIP(hipSetDevice(ctx->device));
        MFIC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        ctx->capturing = true;
    });
}

int mfic_graph_end(mfic_ctx ctx, mfic_graph* out) {
    return mfic::guard_call([&] {
        MFIC_CHECK(ctx && out, MFIC_ERR_ARG, "null argument");
        MFIC_CHECK(ctx->capturing, MFIC_ERR_STATE, "no capture in progress on this context");
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        ctx->capturing = false;
        if (e != hipSuccess || !g)
            throw mfic::Error(MFIC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e) +
                                                " (a call inside the capture failed or synchronised)");
        auto* r = new mfic_graph_s();
        r->ctx = ctx;
        r->graph = g;
        const hipError_t ei = hipGraphInstantiate(&r->exec, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            delete r;
            throw mfic::Error(MFIC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        ++ctx->children;
        *out = r;
    });
}

int mfic_graph_launch(mfic_graph g, int blocking) {
    return mfic::guard_call([&] {
        MFIC_CHECK(g && g->exec, MFIC_ERR_ARG, "null graph");
        MFIC_CHECK(!g->ctx->capturing, MFIC_ERR_STATE, "cannot launch a graph while capturing");
        MFIC_HIP(hipGraphLaunch(g->exec, g->ctx->stream));
        if (blocking) MFIC_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int mfic_graph_destroy(mfic_graph g) {
    return mfic::guard_call([&] {
        if (!g) return;
        (void)hipStreamSynchronize(g->ctx->stream);
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        mfic_ctx ctx = g->ctx;
        delete g;
        --ctx->children;
    });
}

namespace {

// Coordinate descent over the radix plan of every strided pass kind and the
// tile of every transform length a run configuration uses (pencils per
// block for the strided N-grid and padded passes, then for each padded pass
// kind and the axis-0 pass of the shifted sub-lattice, rows per block for
// the contiguous N-grid pass),
// timing whole executions of a temporary run of that configuration with HIP
// events: plans, then tiles, then the plans again under the chosen tiles.
// Leaves the winners in the context and returns the best time.
double autotune_config(mfic_ctx ctx, const mfic_ic_config& cfg, const mfic_cosmo_params& cp) {
    MFIC_CHECK(!ctx->capturing, MFIC_ERR_STATE, "cannot autotune while capturing");
    const int n = (int)cfg.n, prec = cfg.prec;
    mfic_run run = nullptr;
    // a failing call's own status code is passed on (a size or
    // configuration error is not a HIP error)
    auto ck = [](int rc) {
        if (rc != MFIC_OK) throw mfic::Error(rc, mfic_last_error());
    };
    ck(mfic_run_create(ctx, &cfg, &cp, &run));
    hipEvent_t a = nullptr, b = nullptr;
    auto release = [&] {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        (void)mfic_run_destroy(run);
    };
    double result = 0.0;
    try {
        MFIC_HIP(hipEventCreate(&a));
        MFIC_HIP(hipEventCreate(&b));
        auto time_once = [&]() {
            float best = 1e30f;
            for (int rep = 0; rep < 3; ++rep) {
                MFIC_HIP(hipEventRecord(a, ctx->stream));
                ck(mfic_run_execute(run, 0));
                MFIC_HIP(hipEventRecord(b, ctx->stream));
                MFIC_HIP(hipEventSynchronize(b));
                float ms = 0.f;
                MFIC_HIP(hipEventElapsedTime(&ms, a, b));
                best = std::min(best, ms);
            }
            return best;
        };
        const bool shift = cfg.particles && cfg.lattice == MFIC_LATTICE_BCC;
        const int np = 3 * n / 2;
        // plan knobs: (length, pass, variants) with more than one variant
        struct PlanKnob { int length, pass, count; };
        std::vector<PlanKnob> plan_knobs;
        if (mfic::fft_plan_count(n) > 1) {
            plan_knobs.push_back({n, MFIC_PASS_PENCILS, mfic::fft_plan_count(n)});
            if (shift) plan_knobs.push_back({n, MFIC_PASS_AX0_SHIFT, mfic::fft_plan_count(n)});
        }
        if (cfg.lpt_order > 1 && mfic::fft_plan_count(np) > 1)
            for (int pass = MFIC_PASS_PAD_C2R_A; pass <= MFIC_PASS_TRUNC_AX1; ++pass)
                plan_knobs.push_back({np, pass, mfic::fft_plan_count(np)});
        // the two forms of the fused source row passes
        if (cfg.lpt_order > 1) plan_knobs.push_back({np, MFIC_PASS_FUSED_ROWS, 2});
        struct Knob { int length, kind; std::vector<int> cand; };
        std::vector<Knob> tile_knobs = {{n, 0, {4, 8, 16, 32, 64}}, {n / 2, 1, {1, 2, 4, 8, 16}}};
        if (shift) tile_knobs.push_back({n, 2, {4, 8, 16, 32}});
        if (cfg.lpt_order > 1) {
            // one tile for every padded pass, then each pass kind on its own
            tile_knobs.push_back({np, 0, {4, 8, 16, 32, 64}});
            for (int pass = MFIC_PASS_PAD_C2R_A; pass <= MFIC_PASS_TRUNC_AX1; ++pass)
                tile_knobs.push_back({np, pass, {4, 8, 16, 32}});
        }

        ck(mfic_run_execute(run, 1));
        float cur = time_once();
        auto tune_plans = [&] {
            for (const auto& k : plan_knobs) {
                const auto key = std::make_tuple(k.length, k.pass, prec);
                const int had = mfic::plan_variant(ctx, k.length, k.pass, prec);
                int best_v = had;
                float best = cur;
                for (int v = 0; v < k.count; ++v) {
                    if (v == had) continue;
                    if (v) ctx->plans[key] = v;
                    else ctx->plans.erase(key);
                    const float t = time_once();
                    if (t < best) { best = t; best_v = v; }
                }
                if (best_v) ctx->plans[key] = best_v;
                else ctx->plans.erase(key);
                cur = best;
            }
        };
        tune_plans();
        for (const auto& k : tile_knobs) {
            const auto key = std::make_tuple(k.length, k.kind, prec);
            const auto had = ctx->tiles.find(key);
            std::pair<int, int> best_tile = had != ctx->tiles.end() ? had->second : std::make_pair(0, 0);
            float best = cur;
            for (int c : k.cand) {
                ctx->tiles[key] = {c, 0};
                const float t = time_once();
                if (t < best) { best = t; best_tile = {c, 0}; }
            }
            if (best_tile.first > 0) ctx->tiles[key] = best_tile;
            else ctx->tiles.erase(key);
            cur = best;
        }
        tune_plans();
        result = cur;
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        release();
        throw;
    }
    release();
    return result;
}

}  // namespace

// The pipeline of an n^3 box of the given order and precision (sc lattice,
// component fields, the benchmark cosmology).
int mfic_autotune(mfic_ctx ctx, size_t n, int prec, int lpt_order, double* best_ms) {
    return mfic::guard_call([&] {
        MFIC_CHECK(ctx, MFIC_ERR_ARG, "null context");
        MFIC_CHECK(mfic_fft_size_supported(n) && n % 2 == 0, MFIC_ERR_SIZE, "unsupported grid size");
        MFIC_CHECK(lpt_order >= 1 && lpt_order <= 3, MFIC_ERR_CONFIG, "LPT order must be 1..3");
        MFIC_CHECK(prec == MFIC_F32 || prec == MFIC_F64, MFIC_ERR_ARG, "bad precision");
        mfic_cosmo_params cp;
        mfic_cosmo_defaults(&cp);
        cp.Omega_m = 0.3;
        cp.Omega_b = 0.045;
        cp.H0 = 70.0;
        cp.sigma_8 = 0.81;
        cp.A_s = -1.0;
        cp.n_s = 0.961;
        cp.zstart = 24.0;
        mfic_ic_config cfg{};
        cfg.n = n;
        cfg.L = (double)n;
        cfg.lpt_order = lpt_order;
        cfg.lattice = MFIC_LATTICE_SC;
        cfg.prec = prec;
        cfg.seed = 1;
        cfg.lunit = cfg.vunit = 1.0;
        const double ms = autotune_config(ctx, cfg, cp);
        if (best_ms) *best_ms = ms;
    });
}

int mfic_autotune_config(mfic_ctx ctx, const mfic_ic_config* cfg, const mfic_cosmo_params* cosmo, double* best_ms) {
    return mfic::guard_call([&] {
        MFIC_CHECK(ctx && cfg && cosmo, MFIC_ERR_ARG, "null argument");
        const double ms = autotune_config(ctx, *cfg, *cosmo);
        if (best_ms) *best_ms = ms;
    });
}

}  // extern "C"
