// C ABI, batching: HIP-graph capture of any sequence of library calls on a
// context stream (many boxes / seeds replayed by one graph launch) and the
// tile autotuner.  No reference counterpart: monofonIC runs one box per
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

// Coordinate descent over the tile of each transform length a box of size n
// uses (pencils per block for the strided N-grid and padded passes, rows per
// block for the contiguous N-grid pass), timing whole pipeline executions of
// a temporary run with HIP events.  Leaves the winners in the context.
int mfic_autotune(mfic_ctx ctx, size_t n, int prec, int lpt_order, double* best_ms) {
    return mfic::guard_call([&] {
        MFIC_CHECK(ctx, MFIC_ERR_ARG, "null context");
        MFIC_CHECK(!ctx->capturing, MFIC_ERR_STATE, "cannot autotune while capturing");
        MFIC_CHECK(mfic_fft_size_supported(n) && n % 2 == 0, MFIC_ERR_SIZE, "unsupported grid size");
        MFIC_CHECK(lpt_order >= 1 && lpt_order <= 3, MFIC_ERR_CONFIG, "LPT order must be 1..3");
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
        mfic_run run = nullptr;
        // a failing call's own status code is passed on (a size or
        // configuration error is not a HIP error)
        auto ck = [](int rc) {
            if (rc != MFIC_OK) throw mfic::Error(rc, mfic_last_error());
        };
        ck(mfic_run_create(ctx, &cfg, &cp, &run));
        hipEvent_t a, b;
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
        double result = 0.0;
        try {
            ck(mfic_run_execute(run, 1));
            struct Knob { int length, kind; std::vector<int> cand; };
            std::vector<Knob> knobs = {{(int)n, 0, {4, 8, 16, 32, 64}}, {(int)n / 2, 1, {1, 2, 4, 8, 16}}};
            if (lpt_order > 1) knobs.push_back({(int)(3 * n / 2), 0, {4, 8, 16, 32, 64}});
            float cur = time_once();
            for (const auto& k : knobs) {
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
            result = cur;
        } catch (...) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            (void)mfic_run_destroy(run);
            throw;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        ck(mfic_run_destroy(run));
        if (best_ms) *best_ms = result;
    });
}

}  // extern "C"
