// Internal runtime of the mfic library: contexts, device grids, errors and
// the twiddle-table cache.  Host C++ (compiled by hipcc), not exported.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <atomic>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mfic.h"

namespace mfic {

// ---------------------------------------------------------------- errors --
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void set_last_error(const std::string& msg);

#define MFIC_HIP(call)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::mfic::Error(MFIC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define MFIC_CHECK(cond, code, msg)                  \
    do {                                             \
        if (!(cond)) throw ::mfic::Error(code, msg); \
    } while (0)
#define MFIC_LAUNCH_CHECK() MFIC_HIP(hipGetLastError())

// ---------------------------------------------------------------- tables --
// device copy of the twiddle table for one transform length and precision
struct TwTable {
    void* dev = nullptr;
    size_t count = 0;
};

}  // namespace mfic

// ------------------------------------------------------------- handles ----
struct mfic_prof_rec {
    std::string name;
    double bytes;  // algorithmic bytes of the launch
    hipEvent_t a, b;
};

struct mfic_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    // (length, kind [0 = complex, 1 = real-row], precision) -> table
    std::map<std::tuple<int, int, int>, mfic::TwTable> tables;
    std::mutex mu;
    // scratch buffers owned by the context (reused across calls)
    std::map<std::string, std::pair<void*, size_t>> scratch;
    size_t max_lds = 0;  // LDS bytes per workgroup (device attribute)
    // tile overrides from the autotuner: (length, kind, prec) -> (tile, lds pad)
    std::map<std::tuple<int, int, int>, std::pair<int, int>> tiles;
    // radix plan choices: (length, pass [mfic_fft_pass], prec) -> plan variant
    std::map<std::tuple<int, int, int>, int> plans;
    // stream capture into a HIP graph in progress (mfic_graph_begin): no
    // allocation, table creation or event timing may happen meanwhile
    bool capturing = false;
    // per-launch event timing (mfic_prof_enable)
    bool prof = false;
    std::vector<mfic_prof_rec> prof_recs;
    std::vector<hipEvent_t> prof_pool;
    // live handles created on this context (grids, convolvers, PLT tables,
    // runs, graphs, 2LPT plans): mfic_ctx_destroy refuses while any remain,
    // since each of them uses the context's stream when it is destroyed
    std::atomic<long> children{0};
};

struct mfic_grid_s {
    mfic_ctx ctx = nullptr;
    size_t n[3] = {0, 0, 0};
    double L[3] = {0, 0, 0};
    int prec = MFIC_F32;
    int space = MFIC_RSPACE;
    void* d = nullptr;
    size_t nreal = 0;  // n0*n1*(n2+2) reals
    bool owns = true;
    size_t elem_size() const { return prec == MFIC_F64 ? 8 : 4; }
    size_t bytes() const { return nreal * elem_size(); }
    size_t nc() const { return n[2] / 2 + 1; }
};

namespace mfic {

// Records HIP events around the launches in its scope when profiling is on
// (same stream as the kernels, so only the bracketed kernels are timed).
class ProfScope {
  public:
    ProfScope(mfic_ctx ctx, const std::string& name, double bytes);
    ~ProfScope();
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;

  private:
    mfic_ctx ctx_;
    long idx_ = -1;
};

// scratch buffer of at least `bytes`, keyed by name, owned by the context
void* ctx_scratch(mfic_ctx ctx, const std::string& key, size_t bytes);

// refuses (MFIC_ERR_ARG) a caller pointer that kernels would access but that
// is not device or managed memory: a kernel store to pageable host memory
// faults the device
void require_device_memory(const void* p, const char* what);

// twiddles: complex transform of length L (kind 0), or real rows of length
// 2M (kind 1: complex length M table followed by M+1 untangle factors), or
// the half-cell shift phases exp(i pi m / L), m < 2L (kind 2, FP64); the
// inter-stage factors follow radix plan `variant` of the length
const void* twiddles(mfic_ctx ctx, int L, int kind, int prec, int variant = 0);
// the radix plan selected for a pass (mfic_fft_pass) of length L (0 unless
// mfic_fft_set_plan chose another)
int plan_variant(mfic_ctx ctx, int L, int pass, int prec);
// number of radix plans compiled for length L
int fft_plan_count(int L);
// elements per thread and radices of plan `variant` of length L
void fft_plan_info(int L, int variant, int* elems, int radices[4]);

// FFT entry points on raw device buffers of layout [n0][n1][n2+2]
template <typename T> void fft_r2c(mfic_ctx ctx, T* data, const size_t n[3], double scale);
template <typename T> void fft_c2r(mfic_ctx ctx, T* data, const size_t n[3], double scale);
bool fft_length_supported(size_t n);

// pass-level entry points used by fused pipelines
template <typename T> void fft_c2r_cols(mfic_ctx ctx, T* data, const size_t n[3]);
// fft_c2r_cols of a cubic-grid component plus, from the same axis-0 loads,
// the component of the (1/2,1/2,1/2)-cell shifted sub-lattice into `shifted`;
// with weights (wa, wb not null) the in-place output is shifted by wa / 2
// and the second by wb / 2 cells per axis (0, 1 or 2 each)
template <typename T>
void fft_c2r_cols_shift(mfic_ctx ctx, T* data, T* shifted, const size_t n[3], const int* wa = nullptr,
                        const int* wb = nullptr);
template <typename T> void fft_rows_r2c(mfic_ctx ctx, T* data, size_t nrows, int n2, double scale);
template <typename T> void fft_rows_c2r(mfic_ctx ctx, T* data, size_t nrows, int n2, double scale);
// complex pencils of length L: element (o, p, t) at data + o*os + p*ps + t,
// t < tlen, o < nouter (offsets in complex elements)
template <typename T>
void fft_pencils(mfic_ctx ctx, void* data, int L, long ps, long os, int tlen, long nouter, bool fwd,
                 double scale, const char* tag = "");

// a tabulated transfer function must reach the largest |k| of an n[0..2]
// grid of sides L (h/Mpc); beyond its last row the reference's spline fails
inline void check_tk_range(const mfic_cosmo_params& p, const size_t n[3], const double L[3]) {
    if (p.tk_n == 0) return;
    double k2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double ka = 3.14159265358979323846 * (double)n[a] / L[a];
        k2 += ka * ka;
    }
    if (std::sqrt(k2) > p.tk_k[p.tk_n - 1])
        throw Error(MFIC_ERR_CONFIG, "tabulated transfer function ends at k = " + std::to_string(p.tk_k[p.tk_n - 1]) +
                                         " h/Mpc, the grid reaches " + std::to_string(std::sqrt(k2)));
}

// run f, translating exceptions into a status code and the thread-local
// error message (no exception crosses the C ABI)
template <typename F> int guard_call(F&& f) {
    try {
        f();
        return MFIC_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return MFIC_ERR_CONFIG;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return MFIC_ERR_ALLOC;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MFIC_ERR_ARG;
    }
}

}  // namespace mfic
