/*
 * mfic_compat.hpp -- the reference's C++ interfaces for the hot path on top of
 * the mfic C ABI (header only, C++17), so code written against monofonIC's
 * classes compiles unchanged and runs on the device:
 *
 *   Grid_FFT<real_t>           grid_fft.hh:40-330 (non-distributed; float or
 *                              double real grids)
 *   op::assign_to / add_to /   operators.hh:25-35
 *   subtract_from / multiply_add_to
 *   OrszagConvolver<real_t>    convolution.hh:74-275, 412-651
 *   RNG_ngenic                 random_ngenic.cc:39-160 (Fill_Grid)
 *   TransferFunction_eisenstein transfer_eisenstein.cc:186-252 (compute,
 *                              get_kmin, get_kmax)
 *   calculator                 cosmology_calculator.hh:347-432 (the growth
 *                              factor, vfact and amplitude queries)
 *
 * Storage: a Grid_FFT owns one device grid (mfic_grid).  Transforms, the
 * k-space helpers and the convolutions run on the device.  Element access
 * (relem, kelem, the apply_function / assign_function lambdas) works on a
 * host mirror that is copied from the device on first access and back before
 * the next device operation: correct for reference-style code, but every
 * switch between host and device access moves the whole grid over PCIe, so
 * hot loops belong in the device entry points.
 *
 * Errors: every failing mfic call throws std::runtime_error with the
 * library's message, as the reference throws (grid_fft.cc:192).
 */
#ifndef MFIC_COMPAT_HPP
#define MFIC_COMPAT_HPP

#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "mfic.h"

namespace mfic_compat {

inline void check(int rc) {
    if (rc != MFIC_OK) throw std::runtime_error(std::string("mfic: ") + mfic_last_error());
}

/* the process-wide device context (the reference's global FFTW set-up,
 * main.cc:156-182): device 0 unless set_device() was called first */
inline int& device_index() {
    static int d = 0;
    return d;
}
inline void set_device(int d) { device_index() = d; }
struct ContextHolder {
    mfic_ctx c = nullptr;
    ~ContextHolder() {
        if (c) (void)mfic_ctx_destroy(c);
    }
};
inline ContextHolder& context_holder() {
    static ContextHolder h;
    return h;
}
inline mfic_ctx context() {
    ContextHolder& h = context_holder();
    if (!h.c) check(mfic_ctx_create(device_index(), &h.c));
    return h.c;
}
/* destroys the context (every grid and convolver must be gone); the next
 * context() call creates a new one */
inline void release_context() {
    ContextHolder& h = context_holder();
    if (h.c) check(mfic_ctx_destroy(h.c));
    h.c = nullptr;
}

enum space_t { kspace_id = 0, rspace_id = 1 };  // grid_fft.hh:29

/* the few vec3_t members the hot path uses (math/vec3.hh:22-135) */
template <typename T> struct vec3_t {
    T x = 0, y = 0, z = 0;
    vec3_t() = default;
    vec3_t(T a, T b, T c) : x(a), y(b), z(c) {}
    T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    const T& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T norm_squared() const { return x * x + y * y + z * z; }
    T norm() const { return std::sqrt(norm_squared()); }
};

template <typename data_t, bool bdistributed = false> class Grid_FFT {
    static_assert(!bdistributed, "one device holds a whole box: no slab distribution");
    static_assert(std::is_same<data_t, float>::value || std::is_same<data_t, double>::value,
                  "real FP32 or FP64 grids (complex grids: mfic_fft_c2c)");

  public:
    using real_t = data_t;
    using ccomplex_t = std::complex<data_t>;
    static constexpr int prec = std::is_same<data_t, double>::value ? MFIC_F64 : MFIC_F32;

    std::array<size_t, 3> n_;
    std::array<real_t, 3> length_;
    space_t space_;
    std::array<size_t, 4> sizes_{};
    size_t npr_ = 0, npc_ = 0, ntot_ = 0;

    // grid_fft.hh:77-83
    Grid_FFT(const std::array<size_t, 3>& N, const std::array<real_t, 3>& L, bool allocate = true,
             space_t initialspace = rspace_id)
        : n_(N), length_(L), space_(initialspace) {
        npr_ = 2 * (N[2] / 2 + 1);
        npc_ = N[2] / 2 + 1;
        ntot_ = N[0] * N[1] * npr_;
        set_sizes();
        if (allocate) this->allocate();
    }
    Grid_FFT(const Grid_FFT&) = delete;
    Grid_FFT& operator=(const Grid_FFT&) = delete;
    ~Grid_FFT() { reset(); }

    // grid_fft.cc:40-112 (device memory; the host mirror is made on demand)
    void allocate() {
        if (g_) return;
        const size_t n[3] = {n_[0], n_[1], n_[2]};
        const double L[3] = {length_[0], length_[1], length_[2]};
        check(mfic_grid_create(context(), n, L, prec, space_, &g_));
        host_valid_ = false;
        host_dirty_ = false;
    }
    // grid_fft.hh:95-101
    void reset() {
        if (g_) (void)mfic_grid_destroy(g_);
        g_ = nullptr;
        host_.clear();
        host_valid_ = host_dirty_ = false;
    }

    size_t size(size_t i) const { return sizes_[i]; }
    size_t global_size() const { return n_[0] * n_[1] * n_[2]; }
    /* the device layout is [n0][n1][npr] reals in real space and [n0][n1][npc]
     * complex values in k-space (no transposition on one device); for cubic
     * grids this is the reference's (i*sizes_[1]+j)*sizes_[3]+k */
    size_t get_idx(size_t i, size_t j, size_t k) const { return (i * n_[1] + j) * sizes_[3] + k; }
    mfic_grid handle() {
        to_device();
        return g_;
    }

    // grid_fft.cc:207-270: unitary r2c / c2r, a no-op when already there
    void FourierTransformForward(bool do_transform = true) {
        if (do_transform && space_ != kspace_id) check(mfic_fft_forward(handle()));
        else if (g_) check(mfic_grid_set_space(g_, MFIC_KSPACE));
        space_ = kspace_id;
        set_sizes();
        device_changed();
    }
    void FourierTransformBackward(bool do_transform = true) {
        if (do_transform && space_ != rspace_id) check(mfic_fft_backward(handle()));
        else if (g_) check(mfic_grid_set_space(g_, MFIC_RSPACE));
        space_ = rspace_id;
        set_sizes();
        device_changed();
    }

    // grid_fft.hh:140-147, 400-449, 832-850
    void zero() {
        check(mfic_grid_zero(handle()));
        device_changed();
    }
    void zero_DC_mode() {
        if (space_ == kspace_id) {
            check(mfic_grid_zero_dc(handle()));
            device_changed();
        } else {
            // real space: subtract the mean (grid_fft.hh:852-870)
            to_host();
            double s = 0.0;
            for (size_t i = 0; i < n_[0]; ++i)
                for (size_t j = 0; j < n_[1]; ++j)
                    for (size_t k = 0; k < n_[2]; ++k) s += relem(i, j, k);
            const real_t mean = (real_t)(s / (double)global_size());
            for (size_t i = 0; i < n_[0]; ++i)
                for (size_t j = 0; j < n_[1]; ++j)
                    for (size_t k = 0; k < n_[2]; ++k) relem(i, j, k) -= mean;
        }
    }
    void apply_InverseLaplacian() {
        FourierTransformForward();
        check(mfic_grid_inverse_laplacian(handle()));
        device_changed();
    }
    Grid_FFT& operator*=(real_t x) {
        check(mfic_grid_scale(handle(), (double)x));
        device_changed();
        return *this;
    }
    Grid_FFT& operator/=(real_t x) { return *this *= real_t(1) / x; }
    void shift_field(const vec3_t<real_t>& s, bool transform_back = true) {
        const double sh[3] = {s.x, s.y, s.z};
        check(mfic_grid_shift(handle(), sh, transform_back ? 1 : 0));
        space_ = transform_back ? rspace_id : kspace_id;
        set_sizes();
        device_changed();
    }
    // grid_fft.hh:149-170
    void copy_from(Grid_FFT& g) {
        check(mfic_grid_copy(handle(), g.handle()));
        space_ = g.space_;
        set_sizes();
        device_changed();
    }

    // grid_fft.hh:282-296 (index n/2 maps to +n/2) and 376-385 (0 at Nyquist)
    template <typename ft> vec3_t<ft> get_k(size_t i, size_t j, size_t k) const {
        const size_t ijk[3] = {i, j, k};
        vec3_t<ft> kv;
        for (int a = 0; a < 3; ++a) {
            const double s = (double)ijk[a] - (double)(ijk[a] > n_[a] / 2) * (double)n_[a];
            kv[a] = (ft)(s * 2.0 * M_PI / length_[a]);
        }
        return kv;
    }
    ccomplex_t gradient(int idim, std::array<size_t, 3> ijk) const {
        if (ijk[idim] == n_[idim] / 2) return ccomplex_t(0, 0);
        return ccomplex_t(0, get_k<real_t>(ijk[0], ijk[1], ijk[2])[idim]);
    }

    // host-mirror element access (grid_fft.hh:180-216)
    data_t& relem(size_t i, size_t j, size_t k) {
        to_host();
        host_dirty_ = true;
        return host_[get_idx(i, j, k)];
    }
    ccomplex_t& kelem(size_t i, size_t j, size_t k) {
        to_host();
        host_dirty_ = true;
        return reinterpret_cast<ccomplex_t*>(host_.data())[get_idx(i, j, k)];
    }
    ccomplex_t& kelem(size_t idx) {
        to_host();
        host_dirty_ = true;
        return reinterpret_cast<ccomplex_t*>(host_.data())[idx];
    }

    // grid_fft.hh:451-808, the lambda forms used by ic_generator (host mirror)
    template <typename F> void apply_function_k(F f) {
        for_k([&](size_t i, size_t j, size_t k) { kelem(i, j, k) = f(kelem(i, j, k)); });
    }
    template <typename F> void apply_function_k_dep(F f) {
        for_k([&](size_t i, size_t j, size_t k) { kelem(i, j, k) = f(kelem(i, j, k), get_k<real_t>(i, j, k)); });
    }
    template <typename F> void apply_function_r(F f) {
        for_r([&](size_t i, size_t j, size_t k) { relem(i, j, k) = f(relem(i, j, k)); });
    }
    template <typename F, typename... G> void assign_function_of_grids_kdep(F f, G&... g) {
        for_k([&](size_t i, size_t j, size_t k) {
            kelem(i, j, k) = f(get_k<real_t>(i, j, k), g.kelem(i, j, k)...);
        });
    }
    template <typename F, typename... G> void assign_function_of_grids_r(F f, G&... g) {
        for_r([&](size_t i, size_t j, size_t k) { relem(i, j, k) = f(g.relem(i, j, k)...); });
    }

    /* copy the whole buffer ([n0][n1][n2+2] reals) to / from host memory */
    void download(data_t* out) {
        to_device();
        check(mfic_grid_download(g_, out, prec));
    }
    void upload(const data_t* in) {
        host_valid_ = host_dirty_ = false;
        check(mfic_grid_upload(g_, in, prec));
    }

    void to_device() {
        if (!g_) throw std::runtime_error("mfic: grid not allocated");
        if (host_dirty_) {
            check(mfic_grid_upload(g_, host_.data(), prec));
            check(mfic_grid_set_space(g_, space_ == kspace_id ? MFIC_KSPACE : MFIC_RSPACE));
            host_dirty_ = false;
        }
    }
    void to_host() {
        if (!g_) throw std::runtime_error("mfic: grid not allocated");
        if (!host_valid_) {
            host_.resize(ntot_);
            check(mfic_grid_download(g_, host_.data(), prec));
            host_valid_ = true;
        }
    }

  private:
    mfic_grid g_ = nullptr;
    std::vector<data_t> host_;
    bool host_valid_ = false, host_dirty_ = false;

    void device_changed() { host_valid_ = host_dirty_ = false; }
    void set_sizes() {
        if (space_ == rspace_id) sizes_ = {n_[0], n_[1], n_[2], npr_};
        else sizes_ = {n_[1], n_[0], npc_, npc_};  // grid_fft.cc:105-111
    }
    // the loops run over whatever the buffer holds, as the reference's do
    template <typename F> void for_k(F f) {
        to_host();
        for (size_t i = 0; i < n_[0]; ++i)
            for (size_t j = 0; j < n_[1]; ++j)
                for (size_t k = 0; k < npc_; ++k) f(i, j, k);
    }
    template <typename F> void for_r(F f) {
        to_host();
        for (size_t i = 0; i < n_[0]; ++i)
            for (size_t j = 0; j < n_[1]; ++j)
                for (size_t k = 0; k < n_[2]; ++k) f(i, j, k);
    }
};

/* operators.hh:25-35: output operators of the convolver */
namespace op {
struct output {
    void* grid;
    mfic_grid (*handle)(void*);
    void (*done)(void*);
    int op;
    double coef;
};
template <typename field> output make(field& g, int o, double c) {
    return {&g, [](void* p) { return static_cast<field*>(p)->handle(); },
            [](void* p) {
                auto* f = static_cast<field*>(p);
                f->FourierTransformForward(false);
            },
            o, c};
}
template <typename field> output assign_to(field& g) { return make(g, MFIC_OP_ASSIGN, 0.0); }
template <typename field> output add_to(field& g) { return make(g, MFIC_OP_ADD, 0.0); }
template <typename field> output subtract_from(field& g) { return make(g, MFIC_OP_SUB, 0.0); }
template <typename field, typename val> output multiply_add_to(field& g, val x) {
    return make(g, MFIC_OP_MADD, (double)x);
}
}  // namespace op

/* convolution.hh:412-651 with the operand builders of convolution.hh:74-275 */
template <typename data_t> class OrszagConvolver {
  public:
    using grid_t = Grid_FFT<data_t>;
    OrszagConvolver(const std::array<size_t, 3>& N, const std::array<data_t, 3>& L) {
        const size_t n[3] = {N[0], N[1], N[2]};
        const double l[3] = {L[0], L[1], L[2]};
        check(mfic_conv_create(context(), n, l, grid_t::prec, &c_));
    }
    OrszagConvolver(const OrszagConvolver&) = delete;
    OrszagConvolver& operator=(const OrszagConvolver&) = delete;
    ~OrszagConvolver() { (void)mfic_conv_destroy(c_); }

    void convolve_Gradients(grid_t& inl, const std::array<int, 1>& d1l, grid_t& inr, const std::array<int, 1>& d1r,
                            op::output out) {
        run2(opnd(inl, MFIC_OPND_GRAD, {d1l[0], 0, 0, 0}), opnd(inr, MFIC_OPND_GRAD, {d1r[0], 0, 0, 0}), out);
    }
    void convolve_Gradient_and_Hessian(grid_t& inl, const std::array<int, 1>& d1l, grid_t& inr,
                                       const std::array<int, 2>& d2r, op::output out) {
        run2(opnd(inl, MFIC_OPND_GRAD, {d1l[0], 0, 0, 0}), opnd(inr, MFIC_OPND_HESS, {d2r[0], d2r[1], 0, 0}), out);
    }
    void convolve_Hessians(grid_t& inl, const std::array<int, 2>& d2l, grid_t& inr, const std::array<int, 2>& d2r,
                           op::output out) {
        run2(opnd(inl, MFIC_OPND_HESS, {d2l[0], d2l[1], 0, 0}), opnd(inr, MFIC_OPND_HESS, {d2r[0], d2r[1], 0, 0}),
             out);
    }
    void convolve_Hessians(grid_t& inl, const std::array<int, 2>& d2l, grid_t& inr, const std::array<int, 2>& d2r,
                           grid_t& inm, const std::array<int, 2>& d2m, op::output out) {
        const mfic_operand a = opnd(inl, MFIC_OPND_HESS, {d2l[0], d2l[1], 0, 0});
        const mfic_operand b = opnd(inr, MFIC_OPND_HESS, {d2r[0], d2r[1], 0, 0});
        const mfic_operand c = opnd(inm, MFIC_OPND_HESS, {d2m[0], d2m[1], 0, 0});
        check(mfic_conv_convolve3(c_, &a, &b, &c, out.handle(out.grid), out.op, out.coef));
        out.done(out.grid);
    }
    void convolve_SumOfHessians(grid_t& inl, const std::array<int, 2>& d2l, grid_t& inr,
                                const std::array<int, 2>& d2r1, const std::array<int, 2>& d2r2, op::output out) {
        run2(opnd(inl, MFIC_OPND_HESS, {d2l[0], d2l[1], 0, 0}),
             opnd(inr, MFIC_OPND_SUM_HESS, {d2r1[0], d2r1[1], d2r2[0], d2r2[1]}), out);
    }
    void convolve_DifferenceOfHessians(grid_t& inl, const std::array<int, 2>& d2l, grid_t& inr,
                                       const std::array<int, 2>& d2r1, const std::array<int, 2>& d2r2,
                                       op::output out) {
        run2(opnd(inl, MFIC_OPND_HESS, {d2l[0], d2l[1], 0, 0}),
             opnd(inr, MFIC_OPND_DIFF_HESS, {d2r1[0], d2r1[1], d2r2[0], d2r2[1]}), out);
    }

  private:
    mfic_conv c_ = nullptr;
    // the operands are taken to k-space first, as the reference does
    static mfic_operand opnd(grid_t& g, int kind, std::array<int, 4> d) {
        g.FourierTransformForward();
        mfic_operand o;
        o.f = g.handle();
        o.kind = kind;
        for (int i = 0; i < 4; ++i) o.d[i] = d[i];
        return o;
    }
    void run2(const mfic_operand& a, const mfic_operand& b, op::output out) {
        check(mfic_conv_convolve2(c_, &a, &b, out.handle(out.grid), out.op, out.coef));
        out.done(out.grid);
    }
};

/* random_ngenic.cc:39-160: the NGENIC generator's Fill_Grid (bit-exact
 * RANLUX stream, device fill), constructed from the [random] seed */
class RNG_ngenic {
  public:
    explicit RNG_ngenic(long seed) : seed_(seed) {}
    bool isMultiscale() const { return false; }
    template <typename data_t> void Fill_Grid(Grid_FFT<data_t>& g) {
        check(mfic_ngenic_fill(g.handle(), seed_));
        g.FourierTransformForward(false);
    }

  private:
    long seed_;
};

/* cosmology::calculator queries (cosmology_calculator.hh:347-432) and the
 * Eisenstein & Hu transfer plugin (transfer_eisenstein.cc:186-252) */
enum tf_type { delta_matter, delta_cdm, delta_baryon, theta_matter, theta_cdm, theta_baryon, delta_bc, theta_bc };

class TransferFunction_eisenstein {
  public:
    explicit TransferFunction_eisenstein(const mfic_cosmo_params& p) : p_(p) {}
    double compute(double k, tf_type type) const {
        if (type == delta_bc || type == theta_bc) return 0.0;
        double T = 0.0;
        check(mfic_cosmo_eval(&p_, "T", &k, &T, 1));
        return T;
    }
    double get_kmin() const { return 1e-4; }
    double get_kmax() const { return 1e4; }

  private:
    mfic_cosmo_params p_;
};

class calculator {
  public:
    explicit calculator(const mfic_cosmo_params& p) : p_(p), tf_(p) {
        if (p.tk_n > 0)
            throw std::runtime_error("mfic: calculator wraps the Eisenstein & Hu fit; tabulated T(k) goes "
                                     "through mfic_phi1_from_noise / mfic_run_create");
        check(mfic_cosmo_compute(&p_, &s_));
    }
    double get_growth_factor(double a) const { return eval("D", a); }
    double get_f(double a) const { return eval("f", a); }
    double get_vfact(double a) const { return eval("vfact", a); }
    // k^{n_s/2} T(k) sqrt(pnorm) (cosmology_calculator.hh:429-432), k in h/Mpc
    double get_amplitude(double k, tf_type type) const {
        return std::pow(k, 0.5 * p_.n_s) * tf_.compute(k, type) * s_.sqrtpnorm;
    }
    const mfic_cosmo_scalars& scalars() const { return s_; }

  private:
    double eval(const char* what, double x) const {
        double y = 0.0;
        check(mfic_cosmo_eval(&p_, what, &x, &y, 1));
        return y;
    }
    mfic_cosmo_params p_;
    TransferFunction_eisenstein tf_;
    mfic_cosmo_scalars s_{};
};

}  // namespace mfic_compat

#endif /* MFIC_COMPAT_HPP */
