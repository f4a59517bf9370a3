// Host cosmology scalars for the IC hot path (FP64, computed once per run):
// derived parameters, linear growth D+(a) and f(a), v-factor, sigma_8
// normalisation and the Eisenstein & Hu fit parameters consumed on the device.
//
//   parameters   cosmology_parameters.hh:80-241, cosmology_parameters.cc:110-169
//   H(a)         cosmology_calculator.hh:334-342
//   growth       cosmology_calculator.hh:92-152, adaptive Cash-Karp step of
//                math/ode_integrate.hh:39-111 (a rejected trial restarts from
//                the state the rejected trial produced, as there)
//   D(a), f(a)   natural cubic spline in (ln a, ln y), math/interpolate.hh
//   pnorm        cosmology_calculator.hh:487-511 (sigma_8 integral by adaptive
//                Gauss-Kronrod 7/15, rel. 1e-13)
//   E&H          transfer_eisenstein.cc:58-122, 186-201
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "cosmology.hpp"

namespace mfic {

namespace {
constexpr double kPi = 3.141592653589793115997963468544185161590576171875;
constexpr double kMpc = 3.0857e22, kC = 2.99792458e8, kG = 6.6740800e-11;
constexpr double kKB = 1.38064852e-23, kHbar = 1.054571800e-34;

using V3 = std::array<double, 3>;

struct Spline {  // natural cubic spline y(x)
    std::vector<double> x, y, m;
    void init(std::vector<double> xs, std::vector<double> ys) {
        x = std::move(xs);
        y = std::move(ys);
        const size_t n = x.size();
        m.assign(n, 0.0);
        std::vector<double> u(n, 0.0);
        for (size_t i = 1; i + 1 < n; ++i) {
            const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
            const double p = sig * m[i - 1] + 2.0;
            m[i] = (sig - 1.0) / p;
            const double d = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
            u[i] = (6.0 * d / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
        }
        m[n - 1] = 0.0;
        for (size_t k = n - 1; k-- > 0;) m[k] = m[k] * m[k + 1] + u[k];
    }
    double operator()(double xv) const {
        size_t lo = 0, hi = x.size() - 1;
        while (hi - lo > 1) {
            const size_t mid = (lo + hi) / 2;
            (x[mid] > xv ? hi : lo) = mid;
        }
        const double h = x[hi] - x[lo], A = (x[hi] - xv) / h, B = (xv - x[lo]) / h;
        return A * y[lo] + B * y[hi] + ((A * A * A - A) * m[lo] + (B * B * B - B) * m[hi]) * (h * h) / 6.0;
    }
};

struct Calc {
    mfic_cosmo_params p;
    mfic_cosmo_scalars s{};
    EhParams eh{};
    Spline lnD, lnf;

    double H(double a) const {
        double h2 = s.Omega_r / (a * a * a * a) + p.Omega_m / (a * a * a) + p.Omega_k / (a * a) +
                    s.Omega_DE * std::pow(a, -3. * (1. + p.w_0 + p.w_a)) * std::exp(-3. * (1.0 - a) * p.w_a);
        return p.H0 * std::sqrt(h2);
    }
    V3 rhs(const V3& y) const {
        const double a = y[0], Hv = H(a);
        return {a * a * Hv, y[2], -a * Hv * y[2] + 3.0 / 2.0 * p.Omega_m * std::pow(p.H0, 2) * y[1] / a};
    }
    // Cash-Karp 5(4): advances y, returns the embedded error estimate
    V3 ck_step(double h, V3& y) const {
        constexpr double b21 = 0.2, b31 = 3.0 / 40.0, b32 = 9.0 / 40.0, b41 = 0.3, b42 = -0.9, b43 = 1.2,
                         b51 = -11.0 / 54.0, b52 = 2.5, b53 = -70.0 / 27.0, b54 = 35.0 / 27.0,
                         b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0, b63 = 575.0 / 13824.0,
                         b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0, c1 = 37.0 / 378.0,
                         c3 = 250.0 / 621.0, c4 = 125.0 / 594.0, c6 = 512.0 / 1771.0;
        constexpr double dc1 = c1 - 2825.0 / 27648.0, dc3 = c3 - 18575.0 / 48384.0,
                         dc4 = c4 - 13525.0 / 55296.0, dc5 = -277.0 / 14336.0, dc6 = c6 - 0.25;
        auto mul = [h](V3 v) { return V3{h * v[0], h * v[1], h * v[2]}; };
        V3 k1 = mul(rhs(y)), t, k2, k3, k4, k5, k6;
        for (int i = 0; i < 3; ++i) t[i] = y[i] + b21 * k1[i];
        k2 = mul(rhs(t));
        for (int i = 0; i < 3; ++i) t[i] = y[i] + b31 * k1[i] + b32 * k2[i];
        k3 = mul(rhs(t));
        for (int i = 0; i < 3; ++i) t[i] = y[i] + b41 * k1[i] + b42 * k2[i] + b43 * k3[i];
        k4 = mul(rhs(t));
        for (int i = 0; i < 3; ++i) t[i] = y[i] + b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i];
        k5 = mul(rhs(t));
        for (int i = 0; i < 3; ++i)
            t[i] = y[i] + b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i];
        k6 = mul(rhs(t));
        V3 err;
        for (int i = 0; i < 3; ++i) {
            y[i] += c1 * k1[i] + c3 * k3[i] + c4 * k4[i] + c6 * k6[i];
            err[i] = dc1 * k1[i] + dc3 * k3[i] + dc4 * k4[i] + dc5 * k5[i] + dc6 * k6[i];
        }
        return err;
    }
    void growth() {
        const double a0 = 1e-10;
        V3 y = {a0, a0, 2.0 * a0 * H(a0) / std::pow(kC, 2)};
        double t = 1.0 / (a0 * H(a0)), dt = 1e-9;
        std::vector<double> ta, tD, tf;
        while (y[0] < 2.0) {
            const V3 d = rhs(y);
            V3 sc;
            for (int i = 0; i < 3; ++i) sc[i] = std::fabs(y[i]) + dt * std::fabs(d[i]);
            double h = dt, errmax;
            V3 yt = y;  // persists across rejected trials
            for (;;) {
                const V3 e = ck_step(h, yt);
                errmax = 0.0;
                for (int i = 0; i < 3; ++i) errmax = std::max(errmax, std::fabs(e[i] / sc[i]));
                errmax /= 1e-10;
                if (errmax <= 1.0) break;
                h *= std::max(0.1, 0.9 * std::pow(errmax, -0.25));
                if (t + h == t) throw std::runtime_error("growth ODE: stepsize underflow");
            }
            dt = errmax > 1.89e-4 ? h * 0.9 * std::pow(errmax, -0.2) : 5 * h;
            t += h;
            y = yt;
            ta.push_back(y[0]);
            tD.push_back(y[1]);
            tf.push_back(y[2]);
        }
        for (size_t i = 0; i < ta.size(); ++i) {
            tf[i] = std::log(tf[i] / (ta[i] * H(ta[i]) * tD[i]));
            tD[i] = std::log(tD[i]);
            ta[i] = std::log(ta[i]);
        }
        lnD.init(ta, tD);
        lnf.init(ta, tf);
    }
    double Draw(double a) const { return std::exp(lnD(std::log(a))); }
    double f(double a) const { return std::exp(lnf(std::log(a))); }

    std::vector<double> tk;  // host spline nodes of a tabulated transfer function
    double sig8_integrand(double k) const {
        const double x = k * 8.0;
        const double w = x < 0.001 ? 1.0 - 0.1 * x * x : 3.0 * (std::sin(x) - x * std::cos(x)) / (x * x * x);
        const double tf = eh_transfer(eh, k);
        return k * k * w * w * std::pow(k, p.n_s) * tf * tf;
    }
    double integrate(double a, double b, double rel) const;

    void run() {
        s.h = p.H0 / 100.0;
        const double h = s.h;
        const double summ = p.m_nu[0] * p.deg_nu[0] + p.m_nu[1] * p.deg_nu[1] + p.m_nu[2] * p.deg_nu[2];
        s.Omega_nu_massive = summ / (p.O_nu_norm * h * h);
        const double sigmaSB = (kPi * kPi) * (kKB * kKB * kKB * kKB) / 60. / (kHbar * kHbar * kHbar) / (kC * kC);
        const double rhoc = 3 * 1e10 / (8 * kPi * kG) / kMpc / kMpc;
        s.Omega_gamma = 4 * sigmaSB / std::pow(kC, 3) * std::pow(p.Tcmb, 4.0) / rhoc / (h * h);
        s.Omega_r = s.Omega_gamma + p.N_ur * s.Omega_gamma * 7. / 8. * std::pow(4. / 11., 4. / 3.);
        s.Omega_c = p.Omega_m - p.Omega_b - s.Omega_nu_massive;
        if (p.zero_radiation) s.Omega_r = 0.0;
        s.Omega_DE = 1.0 - p.Omega_m - s.Omega_r - p.Omega_k;
        growth();
        s.Dnow = Draw(1.0);
        const double astart = 1.0 / (1.0 + p.zstart), atarget = 1.0 / (1.0 + p.ztarget);
        s.Dplus_start = Draw(astart) / s.Dnow;
        s.Dplus_target = Draw(atarget) / s.Dnow;
        eh = eh_setup(p);
        // sigma_8 over the transfer function's own range: [1e-4, 1e4] h/Mpc
        // for E&H (transfer_eisenstein.cc:245-253), the second to the
        // next-to-last tabulated k for a table (transfer_CAMB_file.cc:156-158)
        double kmin = 1e-4, kmax = 1e4;
        if (p.tk_n > 0) {
            tk = tk_spline_build(p.tk_k, p.tk_T, p.tk_n);
            eh.tk = tk.data();
            eh.tk_n = (int)p.tk_n;
            eh.tk_lkmin = std::log(p.tk_k[1]);
            kmin = p.tk_k[1];
            kmax = p.tk_k[p.tk_n - 2];
        }
        s.sigma8_measured = std::sqrt(4.0 * kPi * integrate(kmin, kmax, 1e-13));
        if (p.A_s > -1.0)
            s.pnorm = p.A_s * std::pow(p.k_p / h, 1.0 - p.n_s) * (h * h * h * h) /
                      (4 * kPi * s.Dplus_target * s.Dplus_target);
        else
            s.pnorm = p.sigma_8 * p.sigma_8 / (s.sigma8_measured * s.sigma8_measured);
        s.sqrtpnorm = std::sqrt(s.pnorm);
        s.f_start = f(astart);
        s.vfac_start = s.f_start * astart * H(astart) / h;
        eh.sqrtpnorm = s.sqrtpnorm;
        eh.n_s = p.n_s;
    }
};

// Gauss-Kronrod 7/15 nodes and weights
constexpr double kXgk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                            0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                            0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                            0.207784955007898467600689403773245, 0.0};
constexpr double kWgk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                            0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                            0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                            0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
constexpr double kWg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                           0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

double Calc::integrate(double a, double b, double rel) const {
    struct Seg { double a, b, r, e; };
    auto gk = [this](double lo, double hi) {
        const double c = 0.5 * (lo + hi), hl = 0.5 * (hi - lo);
        const double fc = sig8_integrand(c);
        double rk = fc * kWgk[7], rg = fc * kWg[3];
        for (int j = 0; j < 7; ++j) {
            const double dx = hl * kXgk[j];
            const double f1 = sig8_integrand(c - dx), f2 = sig8_integrand(c + dx);
            rk += kWgk[j] * (f1 + f2);
            if (j & 1) rg += kWg[j / 2] * (f1 + f2);
        }
        return Seg{lo, hi, rk * hl, std::fabs((rk - rg) * hl)};
    };
    std::vector<Seg> segs{gk(a, b)};
    double tot = segs[0].r, err = segs[0].e;
    while (err > rel * std::fabs(tot) && segs.size() < 20000) {
        size_t w = 0;
        for (size_t i = 1; i < segs.size(); ++i)
            if (segs[i].e > segs[w].e) w = i;
        const Seg s0 = segs[w];
        const double m = 0.5 * (s0.a + s0.b);
        segs[w] = gk(s0.a, m);
        segs.push_back(gk(m, s0.b));
        tot = err = 0.0;
        for (const Seg& s : segs) { tot += s.r; err += s.e; }
    }
    return tot;
}
}  // namespace

EhParams eh_setup(const mfic_cosmo_params& p) {
    EhParams e{};
    const double omhh = p.Omega_m * p.H0 * p.H0 * (0.01 * 0.01);
    const double fb = p.Omega_b / p.Omega_m;
    const double Tcmb = p.Tcmb <= 0.0 ? 2.728 : p.Tcmb;
    const double obhh = omhh * fb, th = Tcmb / 2.7, th4 = th * th * th * th;
    const double zeq = 2.50e4 * omhh / th4, keq = 0.0746 * omhh / (th * th);
    const double zd1 = 0.313 * std::pow(omhh, -0.419) * (1 + 0.607 * std::pow(omhh, 0.674));
    const double zd2 = 0.238 * std::pow(omhh, 0.223);
    const double zd = 1291 * std::pow(omhh, 0.251) / (1 + 0.659 * std::pow(omhh, 0.828)) * (1 + zd1 * std::pow(obhh, zd2));
    const double Rd = 31.5 * obhh / th4 * (1000 / (1 + zd)), Req = 31.5 * obhh / th4 * (1000 / zeq);
    const double s = 2. / 3. / keq * std::sqrt(6. / Req) *
                     std::log((std::sqrt(1 + Rd) + std::sqrt(Rd + Req)) / (1 + std::sqrt(Req)));
    const double ksilk = 1.6 * std::pow(obhh, 0.52) * std::pow(omhh, 0.73) * (1 + std::pow(10.4 * omhh, -0.95));
    const double a1 = std::pow(46.9 * omhh, 0.670) * (1 + std::pow(32.1 * omhh, -0.532));
    const double a2 = std::pow(12.0 * omhh, 0.424) * (1 + std::pow(45.0 * omhh, -0.582));
    const double ac = std::pow(a1, -fb) * std::pow(a2, -(fb * fb * fb));
    const double b1 = 0.944 / (1 + std::pow(458 * omhh, -0.708)), b2 = std::pow(0.395 * omhh, -0.0266);
    const double bc = 1.0 / (1 + b1 * (std::pow(1 - fb, b2) - 1));
    const double y = zeq / (1 + zd);
    const double G = y * (-6. * std::sqrt(1 + y) + (2. + 3. * y) * std::log((std::sqrt(1 + y) + 1) / (std::sqrt(1 + y) - 1)));
    e.keq = keq;
    e.s = s;
    e.ksilk = ksilk;
    e.ac = ac;
    e.bc = bc;
    e.ab = 2.07 * keq * s * std::pow(1 + Rd, -0.75) * G;
    e.bnode = 8.41 * std::pow(omhh, 0.435);
    e.bb = 0.5 + fb + (3. - 2. * fb) * std::sqrt(std::pow(17.2 * omhh, 2.0) + 1);
    e.h = p.H0 * 0.01;
    e.fb = p.Omega_b / p.Omega_m;
    e.fc = (p.Omega_m - p.Omega_b) / p.Omega_m;
    e.alpha_s = p.alpha_s;
    e.beta_s = p.beta_s;
    e.k_p = p.k_p;
    e.n_s = p.n_s;
    return e;
}

void cosmo_defaults(mfic_cosmo_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->H0 = 67.742; p->Omega_m = 0.3099; p->Omega_b = 0.048891054; p->Omega_k = 0.0;
    p->w_0 = -1.0; p->w_a = 0.0; p->n_s = 0.96822; p->A_s = 2.1064e-9; p->alpha_s = 0.0;
    p->beta_s = 0.0; p->k_p = 0.05; p->N_ur = 2.046; p->m_nu[0] = 0.06;
    p->deg_nu[0] = p->deg_nu[1] = p->deg_nu[2] = 1.0; p->Tcmb = 2.7255; p->O_nu_norm = 93.13861;
    p->sigma_8 = -1.0; p->zstart = 100.0; p->ztarget = 0.0;
}

std::vector<double> tk_spline_build(const double* k, const double* T, size_t n) {
    if (n < 6 || !k || !T) throw std::invalid_argument("a tabulated transfer function needs at least 6 rows");
    std::vector<double> x(n), y(n), h(n - 1), m(n, 0.0);
    for (size_t i = 0; i < n; ++i) {
        if (!(k[i] > 0.0) || !(T[i] > 0.0))
            throw std::invalid_argument("tabulated transfer function: k and T must be positive (row " +
                                        std::to_string(i) + ")");
        x[i] = std::log(k[i]);
        y[i] = std::log(T[i]);
        if (i > 0) {
            h[i - 1] = x[i] - x[i - 1];
            if (!(h[i - 1] > 0.0))
                throw std::invalid_argument("tabulated transfer function: k must increase (row " + std::to_string(i) +
                                            ")");
        }
    }
    // natural spline: m_0 = m_{n-1} = 0 and, for 0 < i < n-1,
    // h_{i-1} m_{i-1} + 2 (h_{i-1} + h_i) m_i + h_i m_{i+1} = 6 (s_i - s_{i-1}),
    // s_i the slope of interval i; tridiagonal elimination then back substitution
    std::vector<double> diag(n, 1.0), rhs(n, 0.0);
    for (size_t i = 1; i + 1 < n; ++i) {
        diag[i] = 2.0 * (h[i - 1] + h[i]);
        rhs[i] = 6.0 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1]);
        if (i > 1) {
            const double w = h[i - 1] / diag[i - 1];
            diag[i] -= w * h[i - 1];
            rhs[i] -= w * rhs[i - 1];
        }
    }
    for (size_t i = n - 2; i >= 1; --i) m[i] = (rhs[i] - (i + 2 < n ? h[i] * m[i + 1] : 0.0)) / diag[i];
    // per interval: ln T = a + b t + c t^2 + d t^3, t = ln k - ln k_i
    std::vector<double> nodes(5 * n, 0.0);
    for (size_t i = 0; i < n; ++i) {
        double* c = nodes.data() + 5 * i;
        c[0] = x[i];
        c[1] = y[i];
        if (i + 1 < n) {
            c[2] = (y[i + 1] - y[i]) / h[i] - h[i] * (2.0 * m[i] + m[i + 1]) / 6.0;
            c[3] = 0.5 * m[i];
            c[4] = (m[i + 1] - m[i]) / (6.0 * h[i]);
        }
    }
    return nodes;
}

void* tk_upload(const mfic_cosmo_params& p, EhParams* eh) {
    if (p.tk_n == 0) {
        eh->tk = nullptr;
        eh->tk_n = 0;
        return nullptr;
    }
    const std::vector<double> nodes = tk_spline_build(p.tk_k, p.tk_T, p.tk_n);
    void* d = nullptr;
    if (hipMalloc(&d, nodes.size() * sizeof(double)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    if (hipMemcpy(d, nodes.data(), nodes.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        throw std::runtime_error("hipMemcpy failed");
    }
    eh->tk = static_cast<const double*>(d);
    eh->tk_n = (int)p.tk_n;
    eh->tk_lkmin = std::log(p.tk_k[1]);
    return d;
}

void cosmo_compute(const mfic_cosmo_params& p, mfic_cosmo_scalars* s, EhParams* eh) {
    if (!(p.Omega_m > 0.0) || !(p.Omega_b > 0.0) || !(p.H0 > 0.0))
        throw std::invalid_argument("Omega_m, Omega_b and H0 must be positive");
    if (!(p.A_s > -1.0) && !(p.sigma_8 > 0.0))
        throw std::invalid_argument("sigma_8 normalisation requested without a positive sigma_8");
    Calc c;
    c.p = p;
    c.run();
    if (s) *s = c.s;
    if (eh) {
        *eh = c.eh;
        eh->tk = nullptr;  // host nodes die with c: the device copy comes from tk_upload
    }
}

void cosmo_eval(const mfic_cosmo_params& p, const std::string& what, const double* x, double* y, size_t n) {
    if (what == "T" && p.tk_n == 0) {
        // the E&H fit needs only its constants, not the growth tables and the
        // sigma_8 normalisation (a tabulated T(k) takes the full set-up below)
        const EhParams eh = eh_setup(p);
        for (size_t i = 0; i < n; ++i) y[i] = eh_transfer(eh, x[i]);
        return;
    }
    Calc c;
    c.p = p;
    c.run();
    for (size_t i = 0; i < n; ++i) {
        if (what == "T") y[i] = eh_transfer(c.eh, x[i]);
        else if (what == "D") y[i] = c.Draw(x[i]) / c.s.Dnow;
        else if (what == "f") y[i] = c.f(x[i]);
        else if (what == "H") y[i] = c.H(x[i]);
        else if (what == "A") y[i] = std::pow(x[i], 0.5 * p.n_s) * eh_transfer(c.eh, x[i]) * c.s.sqrtpnorm;
        else if (what == "vfact") y[i] = c.f(x[i]) * x[i] * c.H(x[i]) / c.s.h;
        else throw std::invalid_argument("unknown quantity " + what);
    }
}

}  // namespace mfic
