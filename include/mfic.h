/*
 * mfic -- MI355X-native monofonIC initial-condition hot path, C ABI.
 *
 * Every entry point returns 0 on success or a negative MFIC_ERR_* code;
 * mfic_last_error() returns a thread-local message for the last failure.
 * No C++ exception crosses this boundary.  All work is enqueued on the
 * context's HIP stream in call order; only download / sync / explicit
 * host-result calls block.  Handles are not thread-safe (as Grid_FFT).
 *
 * Reference citations are relative to the monofonIC source tree.
 *
 * Data layout (grid_fft.cc:40-112): a grid of logical size n0 x n1 x n2 is
 * one buffer of n0*n1*(n2+2) reals; in real space element (i,j,k) is at
 * (i*n1 + j)*(n2+2) + k, in k-space it is the complex array
 * [n0][n1][n2/2+1] (kx = i, ky = j, kz = k) over the same memory.
 * FFTs are unitary: both directions scale by 1/sqrt(n0 n1 n2).
 */
#ifndef MFIC_H
#define MFIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- constants --- */
enum mfic_status {
    MFIC_OK = 0,
    MFIC_ERR_ARG = -1,         /* invalid argument (null, mismatched shapes)    */
    MFIC_ERR_SIZE = -2,        /* unsupported grid size / odd padded size       */
    MFIC_ERR_HIP = -3,         /* HIP runtime failure                           */
    MFIC_ERR_ALLOC = -4,       /* device allocation failed                      */
    MFIC_ERR_STATE = -5,       /* wrong space / unallocated grid                */
    MFIC_ERR_CONFIG = -6,      /* invalid configuration value                  */
    MFIC_ERR_UNSUPPORTED = -7, /* feature outside the hot path                  */
    MFIC_ERR_IO = -8           /* file output failed (HDF5 missing or failing)  */
};
enum mfic_prec { MFIC_F32 = 0, MFIC_F64 = 1 };
/* same numbering as the reference's space_t (grid_fft.hh:29) */
enum mfic_space { MFIC_KSPACE = 0, MFIC_RSPACE = 1 };
/* output operators of operators.hh:25-35 */
enum mfic_outop { MFIC_OP_ASSIGN = 0, MFIC_OP_ADD = 1, MFIC_OP_SUB = 2, MFIC_OP_MADD = 3 };
/* operand kinds of the BaseConvolver builders (convolution.hh:74-275) */
enum mfic_opkind {
    MFIC_OPND_HESS = 0,      /* (i k_a)(i k_b) f            convolve_Hessians        */
    MFIC_OPND_SUM_HESS = 1,  /* ((i k_a)(i k_b)+(i k_c)(i k_d)) f  SumOfHessians    */
    MFIC_OPND_DIFF_HESS = 2, /* ((i k_a)(i k_b)-(i k_c)(i k_d)) f  DifferenceOf...  */
    MFIC_OPND_GRAD = 3,      /* (i k_a) f                   convolve_Gradients       */
    MFIC_OPND_IDENT = 4      /* f                                                    */
};
enum mfic_lattice { MFIC_LATTICE_SC = 0, MFIC_LATTICE_BCC = 1, MFIC_LATTICE_FCC = 2 };

typedef struct mfic_ctx_s *mfic_ctx;
typedef struct mfic_grid_s *mfic_grid;

/* ------------------------------------------------------------- context --- */
const char *mfic_last_error(void);
const char *mfic_version(void);
/* context = one HIP device + one stream + twiddle/plan caches.
 * Replaces the process-global FFTW state set up in main.cc:156-182.
 * Destruction order: every handle created on a context (grids, wrapped
 * grids, convolvers, PLT tables, runs, graphs, 2LPT plans) must be destroyed
 * before the context; mfic_ctx_destroy returns MFIC_ERR_STATE and leaves the
 * context intact while any of them is alive. */
int mfic_ctx_create(int device, mfic_ctx *out);
int mfic_ctx_destroy(mfic_ctx ctx);
int mfic_ctx_sync(mfic_ctx ctx);
/* raw hipStream_t of the context (for event timing / interop) */
int mfic_ctx_stream(mfic_ctx ctx, void **stream);
/* per-kernel timing with HIP events recorded on the context stream around
 * every launch (the reference's per-FFT debug timings, grid_fft.cc:220-227).
 * mfic_prof_report synchronises, writes a JSON object
 * {kernel: {count, ms_total, bytes_total}} (bytes = algorithmic bytes) and
 * clears the records. */
int mfic_prof_enable(mfic_ctx ctx, int on);
int mfic_prof_report(mfic_ctx ctx, char *buf, size_t len);

/* ---------------------------------------------------------------- grid --- */
/* Grid_FFT<real_t>(N, L, allocate=true, space) + allocate()
 * (grid_fft.hh:77-83, grid_fft.cc:40-112).  Storage is device memory. */
int mfic_grid_create(mfic_ctx ctx, const size_t n[3], const double L[3], int prec, int space,
                     mfic_grid *out);
/* Grid_FFT::reset / destructor (grid_fft.hh:95-101) */
int mfic_grid_destroy(mfic_grid g);
/* Grid_FFT::FourierTransformForward/Backward(false): flip the space flag only
 * (grid_fft.cc:207-270 with do_transform = false) */
int mfic_grid_set_space(mfic_grid g, int space);
int mfic_grid_get_space(mfic_grid g, int *space);
/* device pointer and element count (n0*n1*(n2+2) reals) */
int mfic_grid_device_ptr(mfic_grid g, void **ptr, size_t *nreal);
/* copy the whole buffer in the layout above; host_prec may differ from the
 * grid precision (converted on the device).  Synchronous. */
int mfic_grid_upload(mfic_grid g, const void *host, int host_prec);
int mfic_grid_download(mfic_grid g, void *host, int host_prec);
/* device-to-device copy, Grid_FFT::copy_from (grid_fft.hh:149-170) */
int mfic_grid_copy(mfic_grid dst, mfic_grid src);
/* Grid_FFT::zero (grid_fft.hh:140-147) */
int mfic_grid_zero(mfic_grid g);
/* operator*= (grid_fft.hh:400-412): every stored real multiplied by c */
int mfic_grid_scale(mfic_grid g, double c);
/* zero_DC_mode, k-space branch (grid_fft.hh:844-850) */
int mfic_grid_zero_dc(mfic_grid g);
/* apply_InverseLaplacian (grid_fft.hh:440-449): -f/|k|^2, then DC := 0 */
int mfic_grid_inverse_laplacian(mfic_grid g);
/* shift_field (grid_fft.hh:832-842): forward, * exp(i sum_a s_a k_a dx_a),
 * backward (when transform_back) */
int mfic_grid_shift(mfic_grid g, const double s[3], int transform_back);

/* ----------------------------------------------------------------- FFT --- */
/* FourierTransformForward / FourierTransformBackward (grid_fft.cc:207-270):
 * r2c / c2r with unitary scaling; a no-op when the grid is already in the
 * target space.  Supported sizes: every n = 2^a 3^b with 4 <= n <= 768 per
 * axis and n2 even (one-GPU hot-path sizes); others return MFIC_ERR_SIZE. */
int mfic_fft_forward(mfic_grid g);
int mfic_fft_backward(mfic_grid g);
int mfic_fft_size_supported(size_t n);
/* autotuner hook: override the tile (pencils per block for kind 0, every
 * strided pass of the length, and for kind 2, the bcc particle path's axis-0
 * pass that also forms the shifted sub-lattice; rows per block for kind 1
 * contiguous-axis passes; kinds 3..8, the padded convolver pass kinds of
 * mfic_fft_pass below, override kind 0 for that pass only) and the LDS pitch
 * padding (kinds 0 and 1) for one transform length and precision; tile <= 0
 * restores the built-in choice.  No reference counterpart (FFTW planning,
 * grid_fft.cc:53-54, plays this role there). */
int mfic_fft_set_tile(mfic_ctx ctx, int length, int kind, int prec, int tile, int lds_pad);
/* the override of mfic_fft_set_tile for a key (tile 0: the built-in choice) */
int mfic_fft_get_tile(mfic_ctx ctx, int length, int kind, int prec, int *tile, int *lds_pad);

/* Radix plans.  Each transform length has a compiled default plan (elements
 * per thread and the radix of each Stockham stage); the padded-pass lengths
 * 512 and 768 also have an alternative with twice the elements per thread
 * and 16 threads per transform.  A strided pass of one kind selects its plan
 * per length and precision (the row passes keep their own plans); the fused
 * source row passes select one of two forms the same way.  No reference
 * counterpart: FFTW picks its codelets while planning (grid_fft.cc:53-54). */
enum mfic_fft_pass {
    MFIC_PASS_PENCILS = 0,    /* strided passes of N-grid transforms (axes 0 and 1) */
    MFIC_PASS_AX0_SHIFT = 2,  /* axis-0 c2r pass that also forms the shifted bcc sub-lattice */
    MFIC_PASS_PAD_C2R_A = 3,  /* convolver padded passes (length 3N/2): axis-0 c2r */
    MFIC_PASS_PAD_C2R_B = 4,  /*   axis-1 c2r */
    MFIC_PASS_PAD_R2C_A = 5,  /*   axis-0 r2c */
    MFIC_PASS_PAD_R2C_B = 6,  /*   axis-1 r2c */
    MFIC_PASS_TRUNC_AX0 = 7,  /*   two-stage truncation filter, axis 0 */
    MFIC_PASS_TRUNC_AX1 = 8,  /*   two-stage truncation filter, axis 1 */
    MFIC_PASS_FUSED_ROWS = 9  /*   fused source row passes: variant 0 transforms two rows per
                                   full-length complex FFT, 1 one row per half-length FFT */
};
/* number of compiled radix plans for a length (0: unsupported length) */
int mfic_fft_plan_count(int length);
/* elements per thread and radices (0 after the last stage) of plan
 * `variant` of a length */
int mfic_fft_plan_info(int length, int variant, int *elems_per_thread, int radices[4]);
/* select plan `variant` for one pass kind, length and precision on a
 * context; variant < 0 restores the default.  mfic_fft_get_plan reports
 * the current choice. */
int mfic_fft_set_plan(mfic_ctx ctx, int length, int pass, int prec, int variant);
int mfic_fft_get_plan(mfic_ctx ctx, int length, int pass, int prec, int *variant);

/* --------------------------------------------------------------- noise --- */
/* RNG_ngenic constructor seed table (random_ngenic.cc:39-75), host only:
 * table has n*n entries. */
int mfic_ngenic_seedtable(long seed, size_t n, uint32_t *table);
/* RNG_ngenic::Fill_Grid (random_ngenic.cc:84-154), non-distributed branch,
 * on the device; leaves g in k-space.  RANLUX stream bit-exact with GSL's
 * ranlxd1; Gaussian transform with correctly rounded log/sin/cos. */
int mfic_ngenic_fill(mfic_grid g, long seed);

/* ----------------------------------------------------------- cosmology --- */
/* cosmology::parameters (cosmology_parameters.hh:80-241) inputs; fill with
 * mfic_cosmo_defaults() (ParameterSet = none, Planck2018EE+BAO+SN defaults,
 * cosmology_parameters.cc:110-133) and override.  A_s <= -1 selects sigma_8
 * normalisation, as a given sigma_8 key does (cosmology_parameters.hh:105). */
typedef struct {
    double Omega_m, Omega_b, H0, n_s, sigma_8, A_s, k_p;
    double Tcmb, N_ur, m_nu[3], deg_nu[3], O_nu_norm;
    double w_0, w_a, Omega_k, alpha_s, beta_s;
    int zero_radiation;
    double zstart, ztarget;
    /* Tabulated transfer function, the CAMB_file plugin
     * (transfer_CAMB_file.cc:36-160, 195-199): when tk_n > 0, T(k) is the
     * natural cubic spline of ln T over ln k through (tk_k[i], tk_T[i])
     * (k in h/Mpc, increasing; T > 0, the total-matter column), with k
     * clamped to >= tk_k[1] and the running factor taken at k h; the sigma_8
     * integral spans [tk_k[1], tk_k[tk_n-2]].  tk_n = 0 selects Eisenstein &
     * Hu.  Host arrays, read (and copied to the device) by each call that
     * takes these parameters; a grid whose largest |k| exceeds the table
     * returns MFIC_ERR_CONFIG where the reference's spline would fail. */
    const double *tk_k, *tk_T;
    size_t tk_n;
} mfic_cosmo_params;
/* cosmology::calculator scalars (cosmology_calculator.hh:167-223, 334-432) */
typedef struct {
    double h, Omega_nu_massive, Omega_gamma, Omega_r, Omega_c, Omega_DE;
    double Dnow, Dplus_start, Dplus_target, f_start, vfac_start;
    double sigma8_measured, pnorm, sqrtpnorm;
} mfic_cosmo_scalars;
void mfic_cosmo_defaults(mfic_cosmo_params *p);
int mfic_cosmo_compute(const mfic_cosmo_params *p, mfic_cosmo_scalars *out);
/* host evaluations, what = "D" D+(a), "f" f(a), "H" H(a), "vfact" f a H / h,
 * "T" T(k) and "A" k^{n_s/2} T(k) sqrt(pnorm) [k in h/Mpc], T from the table
 * when tk_n > 0, else E&H.  "T" with E&H costs only the fit; the others first
 * integrate the growth tables and the sigma_8 normalisation (milliseconds), so
 * pass many x per call. */
int mfic_cosmo_eval(const mfic_cosmo_params *p, const char *what, const double *x, double *y,
                    size_t count);

/* ----------------------------------------------------------------- phi1 --- */
/* noise normalisation (ic_generator.cc:378-383) + phi1 (ic_generator.cc:
 * 509-516): phi(k) = -wn(k) A(|k|) / |k|^2 / volfac, DC := 0, where
 * A = k^{n_s/2} T_EH(k) sqrt(pnorm) (cosmology_calculator.hh:429-432) is
 * evaluated in FP64 on the device.  wn and phi must be k-space grids of the
 * same size (phi may equal wn). */
int mfic_phi1_from_noise(mfic_grid phi, mfic_grid wn, const mfic_cosmo_params *p, int fixing,
                         int inversion);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mfic_phi1_from_noise with the same kernels and results,
 * but its per-|k| amplitude table lives in a fixed device array of the
 * library (442,369 doubles, 3.4 MiB per GPU, kept until the process ends)
 * instead of context scratch, so the call allocates nothing.  Limits: cubic
 * grids up to 768^3 (larger ones return MFIC_ERR_SIZE; non-cubic grids need
 * no table); Eisenstein & Hu only (tk_n > 0 returns MFIC_ERR_UNSUPPORTED,
 * since a tabulated T(k) needs its spline nodes copied to the device).  All
 * calls on one GPU share the array: calls that overlap on one GPU (different
 * streams or host threads) race; calls in one stream's order do not. */
int mfic_phi1_from_noise_noalloc(mfic_grid phi, mfic_grid wn, const mfic_cosmo_params *p, int fixing,
                                 int inversion);

/* ------------------------------------------------------------ convolver --- */
typedef struct {
    mfic_grid f;     /* k-space N-grid field                      */
    int kind;        /* mfic_opkind                               */
    int d[4];        /* directions (HESS: d0 d1; SUM/DIFF: d0..d3) */
} mfic_operand;
/* OrszagConvolver<real_t>(N, L) (convolution.hh:446-478): owns the two
 * padded 3N/2 grids and the N-grid buffer. */
typedef struct mfic_conv_s *mfic_conv;
int mfic_conv_create(mfic_ctx ctx, const size_t n[3], const double L[3], int prec, mfic_conv *out);
int mfic_conv_destroy(mfic_conv c);
/* OrszagConvolver::convolve2 (convolution.hh:495-518) with the operand
 * builders of convolution.hh:74-275 and op(out, v) of operators.hh:25-35:
 * out = / += / -= / += coef * trunc(a * b).  The output grid is left in
 * k-space. */
int mfic_conv_convolve2(mfic_conv c, const mfic_operand *a, const mfic_operand *b, mfic_grid out,
                        int op, double coef);
/* OrszagConvolver::convolve3 (convolution.hh:525-532): trunc(trunc(a b) c) */
int mfic_conv_convolve3(mfic_conv c, const mfic_operand *a, const mfic_operand *b,
                        const mfic_operand *cc, mfic_grid out, int op, double coef);

/* ------------------------------------------------------- LPT potentials --- */
/* The LPT source block of ic_generator::run (ic_generator.cc:523-629):
 * phi2 (2LPT), phi3 and A3x/y/z (3LPT) from phi1, each followed by the
 * inverse Laplacian.  Computed with the reorganised schedule: each Hessian
 * goes through the padded axis-0/axis-1 c2r passes once, one fused row pass
 * per order forms every product sharing a truncation (FP64 sums) in LDS and
 * returns r2c'd rows, the cubic terms keep their two-stage truncation.
 * Growth factors are NOT applied here (mfic_lpt_component applies them).
 * Grids not needed for the order may be NULL. */
int mfic_lpt_sources(mfic_ctx ctx, int order, mfic_grid phi1, mfic_grid phi2, mfic_grid phi3,
                     mfic_grid A3x, mfic_grid A3y, mfic_grid A3z);

/* LPT time coefficients (ic_generator.cc:258-273) */
typedef struct {
    double Dplus0, vfac, g1, g2, g3, g3c, vfac1, vfac2, vfac3, volfac;
} mfic_lpt_coef;
int mfic_lpt_coefficients(const mfic_cosmo_params *p, size_t n, double L, int order,
                          mfic_lpt_coef *out);

/* ------------------------------------------------------------ PLT tables --- */
/* particle::lattice_gradient (particle_plt.hh:520-549): tables on an
 * ng^3 mesh (64 in the reference) for the given lattice. */
typedef struct mfic_plt_s *mfic_plt;
int mfic_plt_create(mfic_ctx ctx, int lattice, int ng, mfic_plt *out);
int mfic_plt_destroy(mfic_plt p);
/* copy the 4 tables (D_r, D_theta, D_phi, vfac; ng*ng*(ng/2+1) doubles each) */
int mfic_plt_download(mfic_plt p, double *Dr, double *Dtheta, double *Dphi, double *vfac);

/* -------------------------------------------------------------- assembly --- */
/* One displacement (comp 0..2) or velocity (comp 3..5) component in k-space
 * (ic_generator.cc:931-970, 987-1047): gradient of the growth-scaled
 * potentials plus the A3 curl, PLT velocity correction, unit / L, DC := 0.
 * Growth factors g1..g3c and vfac1..3 are applied here in FP64 (the
 * reference applies them to the potentials at ic_generator.cc:644-658,
 * which is the same product).  plt may be NULL (Fourier gradient,
 * operators.hh:38-61).  out is left in k-space. */
int mfic_lpt_component(mfic_ctx ctx, int order, int comp, const mfic_lpt_coef *k, mfic_plt plt,
                       double unit, mfic_grid phi1, mfic_grid phi2, mfic_grid phi3, mfic_grid A3x,
                       mfic_grid A3y, mfic_grid A3z, mfic_grid out);

/* ----------------------------------------------------------------- readout --- */
/* lattice_generator::set_positions / set_velocities for Bravais sc / bcc /
 * fcc, lattice index 0 (particle_generator.hh:499-544, 597-639): `field` is the
 * real-space component; positions are ((ijk + shift)/N) * lunit + field in
 * FP64.  dst is device memory for 3*npart doubles (interleaved [ipart][3];
 * host memory is refused with MFIC_ERR_ARG); is_pos = 0 writes velocities
 * (lunit ignored).  Every further sub-lattice (bcc (1/2,1/2,1/2); fcc
 * (0,1/2,1/2), (1/2,0,1/2), (1/2,1/2,0)) is read after a Fourier shift by its
 * offset from the previous one (the field is left shifted, as the reference
 * leaves it). */
int mfic_lattice_readout(mfic_grid field, int lattice, int idim, int is_pos, double lunit,
                         double *dst_device);
/* the same for lattice index 0 or 1: index 1 (the secondary lattice of the
 * baryon species) first shifts the field by second_lattice_shift (sc and
 * fcc (1/2,1/2,1/2), bcc (1/2,1/2,0); particle_generator.hh:58-63, 512-531)
 * and offsets the positions by it */
int mfic_lattice_readout_index(mfic_grid field, int lattice, int lattice_index, int idim, int is_pos,
                               double lunit, double *dst);

/* --------------------------------------------------------------- pipeline --- */
typedef struct {
    size_t n;               /* GridRes                                       */
    double L;               /* BoxLength [Mpc/h]                             */
    int lpt_order;          /* 1..3 (LPTorder)                               */
    int lattice;            /* mfic_lattice (ParticleLoad sc / bcc / fcc)    */
    int plt;                /* PLT corrections (ENABLE_PLT)                  */
    int prec;               /* bulk precision MFIC_F32 / MFIC_F64            */
    int do_fixing, do_inversion;
    long seed;              /* [random] seed                                 */
    int particles;          /* 0: 6 component fields, 1: FP64 particles     */
    double lunit, vunit;    /* output plugin position_unit / velocity_unit   */
    int baryons;            /* DoBaryons: a second, baryon species on the
                             * secondary lattice (particle output, E&H)     */
} mfic_ic_config;

/* The whole hot path of ic_generator::run for one box (ic_generator.cc:
 * 246-1060): white noise, phi1, LPT sources, assembly, c2r, readout, all on
 * the device.  Results stay on the device in the run object. */
typedef struct mfic_run_s *mfic_run;
int mfic_run_create(mfic_ctx ctx, const mfic_ic_config *cfg, const mfic_cosmo_params *cosmo,
                    mfic_run *out);
int mfic_run_destroy(mfic_run r);
/* execute (enqueue) the whole pipeline; blocking = 0 returns immediately */
int mfic_run_execute(mfic_run r, int blocking);
/* component fields (6 grids, dx dy dz vx vy vz, real space) when particles=0 */
int mfic_run_field(mfic_run r, int comp, mfic_grid *out);
/* particle arrays (device pointers, 3*npart doubles each) when particles=1 */
int mfic_run_particles(mfic_run r, double **pos, double **vel, size_t *npart);
/* the baryon species of a run with baryons = 1 (ic_generator.cc:704-745):
 * the same LPT displacements and velocities (delta_bc = theta_bc = 0 with
 * the Eisenstein & Hu fit, so C_species has no effect) read out on the
 * secondary lattice, offset by second_lattice_shift (particle_generator.hh:
 * 58-63: sc and fcc (1/2,1/2,1/2), bcc (1/2,1/2,0)); every sub-lattice at that
 * offset plus its own, positions not wrapped.  Device pointers, [p][3] FP64,
 * npart as for the dark matter.  The dark matter takes Omega_m - Omega_b,
 * the baryons Omega_b (ic_generator.cc:157-166). */
int mfic_run_baryon_particles(mfic_run r, double **pos, double **vel, size_t *npart);
/* k-space potentials phi1, phi2, phi3, A3x, A3y, A3z (which = 0..5), before
 * growth scaling: the assembly applies g1..g3c (ic_generator.cc:644-658) */
int mfic_run_potential(mfic_run r, int which, mfic_grid *out);

/* -------------------------------------------------------------- batching --- */
/* No reference counterpart: monofonIC runs one box per process, and FFTW
 * planning (grid_fft.cc:53-54) plays the autotuner's role.
 *
 * Graph capture: every call enqueued on the context's stream between
 * mfic_graph_begin and mfic_graph_end (for example mfic_run_execute of many
 * runs of different sizes and seeds) is recorded into one HIP graph, replayed
 * by mfic_graph_launch.  Whatever the captured calls allocate lazily (scratch
 * buffers, twiddle tables) must already exist: execute the same sequence once
 * before capturing.  A call that would allocate fails with MFIC_ERR_STATE and
 * mfic_graph_end then reports the broken capture.  Per-kernel timing is
 * suspended while capturing. */
typedef struct mfic_graph_s *mfic_graph;
int mfic_graph_begin(mfic_ctx ctx);
int mfic_graph_end(mfic_ctx ctx, mfic_graph *out);
int mfic_graph_launch(mfic_graph g, int blocking);
int mfic_graph_destroy(mfic_graph g);
/* Autotuner: times whole pipeline executions of a temporary run while it
 * walks, one knob at a time, the radix plan of every strided pass kind that
 * has more than one (mfic_fft_set_plan), the pencils-per-block choices of the
 * N-grid strided passes, of the padded passes (all of a length, then each
 * padded pass kind) and of the shifted sub-lattice pass, and
 * the rows-per-block choices of the N-grid contiguous pass
 * (mfic_fft_set_tile), then the plans once more; the winners stay in the
 * context.  mfic_autotune tunes an n^3 box of precision prec and LPT order
 * (sc lattice, component fields); mfic_autotune_config tunes the given run
 * configuration and cosmology.  best_ms (may be NULL) receives the tuned
 * time of one execution. */
int mfic_autotune(mfic_ctx ctx, size_t n, int prec, int lpt_order, double *best_ms);
int mfic_autotune_config(mfic_ctx ctx, const mfic_ic_config *cfg, const mfic_cosmo_params *cosmo,
                         double *best_ms);

/* --------------------------------------------------- differentiable 2LPT --- */
/* No reference counterpart (monofonIC has no derivative of its IC map).
 * Forward map from a real-space white-noise grid delta to the 2LPT
 * displacement grids, with the NGenIC convention of empty Nyquist planes:
 *   w = P_N r2c(delta); phi1 = -w A(k) / (k^2 volfac) (ic_generator.cc:378-383,
 *   509-516); phi2 = lap^-1 of the Orszag 2LPT source (ic_generator.cc:
 *   523-555); Psi_d = c2r(i k_d (g1 phi1 + g2 phi2)) * unit, DC := 0
 *   (ic_generator.cc:931-970); unit = 1 gives Mpc/h.
 * mfic_lpt2_adjoint maps cotangents of Psi_x, Psi_y, Psi_z back to the
 * gradient with respect to delta (reverse mode, hand-written), recomputing
 * phi1 from delta instead of storing forward intermediates.  All grids are
 * real-space n^3 grids of the plan's precision. */
typedef struct mfic_lpt2_s *mfic_lpt2;
int mfic_lpt2_create(mfic_ctx ctx, size_t n, double L, int prec, const mfic_cosmo_params *cosmo, double unit,
                     mfic_lpt2 *out);
int mfic_lpt2_destroy(mfic_lpt2 p);
int mfic_lpt2_forward(mfic_lpt2 p, mfic_grid noise, mfic_grid psi_x, mfic_grid psi_y, mfic_grid psi_z);
int mfic_lpt2_adjoint(mfic_lpt2 p, mfic_grid noise, mfic_grid cot_x, mfic_grid cot_y, mfic_grid cot_z,
                      mfic_grid grad);
/* a grid over caller-owned device memory in the layout above (for example a
 * framework tensor of shape [n0][n1][n2+2]); destroying it frees nothing */
int mfic_grid_wrap(mfic_ctx ctx, const size_t n[3], const double L[3], int prec, int space, void *device_ptr,
                   mfic_grid *out);

/* --------------------------------------------- complex grids and PPT --- */
/* Grid_FFT<ccomplex_t> transforms (grid_fft.cc:56-63, 207-270) on a device
 * array [n0][n1][n2] of complex values (FP32 or FP64 pairs): in place,
 * unitary (1/sqrt(n0 n1 n2) in both directions), forward e^{-ikx}; every n
 * = 2^a 3^b with 4 <= n <= 768. */
int mfic_fft_c2c(mfic_ctx ctx, void *data, const size_t n[3], int prec, int forward);
/* Semiclassical perturbation theory fields of the field_eulerian output
 * (ic_generator.cc:786-897) from an executed run's potentials: hbar =
 * 1.01 max|d phi| / (pi D+) over neighbouring cells (the reference's loop,
 * which leaves out the last x-plane), psi = exp(i (phi + phi2) / (hbar D+))
 * (phi alone for LPTorder 1), one drift exp(-i hbar k^2 D+ / 2) through the
 * c2c transforms, the kick exp(i phi2 / (hbar D+)), delta = |psi|^2 - 1 and
 * v_d = D+ vfac / L * velocity_unit * hbar Im(psi* d_d psi) / (1 + delta)
 * with d_d = i k_d in k-space (get_k, Nyquist at +n/2).  The run's species
 * is dark matter without baryons (C_species = 0, so sqrt(rho) = 1).  Outputs
 * are real-space grids of the run's size and precision; hbar may be NULL. */
int mfic_ppt_fields(mfic_run r, double velocity_unit, mfic_grid delta, mfic_grid vx, mfic_grid vy, mfic_grid vz,
                    double *hbar);

/* ------------------------------------------------------- particle output --- */
/* The Gadget-HDF5 output plugin (output_gadget_hdf5.cc:36-318, the
 * gadget_hdf5 and gadget_hdf5_double plugins; [output] keys in the options)
 * with HDF_IO.hh's datasets and attributes:
 *   /PartType1/Coordinates, /PartType1/Velocities  N x 3, 32-bit floats (the
 *     reference's vector writer, HDF_IO.hh:850-885, creates them as
 *     H5T_NATIVE_FLOAT whatever the plugin's precision);
 *   /PartType1/ParticleIDs  uint32, or uint64 with long_ids; particle p of an
 *     n^3 lattice with 2^lattice sub-lattices gets 2^lattice * (p mod n^3) +
 *     p / n^3 (particle_generator.hh:229-251);
 *   /Header (output_gadget_hdf5.cc:175-210) and /ICs_parameters.
 * The HDF5 C library is loaded at run time (MFIC_HDF5_LIB, libhdf5.so, or
 * the conda copy); without it the calls return MFIC_ERR_UNSUPPORTED. */
typedef struct {
    int double_precision; /* gadget_hdf5_double (has_64bit_reals)                  */
    int long_ids;         /* [output] UseLongids                                   */
    int gadget2_compat;   /* [output] Gadget2Compatibility: 32-bit NumPart arrays  */
} mfic_gadget_opts;
/* output_plugin::position_unit / velocity_unit / mass_unit of the writer
 * (output_gadget_hdf5.cc:79-85): BoxLength, BoxLength / sqrt(a_start) and
 * rho_crit BoxLength^3 [1e10 Msun/h]; a run written with mfic_gadget_write
 * must use the first two as lunit and vunit */
int mfic_gadget_units(double boxlength, double zstart, double *lunit, double *vunit, double *munit);
/* write_particle_data + the header of a run with particles = 1 (sc, bcc, fcc),
 * streamed chunk by chunk from its device arrays (the run must have
 * executed); cosmological header values come from the run's parameters.  A
 * run with baryons adds /PartType0 (IDs after the dark matter's) and, as the
 * plugin does for individually massive particles, a Masses dataset per
 * species (32-bit floats, 64-bit with double_precision) and a zero MassTable. */
int mfic_gadget_write(mfic_run r, const char *fname, const mfic_gadget_opts *opts);
/* the same file from host arrays: FP64 [p][3] positions and velocities of
 * the 2^lattice * n^3 particles of cfg (cfg->lunit / vunit as above) */
int mfic_gadget_write_host(const char *fname, const mfic_ic_config *cfg, const mfic_cosmo_params *cosmo,
                           const mfic_gadget_opts *opts, const double *pos, const double *vel);

/* The SWIFT output plugin (output_swift.cc:36-377, registered for FP64):
 * /Units, /ICs_parameters, /PartType1 with Coordinates and Velocities
 * [N][3] FP64, ParticleIDs (uint64 with UseLongids, the plugin's default,
 * else uint32) and Masses [N] FP64, every dataset chunked with a Fletcher-32
 * checksum (HDF_IO.hh:888-963; chunks are cut to the dataset length), and
 * /Header as the plugin's destructor writes it.  A run with baryons adds
 * /PartType0 with Masses, InternalEnergy (the gas temperature of
 * output_swift.cc:86-104 from Tcmb, Omega_b, h, YHe and gamma) and
 * SmoothingLength (BoxLength / h / GridRes). */
typedef struct {
    int long_ids; /* [output] UseLongids (the plugin defaults it to 1)            */
    double YHe;   /* [cosmology] YHe for the gas internal energy (<= 0: 0.245421) */
    double gamma; /* [cosmology] gamma, adiabatic index (<= 0: 5/3)              */
} mfic_swift_opts;
/* position_unit / velocity_unit / mass_unit of the SWIFT plugin
 * (output_swift.cc:63-69): BoxLength / h [Mpc], BoxLength [km/s],
 * 27.7536609198 BoxLength^3 / h [1e10 Msun]; a run written with
 * mfic_swift_write must use the first two as lunit and vunit */
int mfic_swift_units(double boxlength, const mfic_cosmo_params *cosmo, double *lunit, double *vunit,
                     double *munit);
int mfic_swift_write(mfic_run r, const char *fname, const mfic_swift_opts *opts);
int mfic_swift_write_host(const char *fname, const mfic_ic_config *cfg, const mfic_cosmo_params *cosmo,
                          const mfic_swift_opts *opts, const double *pos, const double *vel);

#ifdef __cplusplus
}
#endif
#endif /* MFIC_H */
