/*
 * ihg.h -- C ABI of libihg, the MI355X (gfx950) HIP implementation of
 * InstaHeat's right-hand side and time-stepping hot path.
 *
 * Plain C: device pointers are `double *` into HBM (from ihg_malloc), sizes
 * are counts, every call returns 0 or a negative IHG_E* code (message from
 * ihg_last_error()).  The library never exits the process.  All work of a
 * context is queued on that context's own HIP stream; calls that hand data
 * to the host synchronise that stream.  A context is not thread safe (as the
 * reference's globals are not); separate contexts are independent.
 *
 * State layout (src/main_template.h:223-233, src/setup.c:249-271):
 *   y = [phi(N) | dphi(N) | psi(N) | dpsi(N) | h1(Next) | h2(Next) |
 *        dh1(Next) | dh2(Next) | a(1)],  N = Nx*Ny*Nz, Next = 2*Nx*Ny*(Nz/2+1)
 * real fields row-major with z fastest, h fields as FFTW r2c half spectra
 * with interleaved re/im doubles.
 */
#ifndef IHG_H
#define IHG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IHG_OK 0
#define IHG_EINVAL (-22)  /* bad argument / unsupported size */
#define IHG_ENOMEM (-12)  /* device allocation failed */
#define IHG_EHIP (-1000)  /* a HIP runtime call failed */

/* evo_flags (src/toolbox.h:16-24) as a bit set */
#define IHG_FILTER 1u
#define IHG_OUTPUT 2u

/* Runtime form of the compile-time parameters of src/main_template.h:116-162
 * that the hot path reads.  Grid sizes: powers of two, 4..512 per axis. */
typedef struct ihg_params {
    int nx, ny, nz;        /* GRIDPOINTS_X/Y/Z */
    double xa, xb;         /* SPATIAL_LOWER/UPPER_BOUND_X */
    double ya, yb;         /* SPATIAL_LOWER/UPPER_BOUND_Y */
    double za, zb;         /* SPATIAL_LOWER/UPPER_BOUND_Z */
    double mass;           /* MASS */
    int enable_gw;         /* ENABLE_GW */
    int enable_filter;     /* ENABLE_FFT_FILTER */
    int bins;              /* POWER_SPECTRUM_BINS */
    int gw_fp32;           /* precision form, no counterpart in the reference:
                              1 = GW source (products, FFTs, TT projection)
                              and GW spectrum in FP32; the state, h update and
                              everything else stay FP64 */
    int log_bins;          /* no counterpart in the reference (linear bins):
                              1 = the POWER_SPECTRUM_BINS bins of every
                              spectrum of equal width in log |k|, from the
                              fundamental |k| to k_max */
} ihg_params;

typedef struct ihg_ctx ihg_ctx;

/* Set-up and tear-down of the global state of src/setup.c:89-108
 * (sizes, k grid, filter mask, FFT tables, work arrays). */
int ihg_create(const ihg_params *p, ihg_ctx **out);
int ihg_destroy(ihg_ctx *c);
/* out = {N, M, Next, Ntot} (src/setup.c:249-271) */
int ihg_sizes(const ihg_ctx *c, size_t out[4]);
const char *ihg_last_error(void);

/* Device memory and transfers on the context's stream (replacing
 * fftw_malloc of the state arrays, src/setup.c:354-477).  ihg_malloc
 * returns zeroed memory; ihg_free takes only this context's arrays, and
 * ihg_destroy frees every array of the context still allocated. */
int ihg_malloc(ihg_ctx *c, double **ptr, size_t count);
int ihg_free(ihg_ctx *c, double *ptr);
int ihg_upload(ihg_ctx *c, double *dst, const double *src, size_t count);
int ihg_download(ihg_ctx *c, double *dst, const double *src, size_t count);
int ihg_copy(ihg_ctx *c, double *dst, const double *src, size_t count);
int ihg_zero(ihg_ctx *c, double *dst, size_t count);
/* No counterpart in the reference: dst = src by a plain copy kernel on the
 * context's stream (16-byte aligned arrays), timed as "stream_copy" by the
 * live profiler; used for state snapshots and as the HBM ceiling the pass
 * kernels are compared with. */
int ihg_stream_copy(ihg_ctx *c, double *dst, const double *src, size_t count);
int ihg_sync(ihg_ctx *c);

/* fft / ifft (src/toolbox.h:32-33, src/toolbox.c:806-827): 3D r2c and c2r,
 * unnormalised, FFTW sign and layout; `out` of r2c and `in` of c2r are
 * M complex values (2M doubles).  Out of place; c2r keeps its input. */
int ihg_fft_r2c(ihg_ctx *c, const double *in, double *out);
int ihg_fft_c2r(ihg_ctx *c, const double *in, double *out);

/* The array mk_summary reads (the reference's global `field`,
 * src/toolbox.c:638-647); defaults to the `f` of each call. */
int ihg_set_field(ihg_ctx *c, const double *field);

/* mk_rhs (src/toolbox.h:28, src/toolbox.c:75-88): result = dy/dt at f.
 * With IHG_FILTER (and enable_filter) phi, dphi, psi, dpsi in f are replaced
 * by their filtered versions, as in the reference.  With IHG_OUTPUT the
 * observables of output_all are computed (read with ihg_observables). */
int ihg_mk_rhs(ihg_ctx *c, double t, double *f, double *result,
               unsigned flags);
/* rho_mean, pressure_mean of the last mk_rhs (synchronises) */
int ihg_means(ihg_ctx *c, double out[2]);
/* observables of the last IHG_OUTPUT call, packed as
 * smry[6][4] (phi, dphi, psi, dpsi, rho, pressure: mean, var, min, max),
 * cstr[4], phi_ps[bins], psi_ps[bins], rho_ps[bins], gw[bins] */
int ihg_observables(ihg_ctx *c, double *out, size_t len);
/* mon.calls_rhs (src/toolbox.c:77) */
unsigned long ihg_calls_rhs(const ihg_ctx *c);

/* mean (src/toolbox.c:726-734) and variance (:744-754, given the mean:
 * (sum d^2 - (sum d)^2 / n) / (n - 1), d = x - mean) of n doubles of device
 * memory, with deterministic reductions.  Synchronise. */
int ihg_mean(ihg_ctx *c, const double *x, size_t n, double *out);
int ihg_variance(ihg_ctx *c, double mean, const double *x, size_t n,
                 double *out);
/* out[0] = sum (a - b)^2, out[1] = sum (cc - d)^2 over n doubles of device
 * memory: the two sums of check_for_stiffness (src/dopri853.c:550-560).
 * Synchronises. */
int ihg_diff_norms(ihg_ctx *c, const double *a, const double *b,
                   const double *cc, const double *d, size_t n,
                   double out[2]);

/* mk_initial_psi (src/setup.c:825-903): psi, dpsi from phi, dphi, a */
int ihg_mk_initial_psi(ihg_ctx *c, double ti, double *field, double *dfield);

/* RK4 stage arithmetic of run_rk4 (src/rk4.c:78-102):
 *   ihg_rk4_stage: out = y + dt * k / 2 (half = 1) or y + dt * k (half = 0)
 *   ihg_rk4_final: y += dt * (k1 + 2 k2 + 2 k3 + k4) / 6 */
int ihg_rk4_stage(ihg_ctx *c, double *out, const double *y, const double *k,
                  double dt, int half);
int ihg_rk4_final(ihg_ctx *c, double *y, const double *k1, const double *k2,
                  const double *k3, const double *k4, double dt);

/* DOPRI853 stage arithmetic of try_step (src/dopri853.c:331-458).
 * k[0] = dfield (FSAL), k[s-1] = RHS of stage s (s = 1..12).
 *   ihg_dp853_stage: out = y + dt * sum_j a_{s,j} k[j]  (s = 2..12)
 *   ihg_dp853_finish: y_new = y + dt * sum_j b_j k[j] and the two sums of
 *     error() (src/dopri853.c:469-486): sums[0] = sum (yerr2/sk)^2 (5th
 *     order), sums[1] = sum (yerr/sk)^2 (3rd order).  Synchronises. */
int ihg_dp853_stage(ihg_ctx *c, int s, double dt, const double *y,
                    const double *const *k, double *out);
int ihg_dp853_finish(ihg_ctx *c, double dt, const double *y,
                     const double *const *k, double *y_new, double rtol,
                     double atol, double sums[2]);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ihg_dp853_finish without a context.  The state is
 * given by its sizes: nlive = the live prefix the stage kernels stream
 * (Ntot - 1 with the GW blocks, 4 N without them; even), ntot = Ntot (the
 * scale factor is y[ntot - 1]).  The block partials and the two sums live
 * in a fixed device array of the library (67586 doubles, 528 KiB, one per
 * GPU, held until the process ends); nothing is allocated or freed.  Up to
 * 8.4 M element pairs (nlive <= 16777216) its work-groups and the order of
 * its sums are ihg_dp853_finish's, so the results are the same bit for
 * bit; beyond, its 32768 work-groups stride, and the sums agree to
 * rounding.  Work
 * is queued on `stream` (a hipStream_t of the current device, NULL: the
 * null stream), which is synchronised before the sums are returned.  Every
 * call on one GPU shares that GPU's array, so calls that overlap on one GPU
 * race.  nlive above IHG_NOALLOC_MAX_LIVE is rejected with IHG_EINVAL. */
#define IHG_NOALLOC_MAX_LIVE 2147483646L
int ihg_dp853_finish_noalloc(void *stream, long nlive, long ntot, double dt,
                             const double *y, const double *const *k,
                             double *y_new, double rtol, double atol,
                             double sums[2]);
/* the tableau the stage kernels use (row-major a[13][12], c[13], b[12],
 * bhh[3], er[12]) */
int ihg_dp853_tableau(double *a, double *c, double *b, double *bhh,
                      double *er);

/* DOPRI853 dense output (the continuous extension whose coefficients sit
 * unused in src/dopri853_constants.c:23-25, 98-173).  After an accepted step
 * y -> y_new of size dt, k[0..15]: k[0] the step's FSAL derivative, k[j]
 * stage j+1 (j = 1..11), k[12] = f(t + dt, y_new), k[13..15] the three extra
 * stages.  ihg_dp853_dense_stage (s = 14..16): out = y + dt * sum_j
 * a_{s,j} k[j], the input of extra stage s at t + c_s dt.
 * ihg_dp853_dense: out = the interpolant at t + theta dt, theta in [0, 1]
 * (k[1..4] are not read).  ihg_dp853_dense_tableau: c_extra[3],
 * a_extra[3][16], d[4][16], row-major. */
int ihg_dp853_dense_stage(ihg_ctx *c, int s, double dt, const double *y,
                          const double *const *k, double *out);
int ihg_dp853_dense(ihg_ctx *c, double theta, double dt, const double *y,
                    const double *y_new, const double *const *k, double *out);
int ihg_dp853_dense_tableau(double *c_extra, double *a_extra, double *d);

/* Stage arithmetic of the GSL steppers of src/gsl_stepper.c:49-58
 * (gsl_odeiv2_step_rkf45 / _rkck / _rk8pd; method = INTEGRATION_METHOD
 * 1 RKF45, 2 RKCK, 3 DOPRI89).  k[0] = f(t, y), k[s-1] = RHS of stage s.
 *   ihg_gsl_stage: out = y + dt * sum_j a_{s,j} k[j]   (s = 2..stages)
 *   ihg_gsl_finish: y_new = y + dt * sum_j b_j k[j], the embedded error
 *     yerr = dt * sum_j e_j k[j] (RKF45, RKCK) or dt * (sum_j bhat_j k[j] -
 *     sum_j b_j k[j]) (DOPRI89), and the standard control's ratio
 *     *rmax = max(DBL_MIN, max_i |yerr_i| / (eps_rel |y_new_i| + eps_abs))
 *     (control y_new: a_y = 1, a_dydt = 0).  Synchronises.
 *   ihg_gsl_tableau: info = {stages, order, err_mode}, c[13],
 *     a[13][12], b[13], w[13] (w = e for err_mode 0, bhat for 1). */
int ihg_gsl_stage(ihg_ctx *c, int method, int s, double dt, const double *y,
                  const double *const *k, double *out);
int ihg_gsl_finish(ihg_ctx *c, int method, double dt, const double *y,
                   const double *const *k, double *y_new, double eps_abs,
                   double eps_rel, double *rmax);
int ihg_gsl_tableau(int method, int info[3], double *c, double *a, double *b,
                    double *w);

/* Live timing of the kernel families (the runtime_* monitors of
 * src/main_template.h:292-311, per kernel): when enabled, every launch is
 * bracketed by HIP events on the context's stream.  ihg_prof_read returns,
 * per family, the summed milliseconds, the summed algorithmic bytes (each
 * input and output array counted once) and the number of launches since the
 * last reset; it synchronises. */
int ihg_prof_enable(ihg_ctx *c, int on);
int ihg_prof_reset(ihg_ctx *c);
int ihg_prof_read(ihg_ctx *c, double *ms, double *bytes, long *count,
                  int nfam);
int ihg_prof_nfam(void);
const char *ihg_prof_family(int fam);

/* ---- ensembles (the batched form; no counterpart in the reference) ------
 * An ensemble context holds nreal realizations of one grid: their states lie
 * back to back in one device array, `stride` doubles apart (Ntot rounded up
 * to even), and every kernel of ihg_mk_rhs_ensemble runs all of them in one
 * launch (the realization is the third grid dimension).  ihg_rk4_stage /
 * ihg_rk4_final on such a context step all nreal * stride values.  Calls on
 * one realization's slice (ihg_mk_rhs, ihg_observables, ...) use the work
 * arrays of realization 0. */
int ihg_create_ensemble(const ihg_params *p, int nreal, ihg_ctx **out);
/* out = {nreal, stride} */
int ihg_ensemble(const ihg_ctx *c, long out[2]);
/* mk_rhs of every realization: f, result are nreal * stride doubles;
 * flags: IHG_FILTER only */
int ihg_mk_rhs_ensemble(ihg_ctx *c, double t, double *f, double *result,
                        unsigned flags);

/* RK4 folded into mk_rhs (no counterpart in the reference, whose run_rk4
 * sweeps the state once more per stage, src/rk4.c:78-102): the kernels that
 * write the slope k = result also write, for the same elements,
 *   IHG_RK4_STAGE: out = y + dt k / 2 (half = 1) or y + dt k (half = 0),
 *                  the next stage's input (out differs from f, result, y);
 *   IHG_RK4_FINAL: y = y + dt (k1 + 2 k2 + 2 k3 + k) / 6 in place (y differs
 *                  from f and result); the slope k itself is not stored.
 * The values equal ihg_mk_rhs_ensemble followed by ihg_rk4_stage /
 * ihg_rk4_final bit for bit, with one sweep of the state fewer per stage
 * (on contexts without the fused kernels -- the FP32 GW source, or GW
 * with y lines outside 16 to 128 points -- it runs exactly those calls).
 * Arrays span all realizations of an ensemble context, as there.  flags:
 * IHG_FILTER only. */
#define IHG_RK4_STAGE 1
#define IHG_RK4_FINAL 2
typedef struct ihg_rk4_ep {
    int mode;                    /* IHG_RK4_STAGE or IHG_RK4_FINAL */
    int half;                    /* STAGE: dt / 2 (1) or dt (0) */
    double dt;
    const double *y;             /* the step's start state */
    double *out;                 /* STAGE: the next stage's input */
    const double *k1, *k2, *k3;  /* FINAL: the step's first three slopes */
} ihg_rk4_ep;
int ihg_mk_rhs_rk4(ihg_ctx *c, double t, double *f, double *result,
                   unsigned flags, const ihg_rk4_ep *ep);

/* ---- groups: ensembles of several grid sizes in one launch per pass -----
 * (no counterpart in the reference).  A group holds 1 to 4 ensemble (or
 * single) contexts of cubic grids of 16, 32, 64 or 128 points, FP64.
 * ihg_group_mk_rhs does what ihg_mk_rhs_ensemble does on every context, and
 * ihg_group_rk4_stage / ihg_group_rk4_final what ihg_rk4_stage /
 * ihg_rk4_final do, but each pass of all contexts is ONE launch: its work-
 * groups come from a table of segments, one per context, each running the
 * per-context kernel's body on its own grid, so the values are those of the
 * per-context calls bit for bit.  Array arguments hold one pointer per
 * context, in the order of ihg_group_create.  Launches go to the FIRST
 * context's stream (capture them with ihg_graph_begin on that context); the
 * other contexts' streams must be idle (ihg_sync) before grouped calls and
 * are not ordered after them.  The strided passes use the default line
 * plans (autotuned choices of the contexts are not applied), and the live
 * profiler does not time grouped launches. */
typedef struct ihg_group ihg_group;
int ihg_group_create(ihg_ctx *const *ctxs, int n, ihg_group **out);
int ihg_group_destroy(ihg_group *g);
/* grouped launches made so far */
long ihg_group_launches(const ihg_group *g);
/* flags: IHG_FILTER only (applied to the contexts with the filter on) */
int ihg_group_mk_rhs(ihg_group *g, double t, double *const *f,
                     double *const *result, unsigned flags);
/* ihg_mk_rhs_rk4 on every context: ep holds one epilogue per context, all
 * of one mode */
int ihg_group_mk_rhs_rk4(ihg_group *g, double t, double *const *f,
                         double *const *result, unsigned flags,
                         const ihg_rk4_ep *ep);
int ihg_group_rk4_stage(ihg_group *g, double *const *out, double *const *y,
                        double *const *k, double dt, int half);
int ihg_group_rk4_final(ihg_group *g, double *const *y, double *const *k1,
                        double *const *k2, double *const *k3,
                        double *const *k4, double dt);

/* HIP graphs of a context's stream: capture a sequence of library calls
 * that neither synchronises nor allocates (with the live profiler off),
 * replay it as one launch; ihg_graph_join puts independent graphs (e.g. one
 * per ensemble context) side by side in one graph.  ihg_graph_launch runs on
 * the given context's stream only: the other joined contexts' streams must
 * be idle (ihg_sync) before the launch and are not ordered after it. */
typedef struct ihg_graph ihg_graph;
int ihg_graph_begin(ihg_ctx *c);
int ihg_graph_end(ihg_ctx *c, ihg_graph **g);
int ihg_graph_join(ihg_graph *const *parts, int n, ihg_graph **g);
int ihg_graph_launch(ihg_graph *g, ihg_ctx *c);
int ihg_graph_destroy(ihg_graph *g);

/* Seeded Gaussian initial data, generated on the device (the reference
 * generates its initial data on the host, src/setup.c:771-790): field (Ntot
 * doubles) = phi0 + the c2r of a half spectrum of complex Gaussians of
 * variance amplitude^2 / (2 a0^3 omega_k), omega_k = sqrt(k^2 + m^2), for
 * 0 < |k| <= k_cut; dphi0; psi, dpsi and h zero; a0 last.  Each mode's
 * numbers depend on (seed, mode index) only (splitmix64 counters, Box-Muller
 * with a uniform phase).  Uses the context's spectral work arrays. */
int ihg_init_gaussian(ihg_ctx *c, double *field, uint64_t seed,
                      double amplitude, double phi0, double dphi0, double a0,
                      double k_cut);

/* Autotuner of the strided y and x passes (no counterpart in the
 * reference): every candidate line plan, work-group size and tile order for
 * the y and the x line length is timed (three rounds of `reps` passes on the
 * context's work arrays, in the RHS's shape: four arrays, every realization
 * of an ensemble; each candidate's fastest round counts), and a candidate
 * replaces the default only when it is at least 3 % faster.  choice = {y, x}
 * candidate indices in use afterwards; ms (may be NULL) receives up to nms
 * average pass times, y candidates first.  Candidates differ in rounding
 * only. */
int ihg_autotune(ihg_ctx *c, int reps, int choice[2], float *ms, int nms);
/* Number of strided-pass candidates {y, x} and the ones in use (either
 * pointer may be NULL); ihg_set_fft_candidates picks them by hand (an index
 * out of range is an error). */
int ihg_fft_candidates(ihg_ctx *c, int count[2], int current[2]);
int ihg_set_fft_candidates(ihg_ctx *c, int y, int x);

/* Switch the GW precision form (ihg_params.gw_fp32) for later calls. */
int ihg_set_gw_fp32(ihg_ctx *c, int on);

/* ---- learned DOPRI853 step proposer ------------------------------------
 * No counterpart in the reference (its controller is success(),
 * src/dopri853.c:505-540); for hosts that consult a proposer before
 * success() on accepted steps.  The embedded error estimate still judges
 * every step. */

/* State features of f: out = {a, variance of phi (normalised by N - 1),
 * mean of psi^2}.  Synchronises. */
int ihg_features(ihg_ctx *c, const double *f, double out[3]);

/* A 20-64-64-1 tanh MLP in FP32: inference, gradient and Adam training on
 * the GPU.  Parameters are one float array of ihg_mlp_nparams() values:
 * W1[64][20] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3.  x holds n rows
 * of 20 features, y / t n outputs / targets; host arrays.  The loss is the
 * mean of (y - t)^2.  An ihg_mlp has its own HIP stream; calls synchronise
 * it.  ihg_mlp_create takes initial parameters, or NULL for Glorot-uniform
 * weights from `seed`. */
#define IHG_MLP_IN 20
#define IHG_MLP_HIDDEN 64
typedef struct ihg_mlp ihg_mlp;
int ihg_mlp_nparams(void);
int ihg_mlp_create(ihg_mlp **out, const float *params, unsigned long seed);
int ihg_mlp_destroy(ihg_mlp *m);
int ihg_mlp_get_params(ihg_mlp *m, float *params);
int ihg_mlp_set_params(ihg_mlp *m, const float *params);
int ihg_mlp_forward(ihg_mlp *m, const float *x, float *y, int n);
/* gradient of the loss over all n samples (deterministic) */
int ihg_mlp_grad(ihg_mlp *m, const float *x, const float *t, int n,
                 float *grad, float *loss);
/* `epochs` passes of Adam (lr, beta 0.9 / 0.999) over shuffled mini-batches
 * of `batch` samples; epoch_loss[e] = mean loss over epoch e (may be NULL) */
int ihg_mlp_train(ihg_mlp *m, const float *x, const float *t, int n,
                  int epochs, int batch, float lr, unsigned long seed,
                  float *epoch_loss);

/* No counterpart in the reference, whose save() copies every stride-th
 * point on the host (LARGE_OUTPUT, src/io.c:560-598): the same sub-grid of
 * a real field (N doubles, row-major, z fastest) gathered on the device
 * and copied to the host, so that only the output points cross PCIe.
 * src = a device array of N doubles, or NULL for rho of the last mk_rhs;
 * out holds ceil(Nx/sx) ceil(Ny/sy) ceil(Nz/sz) doubles in the same order.
 * The first strided call allocates a staging array (kept by the context).
 * Synchronises. */
int ihg_download_strided(ihg_ctx *c, const double *src, const int stride[3],
                         double *out);

/* Debug access to intermediate arrays of the last mk_rhs, copied to host:
 * 0..2 d_x/y/z phi (GW on; FP32 values with gw_fp32), 3 rho, 4 pressure,
 * 5 tmp.f (output calls) */
int ihg_aux(ihg_ctx *c, int which, double *out);

#ifdef __cplusplus
}
#endif
#endif
