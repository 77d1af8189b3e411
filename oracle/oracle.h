/*
 * oracle.h -- CPU restatement (C99 + OpenMP) of InstaHeat's right-hand side
 * and time steppers.  TEST INFRASTRUCTURE ONLY: it is the checker for the HIP
 * path and the timed CPU baseline, never part of the shipped product.  Only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
 *
 * Every function follows a reference function; the citation is given at the
 * definition in oracle.c.  The FFT is this project's own (radix-2 for powers
 * of two, direct DFT otherwise) behind the same r2c/c2r contract as the
 * reference's FFTW calls (src/toolbox.c:806-827): unnormalised, sign -1
 * forward, half spectrum along the last (z) axis.
 */
#ifndef IHG_ORACLE_H
#define IHG_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runtime form of the compile-time parameters of src/main_template.h:116-162 */
typedef struct orc_params {
    int nx, ny, nz;          /* GRIDPOINTS_X/Y/Z */
    double xa, xb;           /* SPATIAL_LOWER/UPPER_BOUND_X */
    double ya, yb;           /* ..._Y */
    double za, zb;           /* ..._Z */
    double mass;             /* MASS */
    int enable_gw;           /* ENABLE_GW */
    int enable_filter;       /* ENABLE_FFT_FILTER */
    int bins;                /* POWER_SPECTRUM_BINS */
    int threads;             /* THREAD_NUMBER (<=0: OpenMP default) */
    int log_bins;            /* 1: log-spaced spectrum bins (no counterpart
                                in the reference, which bins linearly) */
} orc_params;

/* Observables of one output step, packed in this order:
 * smry[6][4] (phi, dphi, psi, dpsi, rho, pressure: mean, var, min, max),
 * cstr[4], phi_ps[bins], psi_ps[bins], rho_ps[bins], gw[bins] */
#define ORC_OBS_LEN(bins) (24 + 4 + 4 * (size_t)(bins))

typedef struct orc_ctx orc_ctx;

typedef void (*orc_save_fn)(void *user, double t, double a, const double *obs,
                            size_t nobs);

orc_ctx *orc_create(const orc_params *p);
void orc_destroy(orc_ctx *c);
void orc_sizes(const orc_ctx *c, size_t out[4]); /* N, M, Next, Ntot */
int orc_threads(const orc_ctx *c);

void orc_fft_r2c(orc_ctx *c, const double *in, double *out);
void orc_fft_c2r(orc_ctx *c, const double *in, double *out);
void orc_kgrid(const orc_ctx *c, double *kx, double *ky, double *kz,
               double *kxf, double *kyf, double *kzf, double *ksq,
               double *kmax);
void orc_filter_mask(const orc_ctx *c, double *out);

void orc_set_field(orc_ctx *c, double *field);
void orc_mk_rhs(orc_ctx *c, double t, double *f, double *result, int filter,
                int output);
void orc_means(const orc_ctx *c, double out[2]);
void orc_obs(const orc_ctx *c, double *out);
unsigned long orc_calls_rhs(const orc_ctx *c);
void orc_aux(const orc_ctx *c, int which, double *out);

void orc_mk_initial_psi(orc_ctx *c, double ti, double *field, double *dfield);

double orc_mean(const double *f, size_t n);
double orc_variance(double mean, const double *f, size_t n);

/* run_rk4 (src/rk4.c:27-132); returns the number of steps taken */
long orc_run_rk4(orc_ctx *c, double *field, double ti, double tf, double dt,
                 long max_steps, long skip, orc_save_fn save, void *user);

typedef struct orc_dp_opts {
    double ti, tf, dt;           /* INITIAL_TIME, FINAL_TIME, DELTA_T */
    double dt_min;               /* MINIMAL_DELTA_T */
    long max_steps;              /* MAX_STEPS */
    double beta, safe;           /* BETA, SAFE */
    double minscale, maxscale;   /* SMALLEST_SCALING, LARGEST_SCALING */
    double rtol, atol;           /* RELATIVE/ABSOLUTE_TOLERANCE */
    double max_dt_hubble_fraction; /* MAX_DT_HUBBLE_FRACTION (<=0: off) */
    long skip;                   /* TIME_STEP_SKIPS */
    const double *forced_dt;     /* if non-NULL: accept exactly these steps */
    long n_forced;
    int enable_stiffness_check;  /* ENABLE_STIFFNESSCHECK */
} orc_dp_opts;

typedef struct orc_dp_stats {
    long n_stp, n_ok, n_bad;
    double t;
    double last_err;
    double dtlamb;               /* dp.dtlamb of the stiffness check */
    int stiff;                   /* 1: stopped by the stiffness check */
} orc_dp_stats;

/* check_for_stiffness's counters (src/dopri853.c:550-580); returns 1 on
 * the 15th stiff verdict */
typedef struct orc_stiffness {
    double dtlamb;
    int stiff, nonstiff;
} orc_stiffness;
int orc_stiffness_update(orc_stiffness *st, double dt, double num, double den);

/* run_dopri853 (src/dopri853.c:112-158).  dt_log/err_log (may be NULL)
 * receive every attempted dt and its error norm, up to log_cap entries;
 * accepted[] marks accepted attempts.  Returns the number of attempts. */
long orc_run_dopri853(orc_ctx *c, double *field, double *dfield,
                      const orc_dp_opts *o, orc_dp_stats *st,
                      orc_save_fn save, void *user, double *dt_log,
                      double *err_log, int *accepted, long log_cap);

/* success() (src/dopri853.c:505-540) as a pure function of the controller
 * state; exported so that its arithmetic can be checked in isolation. */
typedef struct orc_ctrl {
    double err_old, dt_next;
    int reject;
} orc_ctrl;
int orc_success(orc_ctrl *s, const orc_dp_opts *o, double err, double *dt,
                double rho_mean);

/* DOP853 tableau as used by the steppers (row-major a[13][12], c[13],
 * b[12], bhh[3], er[12]). */
void orc_dp_tableau(double *a, double *c, double *b, double *bhh, double *er);

#ifdef __cplusplus
}
#endif
#endif
