/*
 * instaheat.h -- C99 host of the InstaHeat hot path on MI355X.
 *
 * One ih_sim holds what the reference keeps in globals (pars, field, dfield,
 * field_new, dfield_new, evo_flags, mon; src/main_template.h:313-355) for
 * one parameter set, with the state arrays resident in HBM.  The steppers
 * follow run_rk4 (src/rk4.c:27-132) and run_dopri853 with perform_step /
 * try_step / error / success (src/dopri853.c:112-540) step for step; the
 * arithmetic runs in libihg (include/ihg.h).
 *
 * instaheat_compat.h binds the reference's own names (mk_rhs, run_rk4,
 * run_dopri853, fft, ifft, evo_flags) to one global ih_sim.
 */
#ifndef INSTAHEAT_H
#define INSTAHEAT_H

#include <stddef.h>
#include <stdint.h>

#include "ihg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* INTEGRATION_METHOD values (src/main_template.h:51-56) */
#define IH_RK4 0
#define IH_RKF45 1
#define IH_RKCK 2
#define IH_DOPRI89 3
#define IH_DOPRI853 4

/* evo_flags (src/toolbox.h:16-24) */
struct ih_evolution_flags {
    uint8_t filter;
    uint8_t output;
};

/* the parameters of src/parameters.sh that the hot path reads */
typedef struct ih_config {
    ihg_params grid;          /* GRIDPOINTS_*, SPATIAL_*, MASS, ENABLE_GW,
                                 ENABLE_FFT_FILTER, POWER_SPECTRUM_BINS */
    int integration_method;   /* INTEGRATION_METHOD (IH_RK4, IH_RKF45,
                                 IH_RKCK, IH_DOPRI89, IH_DOPRI853) */
    double ti, tf, dt;        /* INITIAL_TIME, FINAL_TIME, DELTA_T */
    double dt_min;            /* MINIMAL_DELTA_T */
    long max_steps;           /* MAX_STEPS */
    double beta, safe;        /* BETA, SAFE */
    double minscale, maxscale;/* SMALLEST_SCALING, LARGEST_SCALING */
    double rtol, atol;        /* RELATIVE_TOLERANCE, ABSOLUTE_TOLERANCE */
    double max_dt_hubble_fraction; /* MAX_DT_HUBBLE_FRACTION (<= 0: off) */
    long skip;                /* TIME_STEP_SKIPS */
    double a_initial;         /* A_INITIAL */
    double max_runtime;       /* MAX_RUNTIME: run_dopri853 stops after a
                                 step once this many seconds have passed
                                 since ih_create (src/dopri853.c:145-149);
                                 <= 0: no limit */
    int enable_followup;      /* ENABLE_FOLLOWUP: run_dopri853's wrap-up
                                 calls ih_dp_opts.followup
                                 (src/dopri853.c:229-231) */
    int enable_stiffness_check; /* ENABLE_STIFFNESSCHECK
                                   (src/dopri853.c:295-299, 550-580) */
} ih_config;

/* parameters.sh defaults (src/parameters.sh) for the given grid */
void ih_config_defaults(ih_config *cfg);

typedef struct ih_sim ih_sim;

/* save() hook (src/io.c:494-613): time, scale factor and the packed
 * observables of include/ihg.h's ihg_observables */
typedef void (*ih_save_fn)(void *user, double t, double a, const double *obs,
                           size_t nobs);

/* allocate_and_init_all (src/setup.c:89-108) without the initial data */
int ih_create(const ih_config *cfg, ih_sim **out);
void ih_destroy(ih_sim *s);
ihg_ctx *ih_gpu(ih_sim *s);
const ih_config *ih_cfg(const ih_sim *s);
/* device pointers of field / dfield (Ntot doubles each) */
double *ih_field(ih_sim *s);
double *ih_dfield(ih_sim *s);
struct ih_evolution_flags *ih_flags(ih_sim *s);
double ih_time(const ih_sim *s);
/* sets t and clears the step history the proposer reads */
void ih_set_time(ih_sim *s, double t);
/* pars.t.ti: INITIAL_TIME of the next run (the GSL driver's output times,
 * src/gsl_stepper.c:60); the HDF5 readers set it with the time */
void ih_set_ti(ih_sim *s, double ti);
/* pars.t.dt: the first step size of the next run (DELTA_T initially) */
void ih_set_dt(ih_sim *s, double dt);
/* MAX_STEPS of the next run */
void ih_set_max_steps(ih_sim *s, long max_steps);
/* FINAL_TIME of the next run */
void ih_set_tf(ih_sim *s, double tf);

/* mk_rhs(t, f, result) with the flags in ih_flags(s) */
int ih_mk_rhs(ih_sim *s, double t, double *f, double *result);

/* Initial data: a seeded Gaussian random field delta phi with
 * |delta phi_k|^2 = amplitude^2 / (2 a^3 omega_k), omega_k^2 = k^2 + m^2,
 * for 0 < |k| <= k_cut (counter-based RNG, Box-Muller, uniform phases);
 * phi = phi0 + delta phi, dphi = dphi0, a = a0, psi = dpsi = h = 0. */
int ih_init_gaussian(ih_sim *s, uint64_t seed, double amplitude, double phi0,
                     double dphi0, double k_cut);
/* mk_initial_psi (src/setup.c:825-903) on the current field */
int ih_mk_initial_psi(ih_sim *s);
/* initial data from a .dat file (read_initial_data, src/io.c:801-820, and
 * init_from_dat, src/setup.c:802-813): N lines "i j k phi dphi psi dpsi";
 * with_psi = 0 is IC_FROM_DAT_FILE_WITHOUT_PSI (psi from mk_initial_psi) */
int ih_read_dat(ih_sim *s, const char *path, int with_psi);
/* the current phi, dphi, psi, dpsi in that format (no reference writer) */
int ih_write_dat(ih_sim *s, const char *path);

/* prepare_and_save_timeslice (src/toolbox.c:96-104): a filtered output
 * mk_rhs on field into dfield, then save() */
int ih_prepare_and_save_timeslice(ih_sim *s, ih_save_fn save, void *user);

/* run_rk4 (src/rk4.c:27-132); returns the number of steps or < 0 */
long ih_run_rk4(ih_sim *s, ih_save_fn save, void *user);

/* one record per attempted DOPRI853 step; when a log or a proposer is set,
 * a, var_phi and mean_psi2 are ihg_features() of the state after the step
 * (of the state before it for rejected attempts) */
typedef struct ih_dp_record {
    double t, dt, err, rho_mean;
    double a, var_phi, mean_psi2;
    int accepted;
} ih_dp_record;

/* why run_dopri853 stopped (the messages of src/dopri853.c:136-153, 280) */
#define IH_STOP_FINAL_TIME 1   /* "Reached specified final time." */
#define IH_STOP_MAX_STEPS 2    /* "Reached maximal number of steps." */
#define IH_STOP_MAX_RUNTIME 3  /* "Reached specified maximal runtime." */
#define IH_STOP_DT_MIN 4       /* "Stepsize underflow." (dt_next <= dt_min) */
#define IH_STOP_UNDERFLOW 5    /* perform_step's underflow (|dt| <= |t| eps) */
#define IH_STOP_FORCED 6       /* the forced step sequence is done */
#define IH_STOP_STIFF 7        /* "Potential stiffness detected" */
/* run_dopri853's return value when the stiffness check fires (the
 * reference exits the process there, without the wrap-up) */
#define IH_ESTIFF (-2001)

typedef struct ih_dp_stats {
    long n_stp, n_ok, n_bad, attempts;
    double t, dt_next, err_old;
    double integration_seconds; /* mon.integration (src/dopri853.c:117,156) */
    unsigned long calls_rhs;
    int stop;                   /* IH_STOP_* */
    double dtlamb;              /* dp.dtlamb of the stiffness check */
} ih_dp_stats;

/* The stiffness check's counters (check_for_stiffness, src/dopri853.c:
 * 550-580): ih_stiffness_update takes the two sums of ihg_diff_norms (num =
 * |dfield_new - k12|^2, den = |field_new - stage 12 input|^2) and the step's
 * first trial dt, and returns 1 on the 15th stiff verdict in a row (at most
 * 5 non-stiff ones between them). */
typedef struct ih_stiffness {
    double dtlamb;
    int stiff, nonstiff;
} ih_stiffness;
int ih_stiffness_update(ih_stiffness *st, double dt, double num, double den);

/* Optional step proposer, consulted after success() has accepted a step
 * that did not follow a rejection: history holds the last n_history <= 8
 * attempts, oldest first, the accepted one last; dt_classic is success()'s
 * dt_next.  A positive finite return value replaces dt_next, capped by the
 * same Hubble limit; the embedded error estimate still judges every step,
 * and rejections fall back to the classic controller. */
#define IH_HISTORY 8
typedef double (*ih_dt_proposer)(void *user, const ih_dp_record *history,
                                 long n_history, double dt_classic);

typedef struct ih_dp_opts {
    const double *forced_dt;  /* non-NULL: take exactly these steps */
    long n_forced;
    ih_dp_record *log;        /* may be NULL */
    long log_cap;
    ih_dt_proposer proposer;  /* may be NULL */
    void *proposer_user;
    /* time-based output by dense output: after every accepted step from t
     * to t + dt, save() is called for each t_out[i] in (t, t + dt] (the
     * step-based outputs every `skip` steps are kept); t_out increasing */
    const double *t_out;
    long n_out;
    /* The output file (src/io.c:494-613, 822-827), both may be NULL:
     * buffered() = slices saved but not yet written (the reference's
     * pars.file.index); with it the wrap-up saves the final slice only when
     * buffered() != 0 and the last save is older than t, as
     * src/dopri853.c:223-226 do; without it, whenever the last save is
     * older than t.  followup() writes the followup state (h5_write_followup)
     * in the wrap-up when enable_followup is set. */
    long (*buffered)(void *file);
    int (*followup)(void *file);
    void *file;
} ih_dp_opts;

/* the state of the latest dense output (device, Ntot doubles); valid
 * inside the save() call of a time-based output */
double *ih_dense_state(ih_sim *s);
/* inside a save() call: the device state whose observables are passed
 * (field, or the dense output's state); NULL outside save() */
const double *ih_save_state(const ih_sim *s);

/* run_dopri853 (src/dopri853.c:112-158) */
long ih_run_dopri853(ih_sim *s, const ih_dp_opts *opts, ih_save_fn save,
                     void *user, ih_dp_stats *stats);

/* ---- GSL steppers ----------------------------------------------------
 * run_gsl_stepper (src/gsl_stepper.c:41-76): GSL's ODE driver with the
 * stepper of integration_method (IH_RKF45, IH_RKCK, IH_DOPRI89), start
 * step dt, hmin = dt_min, control y_new(atol, rtol), over n_output
 * (GSL_OUTPUT_NUMBER) intervals ending at i * (tf - ti) / n_output (as the
 * reference computes them), prepare_and_save_timeslice() at ti and after
 * each interval.  Within a step the filter is off, as in the reference.
 * The driver stops at the first GSL failure (status, the reference prints
 * it and leaves the loop); the call still returns the accepted steps. */
#define IH_GSL_SUCCESS 0
#define IH_GSL_EINVAL 4    /* an interval ends before the current time */
#define IH_GSL_EMAXITER 11
#define IH_GSL_ENOPROG 27  /* the suggested step fell below hmin */

typedef struct ih_gsl_stats {
    long steps;        /* accepted steps */
    long attempts;     /* step attempts (GSL's evolve count) */
    long failed_steps; /* attempts undone after the control's DEC verdict */
    long intervals;    /* output intervals completed */
    int status;        /* IH_GSL_* of the driver */
    double t, h;       /* time reached, suggested next step */
    double integration_seconds;
    unsigned long calls_rhs; /* RHS evaluations (GSL also evaluates the
                                RHS at the new point of every attempt,
                                unused by this control; not made here) */
} ih_gsl_stats;

long ih_run_gsl(ih_sim *s, long n_output, ih_save_fn save, void *user,
                ih_gsl_stats *stats);
/* the standard control's step adjustment for the error ratio rmax: returns
 * -1 (decrease), 0 (keep) or 1 (increase) and updates *h */
int ih_gsl_hadjust(double rmax, double *h, int order);

/* ---- learned step proposer -------------------------------------------
 * An ih_dt_proposer over libihg's MLP (no counterpart in the reference; the
 * proposer form of the DOPRI853 controller).  Raw inputs of the newest
 * record, oldest first over the last IH_HISTORY attempts (the first attempt
 * repeated while fewer exist):
 *   x[0..7]   log10 err   (floored at 1e-12)
 *   x[8..15]  log10 dt
 *   x[16] log10 a, x[17] log10 H (H = sqrt(rho_mean / 3)),
 *   x[18] log10 (var phi + 1e-300), x[19] log10 (<psi^2> + 1e-300).
 * The network sees (x - in_mean) / in_std; its output y gives
 * log(dt_next / dt) = y * out_std + out_mean, clamped to
 * [log min_ratio, log max_ratio]. */
typedef struct ih_mlp_proposer {
    ihg_mlp *mlp;
    float in_mean[IHG_MLP_IN], in_std[IHG_MLP_IN];
    float out_mean, out_std;
    double min_ratio, max_ratio;
    long calls; /* proposals made */
} ih_mlp_proposer;
void ih_proposer_features(const ih_dp_record *history, long n_history,
                          double x[IHG_MLP_IN]);
/* user = ih_mlp_proposer *; returns 0 (keep dt_classic) on any failure */
double ih_mlp_propose(void *user, const ih_dp_record *history,
                      long n_history, double dt_classic);

/* ---- ensembles: the batched form (no counterpart in the reference) ----
 * nreal realizations of one grid in one ensemble context (ihg.h): fixed-step
 * RK4 of run_rk4 (src/rk4.c:78-102) for all of them at once.  Several
 * ensembles (one per grid size) step together in ih_ens_run_rk4; with
 * use_graph each ensemble's step is captured once as a HIP graph and
 * replayed once per step on that ensemble's own stream, the sizes side by
 * side. */
typedef struct ih_ens ih_ens;
int ih_ens_create(const ih_config *cfg, int nreal, ih_ens **out);
void ih_ens_destroy(ih_ens *e);
ihg_ctx *ih_ens_gpu(ih_ens *e);
double *ih_ens_state(ih_ens *e); /* nreal * stride doubles, device */
long ih_ens_stride(const ih_ens *e);
int ih_ens_nreal(const ih_ens *e);
double ih_ens_time(const ih_ens *e);
unsigned long ih_ens_calls_rhs(const ih_ens *e);
/* Gaussian initial data per realization (ih_init_gaussian's recipe), then
 * mk_initial_psi of each */
int ih_ens_init(ih_ens *e, const uint64_t *seeds, const double *amplitudes,
                double phi0, double dphi0, double k_cut);
int ih_ens_capture_step(ih_ens *e, double dt, ihg_graph **g);
int ih_ens_run_rk4(ih_ens *const *es, int n, long nsteps, double dt,
                   int use_graph);
/* the same steps with every pass of all ensembles in one launch (an
 * ihg_group over their contexts, on the first ensemble's stream); with
 * use_graph one RK4 step of all of them is captured once as a single HIP
 * graph and replayed once per step.  The end states equal ih_ens_run_rk4's
 * bit for bit when the contexts keep their default strided passes (no
 * autotuned choice). */
int ih_ens_run_rk4_grouped(ih_ens *const *es, int n, long nsteps, double dt,
                           int use_graph);
/* one RK4 step of all the ensembles, grouped, captured as one HIP graph to
 * replay on the first ensemble's context (ihg_graph_launch); the replays do
 * not advance the ensembles' times */
int ih_ens_capture_step_grouped(ih_ens *const *es, int n, double dt,
                                ihg_graph **g);
/* the observables of realization r (an output mk_rhs on a copy of it) */
int ih_ens_observables(ih_ens *e, int r, double *obs, size_t len);

/* success() (src/dopri853.c:505-540) on a controller state, exported so its
 * arithmetic can be checked against the oracle's */
typedef struct ih_ctrl {
    double err_old, dt_next;
    int reject;
} ih_ctrl;
int ih_success(ih_ctrl *ctl, const ih_config *cfg, double err, double *dt,
               double rho_mean);
/* error() (src/dopri853.c:469-486) from the two sums of ihg_dp853_finish */
double ih_error_norm(double e5, double e3, double dt, size_t ntot);

#ifdef __cplusplus
}
#endif
#endif
