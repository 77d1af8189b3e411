/*
 * instaheat_h5.h -- the reference's HDF5 output file (src/io.c, src/io.h)
 * for the MI355X host: the same dataset names, ranks, shapes, chunking and
 * layouts, written from the save() hook of include/instaheat.h, plus the
 * readers of IC_FROM_H5_FILE.
 *
 * libinstaheat_h5.so is a separate, optional library (C99 over the HDF5
 * 1.10 C API) so that libinstaheat does not depend on HDF5.  Every call
 * returns 0 or a negative IHG_E* code (message from ih_h5_last_error());
 * nothing exits the process.
 *
 * File contents (src/io.c:47-257, 494-613):
 *   per time slice, datasets of shape (n_slices, dim), chunked by
 *   buf_size slices (WRITE_OUT_BUFFER_NUMBER), time extensible:
 *     time, a                       dim 1 (rank 1)
 *     phi_summary ... pressure_summary   mean, variance, min, max
 *     constraints                   NUMBER_CONSTRAINTS = 4
 *     phi/psi/rho_power_spectrum, gravitational_wave_spectrum   bins
 *     phi, dphi, psi, dpsi, rho     the strided grid (LARGE_OUTPUT)
 *   once, compact 1-D datasets: mass, inflaton_mass, dimension, seed,
 *     threads, fftw_threads, strides_time, tolerances,
 *     gridpoints_internal, spatial_bounds_x/y/z, gridpoints_output,
 *     strides_space, bunch_davies_cutoff, max_dt_hubble_fraction,
 *     filter, gravitational_waves; strings commit_hash (optional) and
 *     method ("hyperbolic")
 *   at the end: steps_total, steps_ok, steps_bad, runtime_stepper,
 *     calls_rhs (ih_h5_write_dp_stats), followup (contiguous, Ntot),
 *     runtime_total and calls_rhs (ih_h5_close).  The ENABLE_TIMING
 *     monitors runtime_fftw ... runtime_writeout (off by default,
 *     src/parameters.sh:66) are not written; ih_h5_write_simple can add
 *     them.
 * The reference's h1_summary / h2_summary ("not fully functional",
 * src/parameters.sh:89-90, off by default) are not written.
 */
#ifndef INSTAHEAT_H5_H
#define INSTAHEAT_H5_H

#include <stddef.h>

#include "instaheat.h"

#ifdef __cplusplus
extern "C" {
#endif

/* dataset names, as in src/io.h:13-71 (the file format is the interface) */
#define IH_H5_PHI_NAME "phi"
#define IH_H5_DPHI_NAME "dphi"
#define IH_H5_PSI_NAME "psi"
#define IH_H5_DPSI_NAME "dpsi"
#define IH_H5_RHO_NAME "rho"
#define IH_H5_PHI_PS_NAME "phi_power_spectrum"
#define IH_H5_PSI_PS_NAME "psi_power_spectrum"
#define IH_H5_RHO_PS_NAME "rho_power_spectrum"
#define IH_H5_GW_PS_NAME "gravitational_wave_spectrum"
#define IH_H5_TIME_NAME "time"
#define IH_H5_A_NAME "a"
#define IH_H5_PHI_SMRY_NAME "phi_summary"
#define IH_H5_DPHI_SMRY_NAME "dphi_summary"
#define IH_H5_PSI_SMRY_NAME "psi_summary"
#define IH_H5_DPSI_SMRY_NAME "dpsi_summary"
#define IH_H5_RHO_SMRY_NAME "rho_summary"
#define IH_H5_PRESSURE_SMRY_NAME "pressure_summary"
#define IH_H5_CONSTRAINTS_NAME "constraints"
#define IH_H5_MASS_NAME "mass"
#define IH_H5_INFLATON_MASS_NAME "inflaton_mass"
#define IH_H5_DIMENSION_NAME "dimension"
#define IH_H5_SEED_NAME "seed"
#define IH_H5_THREAD_NUMBER_NAME "threads"
#define IH_H5_FFTW_THREAD_NUMBER_NAME "fftw_threads"
#define IH_H5_STRIDES_TIME_NAME "strides_time"
#define IH_H5_STRIDES_SPACE_NAME "strides_space"
#define IH_H5_TOLERANCES_NAME "tolerances"
#define IH_H5_GRIDPOINTS_INTERNAL_NAME "gridpoints_internal"
#define IH_H5_GRIDPOINTS_OUTPUT_NAME "gridpoints_output"
#define IH_H5_SPATIAL_BOUNDS_X_NAME "spatial_bounds_x"
#define IH_H5_SPATIAL_BOUNDS_Y_NAME "spatial_bounds_y"
#define IH_H5_SPATIAL_BOUNDS_Z_NAME "spatial_bounds_z"
#define IH_H5_COMMIT_HASH_NAME "commit_hash"
#define IH_H5_METHOD_NAME "method"
#define IH_H5_RUNTIME_TOTAL_NAME "runtime_total"
#define IH_H5_RUNTIME_STEPPER_NAME "runtime_stepper"
#define IH_H5_COUNTER_RHS "calls_rhs"
#define IH_H5_STEPS_TOTAL_NAME "steps_total"
#define IH_H5_STEPS_OK_NAME "steps_ok"
#define IH_H5_STEPS_BAD_NAME "steps_bad"
#define IH_H5_MAX_DT_HUBBLE_FRACTION_NAME "max_dt_hubble_fraction"
#define IH_H5_BUNCH_DAVIES_CUTOFF_NAME "bunch_davies_cutoff"
#define IH_H5_ENABLE_FILTER_NAME "filter"
#define IH_H5_ENABLE_GW_NAME "gravitational_waves"
#define IH_H5_FOLLOWUP_NAME "followup"

/* the OUTPUT_* switches of src/parameters.sh:71-92 as a bit set */
#define IH_H5_OUT_PHI (1u << 0)
#define IH_H5_OUT_DPHI (1u << 1)
#define IH_H5_OUT_PSI (1u << 2)
#define IH_H5_OUT_DPSI (1u << 3)
#define IH_H5_OUT_RHO (1u << 4)
#define IH_H5_OUT_PHI_PS (1u << 5)
#define IH_H5_OUT_PSI_PS (1u << 6)
#define IH_H5_OUT_RHO_PS (1u << 7)
#define IH_H5_OUT_PHI_SMRY (1u << 8)
#define IH_H5_OUT_DPHI_SMRY (1u << 9)
#define IH_H5_OUT_PSI_SMRY (1u << 10)
#define IH_H5_OUT_DPSI_SMRY (1u << 11)
#define IH_H5_OUT_RHO_SMRY (1u << 12)
#define IH_H5_OUT_PRESSURE_SMRY (1u << 13)
#define IH_H5_OUT_CONSTRAINTS (1u << 14)
/* the full-field outputs (LARGE_OUTPUT) */
#define IH_H5_OUT_FIELDS (IH_H5_OUT_PHI | IH_H5_OUT_DPHI | IH_H5_OUT_PSI | \
                          IH_H5_OUT_DPSI | IH_H5_OUT_RHO)
/* src/parameters.sh's defaults: spectra, summaries, constraints on */
#define IH_H5_OUT_DEFAULT (IH_H5_OUT_PHI_PS | IH_H5_OUT_PSI_PS |            \
                           IH_H5_OUT_RHO_PS | IH_H5_OUT_PHI_SMRY |          \
                           IH_H5_OUT_DPHI_SMRY | IH_H5_OUT_PSI_SMRY |       \
                           IH_H5_OUT_DPSI_SMRY | IH_H5_OUT_RHO_SMRY |       \
                           IH_H5_OUT_PRESSURE_SMRY | IH_H5_OUT_CONSTRAINTS)

typedef struct ih_h5_opts {
    unsigned outputs;         /* IH_H5_OUT_* (the GW spectrum follows
                                 ENABLE_GW) */
    long buf_size;            /* WRITE_OUT_BUFFER_NUMBER: time slices held
                                 before a write, and the chunk's length */
    long stride[3];           /* STRIDE_X/Y/Z of the full-field outputs */
    double seed;              /* SEED, written as metadata */
    double inflaton_mass;     /* INFLATON_MASS; -1 for other initial data */
    double bunch_davies_cutoff; /* BUNCH_DAVIES_CUTOFF */
    double threads, fftw_threads; /* THREAD_NUMBER, FFTW_THREAD_NUMBER as
                                 written (the host here is one thread) */
    const char *commit_hash;  /* NULL: no commit_hash dataset
                                 (VERSION_CONTROL_NONE); else at most 15
                                 characters are kept */
} ih_h5_opts;

/* parameters.sh's values: IH_H5_OUT_DEFAULT, buf_size 5, strides 1,
 * seed 0, inflaton_mass -1, cutoff 0, threads 1, no commit hash */
void ih_h5_opts_defaults(ih_h5_opts *o);

typedef struct ih_h5 ih_h5;

/* h5_create_empty_by_path (src/io.c:47-257): create (truncate) the file,
 * its time-series datasets and the parameter datasets of cfg and o */
int ih_h5_create(const char *path, const ih_config *cfg, const ih_h5_opts *o,
                 ih_h5 **out);
/* save() (src/io.c:494-613) from host arrays: obs packed as
 * ihg_observables (nobs = 28 + 4 bins); fields = phi, dphi, psi, dpsi of
 * the saved state (4 N doubles) and rho (N doubles), each read only when a
 * full-field output needs it (may be NULL otherwise).  A full buffer is
 * written to the file (h5_write_all_buffers, src/io.c:336-432). */
int ih_h5_append(ih_h5 *h, double t, double a, const double *obs,
                 size_t nobs, const double *fields, const double *rho);
/* Bind a simulation: ih_h5_save then takes the full fields from its saved
 * state (ih_save_state) and rho from the output call, and
 * ih_h5_write_followup reads its state. */
int ih_h5_bind(ih_h5 *h, ih_sim *s);
/* an ih_save_fn (user = the ih_h5) for ih_run_rk4 / ih_run_dopri853;
 * the first failure is kept (ih_h5_status) and later saves are dropped */
void ih_h5_save(void *h5, double t, double a, const double *obs,
                size_t nobs);
int ih_h5_status(const ih_h5 *h);
/* time slices appended so far (written or buffered) */
long ih_h5_slices(const ih_h5 *h);
/* h5_write_simple (src/io.c:306-323): a 1-D double dataset, compact
 * layout (at most 64 KiB), written once */
int ih_h5_write_simple(ih_h5 *h, const char *name, const double *val,
                       size_t n);
/* the meta data of wrap_up_dopri853 and main (src/dopri853.c:234-246,
 * src/main.c:101-106): steps_total, steps_ok, steps_bad, runtime_stepper,
 * calls_rhs */
int ih_h5_write_dp_stats(ih_h5 *h, const ih_dp_stats *st);
/* h5_write_followup (src/io.c:822-827): the bound simulation's state
 * (Ntot doubles, contiguous layout) */
int ih_h5_write_followup(ih_h5 *h);
/* ih_dp_opts.buffered / .followup for run_dopri853's wrap-up
 * (src/dopri853.c:223-231), file = the ih_h5: slices saved and not yet
 * written (pars.file.index, src/io.c:604-608), and ih_h5_write_followup */
long ih_h5_buffered(void *h5);
int ih_h5_followup(void *h5);
/* h5_close (src/io.c:472-488): write the partial buffer and, as main does
 * at the end of every run (src/main.c:99-106), runtime_total (wall
 * seconds since ih_h5_create) and calls_rhs (the bound simulation's count
 * at the latest ih_h5_save), unless they were written already; close the
 * file, free h (also after an error).  It does not touch the simulation,
 * which may be destroyed first. */
int ih_h5_close(ih_h5 *h);
const char *ih_h5_last_error(void);

/* h5_read_followup (src/io.c:761-801), the IC_FROM_H5_FILE initial data:
 * the state from `followup` (Ntot doubles) and the simulation's time and
 * INITIAL_TIME (ih_set_ti) set to the file's last time */
int ih_h5_read_followup(ih_sim *s, const char *path);
/* h5_read_timeslice (src/io.c:640-711): the first time slice with
 * time + DBL_EPSILON >= t (else the last), phi, dphi, psi, dpsi and a from
 * it (the full-field datasets at stride 1), the other state entries zero;
 * the simulation's time and INITIAL_TIME (ih_set_ti) are set to the
 * slice's time (also in *t_found when not NULL) */
int ih_h5_read_timeslice(ih_sim *s, const char *path, double t,
                         double *t_found);
/* Read a double dataset: dims[0], dims[1] its extent (dims[1] = 1 for
 * rank 1); row < 0 reads all of it (cap >= dims[0] * dims[1]), row >= 0
 * one time slice (cap >= dims[1]).  out may be NULL to ask for the dims
 * only (of any dataset, strings included). */
int ih_h5_read(const char *path, const char *name, long row, double *out,
               size_t cap, size_t dims[2]);
/* Read a string dataset (commit_hash, method) into buf (cap bytes, nul
 * terminated, trailing blanks removed). */
int ih_h5_read_string(const char *path, const char *name, char *buf,
                      size_t cap);

#ifdef __cplusplus
}
#endif
#endif
