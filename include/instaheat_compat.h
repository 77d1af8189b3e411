/*
 * instaheat_compat.h -- the reference's own hot-path names bound to one
 * global simulation, for a C99 host written against the reference's
 * toolbox.h / rk4.h / dopri853.h:
 *
 *   evo_flags            src/toolbox.h:16-26
 *   mk_rhs               src/toolbox.h:28
 *   fft / ifft           src/toolbox.h:32-33 (device pointers here)
 *   run_rk4              src/rk4.h:13
 *   run_dopri853         src/dopri853.h:51
 *   run_gsl_stepper      src/gsl_stepper.h:12 (GSL_OUTPUT_NUMBER from
 *                        ih_compat_gsl_output_number, default 1000)
 *   prepare_and_save_timeslice  src/toolbox.h:29
 *   mean / variance      src/toolbox.h:34-35 (device pointers here)
 *
 * ih_compat_bind() selects the simulation these act on (the reference's
 * globals pars / field / dfield); ih_compat_save_hook() installs the save()
 * callback, ih_compat_file_hooks() the output file's buffer and followup
 * callbacks of run_dopri853's wrap-up (ih_dp_opts.buffered / .followup;
 * with libinstaheat_h5: ih_h5_buffered, ih_h5_followup and the ih_h5).  The
 * reference functions return nothing and exit on fatal errors; here the last
 * error code is kept in ih_compat_status().
 */
#ifndef INSTAHEAT_COMPAT_H
#define INSTAHEAT_COMPAT_H

#include <complex.h>

#include "instaheat.h"

#ifdef __cplusplus
extern "C" {
#endif

struct evolution_flags {
    uint8_t filter;
    uint8_t output;
};
extern struct evolution_flags evo_flags;

void ih_compat_bind(ih_sim *s);
void ih_compat_save_hook(ih_save_fn save, void *user);
void ih_compat_file_hooks(long (*buffered)(void *), int (*followup)(void *),
                          void *file);
int ih_compat_status(void);

void mk_rhs(const double t, double *f, double *result);
void fft(double *in, double _Complex *out);
void ifft(double _Complex *in, double *out);
void run_rk4(void);
void run_dopri853(void);
void run_gsl_stepper(void);
void ih_compat_gsl_output_number(long n);
void prepare_and_save_timeslice(void);
double mean(const double *f, const size_t N);
double variance(const double mean, const double *f, const size_t N);

#ifdef __cplusplus
}
#endif
#endif
