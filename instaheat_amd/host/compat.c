/*
 * compat.c -- the reference's hot-path names (include/instaheat_compat.h)
 * over the C99 host of instaheat.c.
 */
#include "instaheat_compat.h"

#include <math.h>
#include <string.h>

struct evolution_flags evo_flags = {0, 0};

static ih_sim *g_sim = NULL;
static ih_save_fn g_save = NULL;
static void *g_user = NULL;
static int g_status = 0;
static long g_gsl_outputs = 1000; /* GSL_OUTPUT_NUMBER, src/parameters.sh:33 */
static long (*g_buffered)(void *) = NULL;
static int (*g_followup)(void *) = NULL;
static void *g_file = NULL;

void ih_compat_bind(ih_sim *s) { g_sim = s; }

void ih_compat_save_hook(ih_save_fn save, void *user)
{
    g_save = save;
    g_user = user;
}

void ih_compat_file_hooks(long (*buffered)(void *), int (*followup)(void *),
                          void *file)
{
    g_buffered = buffered;
    g_followup = followup;
    g_file = file;
}

int ih_compat_status(void) { return g_status; }

static int bound(void)
{
    if (!g_sim) {
        g_status = IHG_EINVAL;
        return 0;
    }
    return 1;
}

/* mk_rhs, src/toolbox.c:75-88: the flags come from the global evo_flags */
void mk_rhs(const double t, double *f, double *result)
{
    if (!bound()) return;
    struct ih_evolution_flags *fl = ih_flags(g_sim);
    fl->filter = evo_flags.filter;
    fl->output = evo_flags.output;
    g_status = ih_mk_rhs(g_sim, t, f, result);
    fl->filter = fl->output = 0;
}

/* fft / ifft, src/toolbox.c:806-827 (device arrays) */
void fft(double *in, double _Complex *out)
{
    if (!bound()) return;
    g_status = ihg_fft_r2c(ih_gpu(g_sim), in, (double *)out);
}

void ifft(double _Complex *in, double *out)
{
    if (!bound()) return;
    g_status = ihg_fft_c2r(ih_gpu(g_sim), (const double *)in, out);
}

/* run_rk4, src/rk4.c:27-132 */
void run_rk4(void)
{
    if (!bound()) return;
    long n = ih_run_rk4(g_sim, g_save, g_user);
    g_status = n < 0 ? (int)n : 0;
}

/* run_dopri853, src/dopri853.c:112-158 */
void run_dopri853(void)
{
    if (!bound()) return;
    ih_dp_opts opts;
    memset(&opts, 0, sizeof opts);
    opts.buffered = g_buffered;
    opts.followup = g_followup;
    opts.file = g_file;
    long n = ih_run_dopri853(g_sim, &opts, g_save, g_user, NULL);
    g_status = n < 0 ? (int)n : 0;
}

void ih_compat_gsl_output_number(long n) { g_gsl_outputs = n; }

/* run_gsl_stepper, src/gsl_stepper.c:41-76 */
void run_gsl_stepper(void)
{
    if (!bound()) return;
    long n = ih_run_gsl(g_sim, g_gsl_outputs, g_save, g_user, NULL);
    g_status = n < 0 ? (int)n : 0;
}

/* prepare_and_save_timeslice, src/toolbox.c:96-104 */
void prepare_and_save_timeslice(void)
{
    if (!bound()) return;
    g_status = ih_prepare_and_save_timeslice(g_sim, g_save, g_user);
}

/* mean / variance, src/toolbox.c:726-754, of device arrays; NaN (and the
 * error code in ih_compat_status) on failure */
double mean(const double *f, const size_t N)
{
    double m = NAN;
    if (!bound()) return m;
    g_status = ihg_mean(ih_gpu(g_sim), f, N, &m);
    return g_status ? NAN : m;
}

double variance(const double mean, const double *f, const size_t N)
{
    double v = NAN;
    if (!bound()) return v;
    g_status = ihg_variance(ih_gpu(g_sim), mean, f, N, &v);
    return g_status ? NAN : v;
}
