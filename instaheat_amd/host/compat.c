/*
 * compat.c -- the reference's hot-path names (include/instaheat_compat.h)
 * over the C99 host of instaheat.c.
 */
#include "instaheat_compat.h"

struct evolution_flags evo_flags = {0, 0};

static ih_sim *g_sim = NULL;
static ih_save_fn g_save = NULL;
static void *g_user = NULL;
static int g_status = 0;
static long g_gsl_outputs = 1000; /* GSL_OUTPUT_NUMBER, src/parameters.sh:33 */

void ih_compat_bind(ih_sim *s) { g_sim = s; }

void ih_compat_save_hook(ih_save_fn save, void *user)
{
    g_save = save;
    g_user = user;
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
    long n = ih_run_dopri853(g_sim, NULL, g_save, g_user, NULL);
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
