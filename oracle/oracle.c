/*
 * oracle.c -- CPU restatement (C99 + OpenMP) of InstaHeat's RHS and steppers.
 *
 * TEST INFRASTRUCTURE ONLY (the parity reference of the tests and the timed
 * CPU baseline).  The shipped HIP path never links or calls this file.
 *
 * Each function cites the reference function it restates (paths relative to
 * the reference repository root).  Loops are written in the reference's
 * order and with the reference's arithmetic expressions so that rounding
 * differences against it stay at the level of reduction order and FFT
 * rounding.  The FFT is this project's own: radix-2 Cooley-Tukey for powers
 * of two and a direct DFT otherwise, applied axis by axis (z real-to-complex
 * by a full complex transform of the zero-imaginary row, then y and x).
 */
#include "oracle.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI (3.141592653589793238462643383279)
#define ORC_TWOPI (6.283185307179586476925286766559)
#define ORC_MAX(x, y) (((x) > (y)) ? (x) : (y))
#define ORC_MIN(x, y) (((x) > (y)) ? (y) : (x))

typedef double complex cplx;

/* one FFT axis: length and e^{-2 pi i m / n} table */
typedef struct orc_axis {
    size_t n;
    int pow2;
    cplx *tw;
} orc_axis;

struct orc_ctx {
    orc_params p;
    size_t Nx, Ny, Nz, Mz, N, M, Next, Ntot;
    double kx1, ky1, kz1, kx2, ky2, kz2; /* pars.{x,y,z}.k and .k2 */
    double *kx, *ky, *kz, *kxf, *kyf, *kzf, *ksq;
    double k2_max, k_max;
    double k_lo, log_span;   /* log bins: fundamental |k|, log(k_max / k_lo) */
    double *filt;
    cplx *phic, *xphic, *yphic, *zphic, *psic, *fc, *deltarhoc, *work;
    double *xphi, *yphi, *zphi, *grad, *lap, *f, *deltarho;
    double *rho, *pressure;
    double rho_mean, pressure_mean;
    cplx *stt[6];
    double *s_tmp;
    double *field; /* the reference's global `field`, read by mk_summary */
    double smry[6][4];
    double cstr[4];
    double *phi_ps, *psi_ps, *rho_ps, *gw;
    unsigned long calls_rhs;
    orc_axis ax, ay, az;
};

/* ------------------------------------------------------------------------ */
/* FFT (this project's own; contract of src/toolbox.c:806-827)              */
/* ------------------------------------------------------------------------ */

static int is_pow2(size_t n) { return n && !(n & (n - 1)); }

static void axis_init(orc_axis *a, size_t n)
{
    a->n = n;
    a->pow2 = is_pow2(n);
    a->tw = malloc(n * sizeof *a->tw);
    for (size_t m = 0; m < n; ++m) {
        long double ang = -2.0L * 3.14159265358979323846264338327950288L *
                          (long double)m / (long double)n;
        a->tw[m] = (double)cosl(ang) + I * (double)sinl(ang);
    }
}

/* in-place 1D complex DFT of length a->n; sign -1 forward, +1 backward;
 * `scratch` holds n entries (used by the direct DFT) */
static void fft1d(const orc_axis *a, cplx *x, int sign, cplx *scratch)
{
    const size_t n = a->n;
    if (n == 1) return;
    if (!a->pow2) {
        for (size_t k = 0; k < n; ++k) {
            cplx s = 0.0;
            for (size_t j = 0; j < n; ++j) {
                cplx w = a->tw[(j * k) % n];
                if (sign > 0) w = conj(w);
                s += x[j] * w;
            }
            scratch[k] = s;
        }
        memcpy(x, scratch, n * sizeof *x);
        return;
    }
    /* bit reversal */
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            cplx t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1, step = n / len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t j = 0; j < half; ++j) {
                cplx w = a->tw[j * step];
                if (sign > 0) w = conj(w);
                cplx u = x[i + j], v = x[i + j + half] * w;
                x[i + j] = u + v;
                x[i + j + half] = u - v;
            }
        }
    }
}

/* complex transform of every line along one axis of a complex array laid out
 * [Nx][Ny][Mz]; `stride` is the distance between consecutive line elements */
static void fft_lines(orc_ctx *c, cplx *data, const orc_axis *a, size_t stride,
                      size_t nlines, size_t inner, size_t outer_stride,
                      int sign)
{
    #pragma omp parallel
    {
        cplx *line = malloc(a->n * sizeof *line);
        cplx *scr = malloc(a->n * sizeof *scr);
        #pragma omp for
        for (size_t l = 0; l < nlines; ++l) {
            size_t base = (l / inner) * outer_stride + (l % inner);
            for (size_t m = 0; m < a->n; ++m) line[m] = data[base + m * stride];
            fft1d(a, line, sign, scr);
            for (size_t m = 0; m < a->n; ++m) data[base + m * stride] = line[m];
        }
        free(line);
        free(scr);
    }
    (void)c;
}

/* r2c 3D, out-of-place, unnormalised (FFTW r2c_3d semantics) */
static void fft_r2c(orc_ctx *c, const double *in, cplx *out)
{
    const size_t Nz = c->Nz, Mz = c->Mz, rows = c->Nx * c->Ny;
    #pragma omp parallel
    {
        cplx *line = malloc(Nz * sizeof *line);
        cplx *scr = malloc(Nz * sizeof *scr);
        #pragma omp for
        for (size_t r = 0; r < rows; ++r) {
            for (size_t z = 0; z < Nz; ++z) line[z] = in[r * Nz + z];
            fft1d(&c->az, line, -1, scr);
            for (size_t k = 0; k < Mz; ++k) out[r * Mz + k] = line[k];
        }
        free(line);
        free(scr);
    }
    /* y lines: for each (x, kz); stride Mz */
    fft_lines(c, out, &c->ay, Mz, c->Nx * Mz, Mz, c->Ny * Mz, -1);
    /* x lines: for each (y, kz); stride Ny*Mz */
    fft_lines(c, out, &c->ax, c->Ny * Mz, c->Ny * Mz, c->Ny * Mz, 0, -1);
}

/* c2r 3D, unnormalised; the output is the real part of the inverse DFT of
 * the Hermitian extension of `in` (FFTW c2r_3d semantics).  `in` is kept. */
static void fft_c2r(orc_ctx *c, const cplx *in, double *out)
{
    const size_t Nz = c->Nz, Mz = c->Mz, rows = c->Nx * c->Ny;
    cplx *w = c->work;
    memcpy(w, in, c->M * sizeof *w);
    fft_lines(c, w, &c->ax, c->Ny * Mz, c->Ny * Mz, c->Ny * Mz, 0, +1);
    fft_lines(c, w, &c->ay, Mz, c->Nx * Mz, Mz, c->Ny * Mz, +1);
    #pragma omp parallel
    {
        cplx *line = malloc(Nz * sizeof *line);
        cplx *scr = malloc(Nz * sizeof *scr);
        #pragma omp for
        for (size_t r = 0; r < rows; ++r) {
            for (size_t k = 0; k < Mz; ++k) line[k] = w[r * Mz + k];
            for (size_t k = Mz; k < Nz; ++k) line[k] = conj(w[r * Mz + Nz - k]);
            fft1d(&c->az, line, +1, scr);
            for (size_t z = 0; z < Nz; ++z) out[r * Nz + z] = creal(line[z]);
        }
        free(line);
        free(scr);
    }
}

void orc_fft_r2c(orc_ctx *c, const double *in, double *out)
{
    fft_r2c(c, in, (cplx *)out);
}

void orc_fft_c2r(orc_ctx *c, const double *in, double *out)
{
    fft_c2r(c, (const cplx *)in, out);
}

/* ------------------------------------------------------------------------ */
/* setup: sizes (src/setup.c:192-278), k grid (:609-666), filter (:710-754) */
/* ------------------------------------------------------------------------ */

/* mk_k_grid, src/setup.c:609-666 (3D branch) */
static void mk_k_grid(orc_ctx *c)
{
    const size_t Nx = c->Nx, Ny = c->Ny, Nz = c->Nz;
    const size_t Mx = c->Nx, My = c->Ny, Mz = c->Mz;
    #pragma omp parallel for
    for (size_t i = 0; i < Mx; ++i) {
        size_t osx = i * My * Mz;
        for (size_t j = 0; j < My; ++j) {
            size_t osy = osx + j * Mz;
            for (size_t k = 0; k < Mz; ++k) {
                size_t id = osy + k;
                double k2 = c->kz2 * k * k;
                if (i > Nx / 2) {
                    c->kx[id] = c->kx1 * ((int)i - (int)Nx);
                    c->kxf[id] = c->kx[id];
                    k2 += c->kx2 * (Nx - i) * (Nx - i);
                } else if (2 * i == Nx) {
                    c->kx[id] = 0.0;
                    c->kxf[id] = c->kx1 * i;
                    k2 += c->kx2 * i * i;
                } else {
                    c->kx[id] = c->kx1 * i;
                    c->kxf[id] = c->kx[id];
                    k2 += c->kx2 * i * i;
                }
                if (j > Ny / 2) {
                    c->ky[id] = c->ky1 * ((int)j - (int)Ny);
                    c->kyf[id] = c->ky[id];
                    k2 += c->ky2 * (Ny - j) * (Ny - j);
                } else if (2 * j == Ny) {
                    c->ky[id] = 0.0;
                    c->kyf[id] = c->ky1 * j;
                    k2 += c->ky2 * j * j;
                } else {
                    c->ky[id] = c->ky1 * j;
                    c->kyf[id] = c->ky[id];
                    k2 += c->ky2 * j * j;
                }
                if (2 * k == Nz) {
                    c->kz[id] = 0.0;
                    c->kzf[id] = c->kz1 * k;
                } else {
                    c->kz[id] = c->kz1 * k;
                    c->kzf[id] = c->kz[id];
                }
                c->ksq[id] = k2;
            }
        }
    }
    c->k2_max = c->kx2 * (Nx / 2) * (Nx / 2) + c->ky2 * (Ny / 2) * (Ny / 2) +
                c->kz2 * (Nz / 2) * (Nz / 2);
    c->k_max = sqrt(c->k2_max);
    c->k_lo = fmin(c->kx1, fmin(c->ky1, c->kz1));
    c->log_span = log(c->k_max / c->k_lo);
}

/* Spectrum bin of a mode with |k| = k: the reference's linear bins
 * (src/toolbox.c:579, :477), or, with log_bins, bins of equal width in
 * log |k| from the fundamental |k| (k_lo, the smallest nonzero |k|) to
 * k_max, index trunc(bins log(k / k_lo) / log(k_max / k_lo) - 1e-14). */
static size_t bin_index(const orc_ctx *c, double k)
{
    const size_t bins = c->p.bins;
    const double x = c->p.log_bins ? bins * log(k / c->k_lo) / c->log_span
                                   : bins * k / c->k_max;
    long b = (long)trunc(x - 1.0e-14);
    if (b < 0) b = 0;
    if (b >= (long)bins) b = (long)bins - 1;
    return (size_t)b;
}

/* filter_window, src/setup.c:747-754 */
static double filter_window(const double x)
{
    return exp(-36.0 * pow(x, 36));
}

/* mk_filter_mask, src/setup.c:710-727 (3D) */
static void mk_filter_mask(orc_ctx *c)
{
    const double kxmax = (c->Nx / 2) * c->kx1;
    const double kymax = (c->Ny / 2) * c->ky1;
    const double kzmax = (c->Nz / 2) * c->kz1;
    #pragma omp parallel for
    for (size_t i = 0; i < c->M; ++i) {
        double frac = fabs(c->kxf[i]) / kxmax;
        frac = ORC_MAX(frac, fabs(c->kyf[i]) / kymax);
        frac = ORC_MAX(frac, fabs(c->kzf[i]) / kzmax);
        c->filt[i] = filter_window(frac);
    }
}

static void *zalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}

orc_ctx *orc_create(const orc_params *p)
{
    if (p->nx < 2 || p->ny < 2 || p->nz < 2) return NULL;
    orc_ctx *c = zalloc(1, sizeof *c);
    c->p = *p;
#ifdef _OPENMP
    if (p->threads > 0) omp_set_num_threads(p->threads);
#endif
    /* init_grid_pars, src/setup.c:192-278 */
    c->Nx = p->nx;
    c->Ny = p->ny;
    c->Nz = p->nz;
    c->kx1 = ORC_TWOPI / (p->xb - p->xa);
    c->kx2 = ORC_TWOPI * ORC_TWOPI / ((p->xb - p->xa) * (p->xb - p->xa));
    c->ky1 = ORC_TWOPI / (p->yb - p->ya);
    c->ky2 = ORC_TWOPI * ORC_TWOPI / ((p->yb - p->ya) * (p->yb - p->ya));
    c->kz1 = ORC_TWOPI / (p->zb - p->za);
    c->kz2 = ORC_TWOPI * ORC_TWOPI / ((p->zb - p->za) * (p->zb - p->za));
    c->N = c->Nx * c->Ny * c->Nz;
    c->Mz = c->Nz / 2 + 1;
    c->M = c->Nx * c->Ny * c->Mz;
    c->Next = 2 * c->M;
    c->Ntot = 4 * c->N + 4 * c->Next + 1;
    const size_t N = c->N, M = c->M;
    double **kk[] = {&c->kx, &c->ky, &c->kz, &c->kxf, &c->kyf, &c->kzf,
                     &c->ksq, &c->filt};
    for (size_t i = 0; i < sizeof kk / sizeof kk[0]; ++i)
        *kk[i] = zalloc(M, sizeof(double));
    cplx **cc[] = {&c->phic, &c->xphic, &c->yphic, &c->zphic, &c->psic,
                   &c->fc, &c->deltarhoc, &c->work};
    for (size_t i = 0; i < sizeof cc / sizeof cc[0]; ++i)
        *cc[i] = zalloc(M, sizeof(cplx));
    double **rr[] = {&c->xphi, &c->yphi, &c->zphi, &c->grad, &c->lap, &c->f,
                     &c->deltarho, &c->rho, &c->pressure, &c->s_tmp};
    for (size_t i = 0; i < sizeof rr / sizeof rr[0]; ++i)
        *rr[i] = zalloc(N, sizeof(double));
    for (int i = 0; i < 6; ++i) c->stt[i] = zalloc(M, sizeof(cplx));
    const size_t bins = p->bins > 0 ? (size_t)p->bins : 1;
    c->phi_ps = zalloc(bins, sizeof(double));
    c->psi_ps = zalloc(bins, sizeof(double));
    c->rho_ps = zalloc(bins, sizeof(double));
    c->gw = zalloc(bins, sizeof(double));
    axis_init(&c->ax, c->Nx);
    axis_init(&c->ay, c->Ny);
    axis_init(&c->az, c->Nz);
    mk_k_grid(c);
    mk_filter_mask(c);
    return c;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    double *r[] = {c->kx, c->ky, c->kz, c->kxf, c->kyf, c->kzf, c->ksq,
                   c->filt, c->xphi, c->yphi, c->zphi, c->grad, c->lap, c->f,
                   c->deltarho, c->rho, c->pressure, c->s_tmp, c->phi_ps,
                   c->psi_ps, c->rho_ps, c->gw};
    for (size_t i = 0; i < sizeof r / sizeof r[0]; ++i) free(r[i]);
    cplx *q[] = {c->phic, c->xphic, c->yphic, c->zphic, c->psic, c->fc,
                 c->deltarhoc, c->work};
    for (size_t i = 0; i < sizeof q / sizeof q[0]; ++i) free(q[i]);
    for (int i = 0; i < 6; ++i) free(c->stt[i]);
    free(c->ax.tw);
    free(c->ay.tw);
    free(c->az.tw);
    free(c);
}

void orc_sizes(const orc_ctx *c, size_t out[4])
{
    out[0] = c->N;
    out[1] = c->M;
    out[2] = c->Next;
    out[3] = c->Ntot;
}

int orc_threads(const orc_ctx *c)
{
    (void)c;
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_kgrid(const orc_ctx *c, double *kx, double *ky, double *kz,
               double *kxf, double *kyf, double *kzf, double *ksq,
               double *kmax)
{
    const size_t b = c->M * sizeof(double);
    memcpy(kx, c->kx, b);
    memcpy(ky, c->ky, b);
    memcpy(kz, c->kz, b);
    memcpy(kxf, c->kxf, b);
    memcpy(kyf, c->kyf, b);
    memcpy(kzf, c->kzf, b);
    memcpy(ksq, c->ksq, b);
    *kmax = c->k_max;
}

void orc_filter_mask(const orc_ctx *c, double *out)
{
    memcpy(out, c->filt, c->M * sizeof(double));
}

void orc_set_field(orc_ctx *c, double *field) { c->field = field; }

/* ------------------------------------------------------------------------ */
/* reductions (src/toolbox.c:676-754)                                       */
/* ------------------------------------------------------------------------ */

/* mean, src/toolbox.c:726-734 */
double orc_mean(const double *f, size_t N)
{
    double mean = 0.0;
    #pragma omp parallel for reduction(+: mean)
    for (size_t i = 0; i < N; ++i) mean += f[i];
    return mean / N;
}

/* variance, src/toolbox.c:744-754 */
double orc_variance(double mean, const double *f, size_t N)
{
    double sum1 = 0.0, sum2 = 0.0;
    #pragma omp parallel for reduction(+: sum1, sum2)
    for (size_t i = 0; i < N; ++i) {
        double tmp = f[i] - mean;
        sum1 += tmp * tmp;
        sum2 += tmp;
    }
    return (sum1 - sum2 * sum2 / N) / (N - 1);
}

/* mean_var_min_max, src/toolbox.c:676-690 */
static void mean_var_min_max(const double *f, double *smry, size_t N)
{
    double mean = 0.0, min_val = f[0], max_val = f[0];
    #pragma omp parallel for reduction(+: mean) reduction(max: max_val) \
        reduction(min: min_val)
    for (size_t i = 0; i < N; ++i) {
        mean += f[i];
        min_val = ORC_MIN(min_val, f[i]);
        max_val = ORC_MAX(max_val, f[i]);
    }
    smry[0] = mean / N;
    smry[1] = orc_variance(smry[0], f, N);
    smry[2] = min_val;
    smry[3] = max_val;
}

/* ------------------------------------------------------------------------ */
/* right-hand side (src/toolbox.c:75-626)                                   */
/* ------------------------------------------------------------------------ */

/* potential / potential_prime, src/toolbox.c:493-510 */
static double potential(const orc_ctx *c, const double f)
{
    return c->p.mass * c->p.mass * f * f / 2.0;
}

static double potential_prime(const orc_ctx *c, const double f)
{
    return c->p.mass * c->p.mass * f;
}

/* capply_filter, src/toolbox.c:616-626 */
static void capply_filter(orc_ctx *c, cplx *in, double *out)
{
    #pragma omp parallel for
    for (size_t i = 0; i < c->M; ++i) {
        in[i] *= c->filt[i];
        c->deltarhoc[i] = in[i] / c->N;
    }
    fft_c2r(c, c->deltarhoc, out);
}

/* apply_filter, src/toolbox.c:596-602 */
static void apply_filter(orc_ctx *c, double *f)
{
    fft_r2c(c, f, c->fc);
    capply_filter(c, c->fc, f);
}

/* mk_ffts_and_filter, src/toolbox.c:160-186 (filter + constraints on).
 * Every call site of the reference sets filter whenever it sets output
 * ("output == 1 => filter == 1", :165), so psi_hat is always fresh on an
 * output call there.  For the combination output without filter, which the
 * reference never issues and where it would read a stale tmp.psic, psi_hat
 * is recomputed here, as on the device. */
static void mk_ffts_and_filter(orc_ctx *c, double *f, int filter, int output)
{
    const size_t N = c->N;
    fft_r2c(c, f, c->phic);
    if (c->p.enable_filter && filter == 1) {
        fft_r2c(c, f + 2 * N, c->psic);
        capply_filter(c, c->phic, f);
        capply_filter(c, c->psic, f + 2 * N);
        apply_filter(c, f + N);
        apply_filter(c, f + 3 * N);
    } else if (output == 1) {
        fft_r2c(c, f + 2 * N, c->psic);
    }
    if (output == 1) {
        #pragma omp parallel for
        for (size_t i = 0; i < c->M; ++i) c->fc[i] = -c->psic[i] * c->ksq[i] / N;
        fft_c2r(c, c->fc, c->f);
    }
}

/* assemble_gradient_squared, src/toolbox.c:196-208 (3D) */
static void assemble_gradient_squared(orc_ctx *c)
{
    #pragma omp parallel for
    for (size_t i = 0; i < c->N; ++i) {
        c->grad[i] = c->xphi[i] * c->xphi[i];
        c->grad[i] += c->yphi[i] * c->yphi[i];
        c->grad[i] += c->zphi[i] * c->zphi[i];
    }
}

/* mk_gradient_squared_and_laplacian, src/toolbox.c:129-149 */
static void mk_gradient_squared_and_laplacian(orc_ctx *c, double *in,
                                              int filter, int output)
{
    mk_ffts_and_filter(c, in, filter, output);
    #pragma omp parallel for
    for (size_t i = 0; i < c->M; ++i) {
        cplx pre = c->phic[i] / c->N;
        c->xphic[i] = pre * I * c->kx[i];
        c->yphic[i] = pre * I * c->ky[i];
        c->zphic[i] = pre * I * c->kz[i];
        c->fc[i] = -pre * c->ksq[i];
    }
    fft_c2r(c, c->xphic, c->xphi);
    fft_c2r(c, c->yphic, c->yphi);
    fft_c2r(c, c->zphic, c->zphi);
    fft_c2r(c, c->fc, c->lap);
    assemble_gradient_squared(c);
}

/* mk_rho_and_p, src/toolbox.c:225-243 */
static void mk_rho_and_p(orc_ctx *c, const double *f)
{
    const size_t N = c->N;
    const double a2 = f[c->Ntot - 1] * f[c->Ntot - 1];
    double rho_mean = 0.0, pressure_mean = 0.0;
    #pragma omp parallel for reduction(+: rho_mean, pressure_mean)
    for (size_t i = 0; i < N; ++i) {
        double dph = f[N + i], ps = f[2 * N + i];
        double t1 = (0.5 - ps) * dph * dph;
        double t2 = (0.5 + ps) * c->grad[i] / a2;
        c->rho[i] = t1 + t2 + potential(c, f[i]);
        c->pressure[i] = t1 - t2 / 3.0 - potential(c, f[i]);
        pressure_mean += c->pressure[i];
        rho_mean += c->rho[i];
    }
    c->rho_mean = rho_mean / N;
    c->pressure_mean = pressure_mean / N;
}

/* mk_power_spectrum, src/toolbox.c:567-582 */
static void mk_power_spectrum(const orc_ctx *c, const cplx *in, double *out)
{
    const size_t bins = c->p.bins;
    for (size_t i = 0; i < bins; ++i) out[i] = 0.0;
    for (size_t i = 1; i < c->M; ++i) {
        double pow2_tmp = creal(in[i] * conj(in[i]));
        if (fabs(c->kz[i]) > DBL_EPSILON) pow2_tmp *= 2.0;
        out[bin_index(c, sqrt(c->ksq[i]))] += pow2_tmp / c->N;
    }
}

/* mk_constraints, src/toolbox.c:524-551 */
static void mk_constraints(orc_ctx *c, const double *f)
{
    const size_t N = c->N;
    const double a2 = f[c->Ntot - 1] * f[c->Ntot - 1];
    const double hubble = sqrt(c->rho_mean / 3.0);
    const double h3 = 3.0 * hubble;
    const double phi_mean = orc_mean(f, N);
    const double dphi_mean = orc_mean(f + N, N);
    double ham_l2 = 0.0, ham_max = 0.0, mom_l2 = 0.0, mom_max = 0.0;
    #pragma omp parallel for reduction(max: ham_max, mom_max) \
        reduction(+: ham_l2, mom_l2)
    for (size_t i = 0; i < N; ++i) {
        double tmp1 = hubble * f[2 * N + i] + f[3 * N + i];
        double ham = c->f[i] / a2 - h3 * tmp1 - 0.5 * (c->rho[i] - c->rho_mean);
        double mom = tmp1 - 0.5 * dphi_mean * (f[i] - phi_mean);
        ham_l2 += ham * ham;
        mom_l2 += mom * mom;
        ham_max = ORC_MAX(ham_max, fabs(ham));
        mom_max = ORC_MAX(mom_max, fabs(mom));
    }
    c->cstr[0] = ham_l2 / N;
    c->cstr[1] = ham_max;
    c->cstr[2] = mom_l2 / N;
    c->cstr[3] = mom_max;
}

/* mk_summary, src/toolbox.c:634-663: reads the global `field`, not `f` */
static void mk_summary(orc_ctx *c)
{
    const size_t N = c->N;
    const double *field = c->field;
    mean_var_min_max(field, c->smry[0], N);
    mean_var_min_max(field + N, c->smry[1], N);
    mean_var_min_max(field + 2 * N, c->smry[2], N);
    mean_var_min_max(field + 3 * N, c->smry[3], N);
    mean_var_min_max(c->rho, c->smry[4], N);
    mean_var_min_max(c->pressure, c->smry[5], N);
}

/* mk_gw_spectrum, src/toolbox.c:435-484 */
static void mk_gw_spectrum(orc_ctx *c, const double *f)
{
    const size_t Ndh1 = 4 * c->N + 2 * c->Next;
    const size_t Ndh2 = Ndh1 + c->Next;
    const double lx = c->p.xb - c->p.xa;
    const double ly = c->p.yb - c->p.ya;
    const double lz = c->p.zb - c->p.za;
    const double hubble = sqrt(c->rho_mean / 3.0);
    const double rat = pow(0.01, 1.0 / 3.0);
    const double fac = ORC_PI * rat / (3.0 * hubble * hubble * lx * ly * lz);
    const size_t bins = c->p.bins;
    for (size_t i = 0; i < bins; ++i) c->gw[i] = 0.0;
    for (size_t i = 1; i < c->M; ++i) {
        double kx = c->kxf[i], ky = c->kyf[i], kz = c->kzf[i];
        double kx2 = kx * kx, ky2 = ky * ky, kz2 = kz * kz;
        double k2 = c->ksq[i], k = sqrt(k2);
        double dh1r = f[Ndh1 + 2 * i], dh2r = f[Ndh2 + 2 * i];
        double dh1i = f[Ndh1 + 2 * i + 1], dh2i = f[Ndh2 + 2 * i + 1];
        double pw;
        if (fabs(kz) > DBL_EPSILON) {
            pw = 2.0 * k2 / (kz2 * (ky2 + kz2)) *
                 ((kx2 + kz2) * (dh1r * dh1r + dh1i * dh1i) +
                  2.0 * kx * ky * (dh1r * dh2r + dh1i * dh2i) +
                  (ky2 + kz2) * (dh2r * dh2r + dh2i * dh2i));
        } else if (fabs(ky) > DBL_EPSILON) {
            pw = 2.0 * (kx2 + ky2) / (ky2 * ky2) *
                 ((kx2 + ky2) * (dh1r * dh1r + dh1i * dh1i) +
                  ky2 * (dh2r * dh2r + dh2i * dh2i));
        } else {
            pw = 2.0 * (dh1r * dh1r + dh1i * dh1i + dh2r * dh2r + dh2i * dh2i);
        }
        if (fabs(c->kz[i]) > DBL_EPSILON) pw *= 2.0;
        c->gw[bin_index(c, k)] += fac * k2 * k * pw;
    }
    for (size_t i = 0; i < bins; ++i) c->gw[i] /= c->N;
}

/* output_all, src/toolbox.c:251-270 (all outputs on) */
static void output_all(orc_ctx *c, const double *f)
{
    mk_power_spectrum(c, c->phic, c->phi_ps);
    mk_power_spectrum(c, c->psic, c->psi_ps);
    fft_r2c(c, c->rho, c->fc);
    mk_power_spectrum(c, c->fc, c->rho_ps);
    mk_constraints(c, f);
    mk_summary(c);
    if (c->p.enable_gw) mk_gw_spectrum(c, f);
}

/* update_phi_psi, src/toolbox.c:282-299 */
static void update_phi_psi(orc_ctx *c, const double *f, double *result)
{
    const size_t N = c->N, N2 = 2 * N, N3 = 3 * N;
    const double a2 = f[c->Ntot - 1] * f[c->Ntot - 1];
    const double hubble = sqrt(c->rho_mean / 3.0);
    const double h3 = 3.0 * hubble;
    const double pm = c->pressure_mean;
    #pragma omp parallel for
    for (size_t i = 0; i < N; ++i) {
        double dph = f[N + i], ps = f[N2 + i], dps = f[N3 + i];
        result[i] = dph;
        result[N + i] = (1.0 + 4.0 * ps) * c->lap[i] / a2 -
                        (h3 - 4.0 * dps) * dph -
                        (1.0 + 2.0 * ps) * potential_prime(c, f[i]);
        result[N2 + i] = dps;
        result[N3 + i] = 0.5 * c->pressure[i] + (ps - 0.5) * pm -
                         4.0 * hubble * dps;
    }
}

/* mk_gw_sources, src/toolbox.c:410-424 */
static void mk_gw_sources(orc_ctx *c, cplx **s)
{
    double *t1[] = {c->xphi, c->xphi, c->xphi, c->yphi, c->yphi, c->zphi};
    double *t2[] = {c->xphi, c->yphi, c->zphi, c->yphi, c->zphi, c->zphi};
    for (size_t j = 0; j < 6; ++j) {
        #pragma omp parallel for
        for (size_t i = 0; i < c->N; ++i) c->s_tmp[i] = t1[j][i] * t2[j][i];
        fft_r2c(c, c->s_tmp, s[j]);
    }
}

/* mk_gw_sources_tt, src/toolbox.c:363-397 */
static void mk_gw_sources_tt(orc_ctx *c, cplx **s)
{
    mk_gw_sources(c, s);
    #pragma omp parallel for
    for (size_t i = 1; i < c->M; ++i) {
        double kx = c->kxf[i], ky = c->kyf[i], kz = c->kzf[i];
        double ksq = c->ksq[i];
        double fx = kx / ksq, fy = ky / ksq, fz = kz / ksq;
        cplx t1 = kx * fx * s[0][i] + ky * fy * s[3][i] + kz * fz * s[5][i] +
                  2.0 * (kx * fy * s[1][i] + kx * fz * s[2][i] +
                         ky * fz * s[4][i]);
        cplx t2 = s[0][i] + s[3][i] + s[5][i];
        cplx s1 = t1 + t2, s2 = t1 - t2;
        if (fabs(kz) > DBL_EPSILON) {
            cplx k1 = kx * s[0][i] + ky * s[1][i] + kz * s[2][i];
            cplx k2 = kx * s[1][i] + ky * s[3][i] + kz * s[4][i];
            s[0][i] = s[0][i] - 2.0 * fx * k1 + 0.5 * (fx * kx * s1 + s2);
            s[1][i] = s[1][i] - fx * k2 - fy * k1 + 0.5 * fx * ky * s1;
        } else if (fabs(ky) > DBL_EPSILON) {
            cplx k1 = kx * s[0][i] + ky * s[1][i] + kz * s[2][i];
            cplx k3 = kx * s[2][i] + ky * s[4][i] + kz * s[5][i];
            s[0][i] = s[0][i] - 2.0 * fx * k1 + 0.5 * (fx * kx * s1 + s2);
            s[1][i] = s[2][i] - fx * k3 - fz * k1 + 0.5 * fx * kz * s1;
        } else {
            cplx k2 = kx * s[1][i] + ky * s[3][i] + kz * s[4][i];
            cplx k3 = kx * s[2][i] + ky * s[4][i] + kz * s[5][i];
            s[0][i] = s[3][i] - 2.0 * fy * k2 + 0.5 * (fy * ky * s1 + s2);
            s[1][i] = s[4][i] - fy * k3 - fz * k2 + 0.5 * fy * kz * s1;
        }
    }
    s[0][0] = 0.0;
    s[1][0] = 0.0;
}

/* update_h, src/toolbox.c:311-346 */
static void update_h(orc_ctx *c, const double *f, double *result)
{
    const size_t N = c->N, Next = c->Next, Nh1 = 4 * N, Nh2 = Nh1 + Next;
    const size_t Ndh1 = Nh2 + Next, Ndh2 = Ndh1 + Next;
    const double a2 = f[c->Ntot - 1] * f[c->Ntot - 1];
    const double hubble = sqrt(c->rho_mean / 3.0);
    const double h3 = 3.0 * hubble;
    const double pm = c->pressure_mean;
    #pragma omp parallel for
    for (size_t i = 0; i < 2 * Next; ++i) result[Nh1 + i] = f[Ndh1 + i];
    cplx **stt = c->stt;
    mk_gw_sources_tt(c, stt);
    #pragma omp parallel for
    for (size_t i = 0; i < c->M; ++i) {
        size_t i1 = 2 * i, i2 = i1 + 1;
        double tmp = -(2.0 * pm + c->ksq[i] / a2);
        result[Ndh1 + i1] = tmp * f[Nh1 + i1] - h3 * f[Ndh1 + i1] +
                            2.0 * creal(stt[0][i]) / a2;
        result[Ndh1 + i2] = tmp * f[Nh1 + i2] - h3 * f[Ndh1 + i2] +
                            2.0 * cimag(stt[0][i]) / a2;
        result[Ndh2 + i1] = tmp * f[Nh2 + i1] - h3 * f[Ndh2 + i1] +
                            2.0 * creal(stt[1][i]) / a2;
        result[Ndh2 + i2] = tmp * f[Nh2 + i2] - h3 * f[Ndh2 + i2] +
                            2.0 * cimag(stt[1][i]) / a2;
    }
}

/* mk_rhs, src/toolbox.c:75-88 */
void orc_mk_rhs(orc_ctx *c, double t, double *f, double *result, int filter,
                int output)
{
    (void)t;
    c->calls_rhs += 1;
    mk_gradient_squared_and_laplacian(c, f, filter, output);
    mk_rho_and_p(c, f);
    if (output == 1) output_all(c, f);
    update_phi_psi(c, f, result);
    if (c->p.enable_gw) update_h(c, f, result);
    result[c->Ntot - 1] = f[c->Ntot - 1] * sqrt(c->rho_mean / 3.0);
}

void orc_means(const orc_ctx *c, double out[2])
{
    out[0] = c->rho_mean;
    out[1] = c->pressure_mean;
}

void orc_obs(const orc_ctx *c, double *out)
{
    const size_t bins = c->p.bins;
    memcpy(out, c->smry, 24 * sizeof(double));
    memcpy(out + 24, c->cstr, 4 * sizeof(double));
    memcpy(out + 28, c->phi_ps, bins * sizeof(double));
    memcpy(out + 28 + bins, c->psi_ps, bins * sizeof(double));
    memcpy(out + 28 + 2 * bins, c->rho_ps, bins * sizeof(double));
    memcpy(out + 28 + 3 * bins, c->gw, bins * sizeof(double));
}

unsigned long orc_calls_rhs(const orc_ctx *c) { return c->calls_rhs; }

/* intermediate arrays of the last mk_rhs, for tests:
 * 0 xphi, 1 yphi, 2 zphi, 3 lap, 4 grad, 5 rho, 6 pressure, 7 tmp.f (N each);
 * 8 stt[0], 9 stt[1] (2M each) */
void orc_aux(const orc_ctx *c, int which, double *out)
{
    const double *src[] = {c->xphi, c->yphi, c->zphi, c->lap, c->grad,
                           c->rho, c->pressure, c->f};
    if (which >= 0 && which < 8)
        memcpy(out, src[which], c->N * sizeof(double));
    else if (which == 8 || which == 9)
        memcpy(out, c->stt[which - 8], c->M * sizeof(cplx));
}

/* mk_initial_psi, src/setup.c:825-903 (the "new routine") */
void orc_mk_initial_psi(orc_ctx *c, double ti, double *field, double *dfield)
{
    const size_t N = c->N;
    for (size_t i = 2 * N; i < 4 * N; ++i) field[i] = 0.0;
    c->field = field;
    orc_mk_rhs(c, ti, field, dfield, 0, 0);
    const double a2 = field[c->Ntot - 1] * field[c->Ntot - 1];
    const double hubble = sqrt(c->rho_mean / 3.0);
    const double phi_mean = orc_mean(field, N);
    const double dphi_mean = orc_mean(field + N, N);
    const double ddphi_mean = orc_mean(dfield + N, N);
    #pragma omp parallel for
    for (size_t i = 0; i < N; ++i) {
        c->deltarho[i] = field[N + i] - dphi_mean;
        c->f[i] = field[i] - phi_mean;
    }
    fft_r2c(c, c->deltarho, c->deltarhoc);
    fft_r2c(c, c->f, c->fc);
    #pragma omp parallel for
    for (size_t i = 1; i < c->M; ++i) {
        c->phic[i] = (-dphi_mean * c->deltarhoc[i] + ddphi_mean * c->fc[i]) /
                     ((-dphi_mean * dphi_mean + 2.0 * c->ksq[i] / a2) * N);
    }
    c->phic[0] = 0.0;
    fft_c2r(c, c->phic, field + 2 * N);
    #pragma omp parallel for
    for (size_t i = 0; i < N; ++i)
        field[3 * N + i] = 0.5 * dphi_mean * c->f[i] - hubble * field[2 * N + i];
}

/* ------------------------------------------------------------------------ */
/* RK4, src/rk4.c:27-132                                                    */
/* ------------------------------------------------------------------------ */

static void do_save(orc_ctx *c, orc_save_fn save, void *user, double t,
                    const double *field, double *obsbuf)
{
    if (!save) return;
    orc_obs(c, obsbuf);
    save(user, t, field[c->Ntot - 1], obsbuf, ORC_OBS_LEN(c->p.bins));
}

long orc_run_rk4(orc_ctx *c, double *field, double ti, double tf, double dt,
                 long max_steps, long skip, orc_save_fn save, void *user)
{
    const size_t Ntot = c->Ntot;
    double *k1 = zalloc(Ntot, sizeof(double));
    double *k2 = zalloc(Ntot, sizeof(double));
    double *k3 = zalloc(Ntot, sizeof(double));
    double *k4 = zalloc(Ntot, sizeof(double));
    double *tmp_k = zalloc(Ntot, sizeof(double));
    double *obs = zalloc(ORC_OBS_LEN(c->p.bins), sizeof(double));
    double t = ti;
    long nt;
    c->field = field;
    for (nt = 0; t < tf && nt < max_steps; ++nt) {
        if (t + dt * 1.0001 > tf) dt = tf - t;
        int filter = c->p.enable_filter ? 1 : 0;
        if (nt % skip == 0) {
            orc_mk_rhs(c, t, field, k1, filter, 1);
            do_save(c, save, user, t, field, obs);
        } else {
            orc_mk_rhs(c, t, field, k1, filter, 0);
        }
        #pragma omp parallel for
        for (size_t i = 0; i < Ntot; ++i) tmp_k[i] = field[i] + dt * k1[i] / 2.0;
        orc_mk_rhs(c, t + dt / 2.0, tmp_k, k2, 0, 0);
        #pragma omp parallel for
        for (size_t i = 0; i < Ntot; ++i) tmp_k[i] = field[i] + dt * k2[i] / 2.0;
        orc_mk_rhs(c, t + dt / 2.0, tmp_k, k3, 0, 0);
        #pragma omp parallel for
        for (size_t i = 0; i < Ntot; ++i) tmp_k[i] = field[i] + dt * k3[i];
        orc_mk_rhs(c, t + dt, tmp_k, k4, 0, 0);
        #pragma omp parallel for
        for (size_t i = 0; i < Ntot; ++i)
            field[i] += dt * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]) / 6.0;
        t += dt;
    }
    /* prepare_and_save_timeslice, src/toolbox.c:96-104 */
    orc_mk_rhs(c, t, field, k1, 1, 1);
    do_save(c, save, user, t, field, obs);
    free(k1);
    free(k2);
    free(k3);
    free(k4);
    free(tmp_k);
    free(obs);
    return nt;
}

/* ------------------------------------------------------------------------ */
/* DOPRI853, src/dopri853.c + src/dopri853_constants.c                      */
/* ------------------------------------------------------------------------ */

/* Dormand-Prince 8(5,3) coefficients (Hairer, Norsett & Wanner, "Solving
 * Ordinary Differential Equations I", DOP853; the same published table as
 * src/dopri853_constants.c:13-121).  Stage s (0-based) combines stages
 * j < s with A[s][j]; stage 11 is evaluated at t + dt. */
static const double A[12][12] = {
    {0},
    {5.26001519587677318785587544488e-2},
    {1.97250569845378994544595329183e-2, 5.91751709536136983633785987549e-2},
    {2.95875854768068491816892993775e-2, 0.0,
     8.87627564304205475450678981324e-2},
    {2.41365134159266685502369798665e-1, 0.0,
     -8.84549479328286085344864962717e-1, 9.24834003261792003115737966543e-1},
    {3.7037037037037037037037037037e-2, 0.0, 0.0,
     1.70828608729473871279604482173e-1, 1.25467687566822425016691814123e-1},
    {3.7109375e-2, 0.0, 0.0, 1.70252211019544039314978060272e-1,
     6.02165389804559606850219397283e-2, -1.7578125e-2},
    {3.70920001185047927108779319836e-2, 0.0, 0.0,
     1.70383925712239993810214054705e-1, 1.07262030446373284651809199168e-1,
     -1.53194377486244017527936158236e-2, 8.27378916381402288758473766002e-3},
    {6.24110958716075717114429577812e-1, 0.0, 0.0,
     -3.36089262944694129406857109825e0, -8.68219346841726006818189891453e-1,
     2.75920996994467083049415600797e1, 2.01540675504778934086186788979e1,
     -4.34898841810699588477366255144e1},
    {4.77662536438264365890433908527e-1, 0.0, 0.0,
     -2.48811461997166764192642586468e0, -5.90290826836842996371446475743e-1,
     2.12300514481811942347288949897e1, 1.52792336328824235832596922938e1,
     -3.32882109689848629194453265587e1, -2.03312017085086261358222928593e-2},
    {-9.3714243008598732571704021658e-1, 0.0, 0.0,
     5.18637242884406370830023853209e0, 1.09143734899672957818500254654e0,
     -8.14978701074692612513997267357e0, -1.85200656599969598641566180701e1,
     2.27394870993505042818970056734e1, 2.49360555267965238987089396762e0,
     -3.0467644718982195003823669022e0},
    {2.27331014751653820792359768449e0, 0.0, 0.0,
     -1.05344954667372501984066689879e1, -2.00087205822486249909675718444e0,
     -1.79589318631187989172765950534e1, 2.79488845294199600508499808837e1,
     -2.85899827713502369474065508674e0, -8.87285693353062954433549289258e0,
     1.23605671757943030647266201528e1, 6.43392746015763530355970484046e-1},
};
static const double C[12] = {
    0.0, 0.526001519587677318785587544488e-01,
    0.789002279381515978178381316732e-01, 0.118350341907227396726757197510e+00,
    0.281649658092772603273242802490e+00, 0.333333333333333333333333333333e+00,
    0.25e+00, 0.307692307692307692307692307692e+00,
    0.651282051282051282051282051282e+00, 0.6e+00,
    0.857142857142857142857142857142e+00, 1.0};
static const double Bw[12] = {
    5.42937341165687622380535766363e-2, 0.0, 0.0, 0.0, 0.0,
    4.45031289275240888144113950566e0, 1.89151789931450038304281599044e0,
    -5.8012039600105847814672114227e0, 3.1116436695781989440891606237e-1,
    -1.52160949662516078556178806805e-1, 2.01365400804030348374776537501e-1,
    4.47106157277725905176885569043e-2};
static const double BHH[3] = {0.244094488188976377952755905512e+00,
                              0.733846688281611857341361741547e+00,
                              0.220588235294117647058823529412e-01};
static const double ER[12] = {
    0.1312004499419488073250102996e-01, 0.0, 0.0, 0.0, 0.0,
    -0.1225156446376204440720569753e+01, -0.4957589496572501915214079952e+00,
    0.1664377182454986536961530415e+01, -0.3503288487499736816886487290e+00,
    0.3341791187130174790297318841e+00, 0.8192320648511571246570742613e-01,
    -0.2235530786388629525884427845e-01};

void orc_dp_tableau(double *a, double *c, double *b, double *bhh, double *er)
{
    for (int s = 0; s < 13; ++s)
        for (int j = 0; j < 12; ++j) a[s * 12 + j] = 0.0;
    for (int s = 0; s < 12; ++s)
        for (int j = 0; j < 12; ++j) a[s * 12 + j] = A[s][j];
    for (int j = 0; j < 12; ++j) a[12 * 12 + j] = Bw[j];
    for (int s = 0; s < 12; ++s) c[s] = C[s];
    c[12] = 1.0;
    for (int j = 0; j < 12; ++j) b[j] = Bw[j];
    for (int j = 0; j < 3; ++j) bhh[j] = BHH[j];
    for (int j = 0; j < 12; ++j) er[j] = ER[j];
}

/* success, src/dopri853.c:505-540 */
int orc_success(orc_ctrl *s, const orc_dp_opts *o, double err, double *dt,
                double rho_mean)
{
    const double alpha = 1.0 / 8.0 - o->beta * 0.2;
    double scale;
    if (err <= 1.0) {
        if (err == 0.0) {
            scale = o->maxscale;
        } else {
            scale = o->safe * pow(err, -alpha) * pow(s->err_old, o->beta);
            if (scale < o->minscale) scale = o->minscale;
            if (scale > o->maxscale) scale = o->maxscale;
        }
        if (s->reject)
            s->dt_next = (*dt) * ORC_MIN(scale, 1.0);
        else
            s->dt_next = (*dt) * scale;
        if (o->max_dt_hubble_fraction > 0.0) {
            const double minstep =
                o->max_dt_hubble_fraction * sqrt(3.0 / rho_mean);
            if (s->dt_next > minstep) s->dt_next = minstep;
        }
        s->err_old = ORC_MAX(err, 1.0e-4);
        s->reject = 0;
        return 1;
    }
    scale = ORC_MAX(o->safe * pow(err, -alpha), o->minscale);
    (*dt) *= scale;
    s->reject = 1;
    return 0;
}

typedef struct dp_bufs {
    double *k[12]; /* k[0] = dfield (FSAL); k[1..11] stages 2..12 */
    double *k_tmp, *yerr, *yerr2, *field_new, *dfield_new;
} dp_bufs;

/* try_step, src/dopri853.c:331-458; stages are stored in their own arrays
 * instead of the reference's reuse of k2/k3/k4, which does not change any
 * value that is read afterwards */
static void try_step(orc_ctx *c, dp_bufs *b, const double *field, double t,
                     double dt)
{
    const size_t Ntot = c->Ntot;
    for (int s = 1; s < 12; ++s) {
        #pragma omp parallel for
        for (size_t i = 0; i < Ntot; ++i) {
            double acc = A[s][0] * b->k[0][i];
            for (int j = 1; j < s; ++j)
                if (A[s][j] != 0.0) acc += A[s][j] * b->k[j][i];
            b->k_tmp[i] = field[i] + dt * acc;
        }
        double ts = (s == 11) ? t + dt : t + C[s] * dt;
        orc_mk_rhs(c, ts, b->k_tmp, b->k[s], 0, 0);
    }
    #pragma omp parallel for
    for (size_t i = 0; i < Ntot; ++i) {
        double bs = Bw[0] * b->k[0][i];
        for (int j = 5; j < 12; ++j) bs += Bw[j] * b->k[j][i];
        b->field_new[i] = field[i] + dt * bs;
        b->yerr[i] = bs - BHH[0] * b->k[0][i] - BHH[1] * b->k[8][i] -
                     BHH[2] * b->k[11][i];
        double e2 = ER[0] * b->k[0][i];
        for (int j = 5; j < 12; ++j) e2 += ER[j] * b->k[j][i];
        b->yerr2[i] = e2;
    }
}

/* error, src/dopri853.c:469-486 */
static double dp_error(const orc_ctx *c, const dp_bufs *b, const double *field,
                       const orc_dp_opts *o, double dt)
{
    double err = 0.0, err2 = 0.0;
    #pragma omp parallel for reduction(+: err, err2)
    for (size_t i = 0; i < c->Ntot; ++i) {
        double sk = o->atol +
                    o->rtol * ORC_MAX(fabs(field[i]), fabs(b->field_new[i]));
        double tmp = b->yerr[i] / sk;
        err2 += tmp * tmp;
        tmp = b->yerr2[i] / sk;
        err += tmp * tmp;
    }
    double deno = err + 0.01 * err2;
    if (deno <= 0.0) deno = 1.0;
    return dt * err * sqrt(1.0 / (c->Ntot * deno));
}

/* check_for_stiffness, src/dopri853.c:550-580 */
int orc_stiffness_update(orc_stiffness *st, double dt, double num, double den)
{
    if (den > 0.0) st->dtlamb = dt * sqrt(num / den);
    if (st->dtlamb > 6.1) {
        st->nonstiff = 0;
        st->stiff += 1;
        if (st->stiff == 15) return 1;
    } else {
        st->nonstiff += 1;
        if (st->nonstiff == 6) st->stiff = 0;
    }
    return 0;
}

long orc_run_dopri853(orc_ctx *c, double *field, double *dfield,
                      const orc_dp_opts *o, orc_dp_stats *st,
                      orc_save_fn save, void *user, double *dt_log,
                      double *err_log, int *accepted, long log_cap)
{
    const size_t Ntot = c->Ntot;
    dp_bufs b;
    b.k[0] = dfield;
    for (int s = 1; s < 12; ++s) b.k[s] = zalloc(Ntot, sizeof(double));
    b.k_tmp = zalloc(Ntot, sizeof(double));
    b.yerr = zalloc(Ntot, sizeof(double));
    b.yerr2 = zalloc(Ntot, sizeof(double));
    b.field_new = zalloc(Ntot, sizeof(double));
    b.dfield_new = zalloc(Ntot, sizeof(double));
    double *obs = zalloc(ORC_OBS_LEN(c->p.bins), sizeof(double));
    orc_ctrl ctl = {1.0e-4, o->dt, 0};
    double t = o->ti, dt = o->dt, dt_did = 0.0, last_save_t = o->ti;
    long n_stp, n_ok = 0, n_bad = 0, attempts = 0;
    int filter = c->p.enable_filter ? 1 : 0, stiff_stop = 0;
    orc_stiffness stiff = {0.0, 0, 0};
    c->field = field;

    /* prepare_and_save_timeslice, src/toolbox.c:96-104 */
    orc_mk_rhs(c, t, field, dfield, 1, 1);
    do_save(c, save, user, t, field, obs);

    for (n_stp = 0;; ++n_stp) {
        if (o->forced_dt) {
            if (n_stp >= o->n_forced) break;
            dt = o->forced_dt[n_stp];
        } else if (t + dt * 1.0001 > o->tf) {
            dt = o->tf - t;
        }
        /* perform_step, src/dopri853.c:271-311 */
        double h = dt;
        int underflow = 0;
        for (;;) {
            try_step(c, &b, field, t, h);
            double err = dp_error(c, &b, field, o, h);
            double tried = h;
            int ok = 1;
            if (!o->forced_dt) ok = orc_success(&ctl, o, err, &h, c->rho_mean);
            if (attempts < log_cap) {
                if (dt_log) dt_log[attempts] = tried;
                if (err_log) err_log[attempts] = err;
                if (accepted) accepted[attempts] = ok;
            }
            if (st) st->last_err = err;
            attempts++;
            if (ok) break;
            if (fabs(h) <= fabs(t) * DBL_EPSILON) {
                underflow = 1;
                break;
            }
        }
        if (underflow) break;
        int output = ((n_stp + 1) % o->skip == 0) ? 1 : 0;
        orc_mk_rhs(c, t + h, b.field_new, b.dfield_new, filter, output);
        /* src/dopri853.c:295-299 */
        if (o->enable_stiffness_check && ((n_ok % 10) == 0 || stiff.stiff > 0)) {
            double num = 0.0, den = 0.0;
            for (size_t i = 0; i < Ntot; ++i) {
                double tmp = b.dfield_new[i] - b.k[11][i];
                num += tmp * tmp;
                tmp = b.field_new[i] - b.k_tmp[i];
                den += tmp * tmp;
            }
            if (orc_stiffness_update(&stiff, dt, num, den)) {
                stiff_stop = 1; /* the reference exits here */
                break;
            }
        }
        memcpy(field, b.field_new, Ntot * sizeof(double));
        memcpy(dfield, b.dfield_new, Ntot * sizeof(double));
        t += (dt_did = h);
        if (dt_did == dt)
            ++n_ok;
        else
            ++n_bad;
        if (output) {
            do_save(c, save, user, t, field, obs);
            last_save_t = t;
        }
        if (t >= o->tf) break;
        if (n_stp >= o->max_steps) break;
        if (!o->forced_dt && fabs(ctl.dt_next) <= o->dt_min) break;
        if (!o->forced_dt) dt = ctl.dt_next;
    }
    /* wrap_up_dopri853, src/dopri853.c:223-228 (no file buffer here: the
     * final slice whenever the last save is older than t) */
    if (!stiff_stop && last_save_t < t) {
        orc_mk_rhs(c, t, field, dfield, 1, 1);
        do_save(c, save, user, t, field, obs);
    }
    if (st) {
        st->n_stp = n_stp;
        st->n_ok = n_ok;
        st->n_bad = n_bad;
        st->t = t;
        st->dtlamb = stiff.dtlamb;
        st->stiff = stiff_stop;
    }
    for (int s = 1; s < 12; ++s) free(b.k[s]);
    free(b.k_tmp);
    free(b.yerr);
    free(b.yerr2);
    free(b.field_new);
    free(b.dfield_new);
    free(obs);
    return attempts;
}
