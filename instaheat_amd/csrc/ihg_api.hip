// ihg_api.hip -- context set-up and the C ABI of libihg (include/ihg.h).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <cstdint>

#include "ihg_internal.h"
#include "rk4_kernels.h"
#include "mem.h"

namespace ihg {

static thread_local char g_err[512] = "";

int set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return IHG_EINVAL;
}

int hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
    return (e == hipErrorOutOfMemory) ? IHG_ENOMEM : IHG_EHIP;
}

static bool pow2_in_range(int n) { return n >= 4 && n <= 512 && !(n & (n - 1)); }

static int check_ptr(const void *p, const char *name)
{
    if (!p) return set_error("%s is NULL", name);
    if (((size_t)p) & 15) return set_error("%s is not 16-byte aligned", name);
    return 0;
}

template <class T>
static int dalloc(T **p, size_t count, hipStream_t stream)
{
    // zeroed on the context's stream and finished before returning: the
    // context's stream does not wait for the legacy null stream, so a
    // null-stream memset could still be running when the stream's next
    // upload or kernel writes the array
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) return hip_check(e, "hipMalloc");
    IHG_TRY(hip_check(hipMemsetAsync(*p, 0, count * sizeof(T), stream),
                      "hipMemsetAsync"));
    return hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// host -> device copy at set-up, on the context's stream and finished
// before returning (a null-stream copy from pageable memory may still be in
// flight when it returns, unordered with the context's stream)
static int upload_now(void *dst, const void *src, size_t bytes,
                      hipStream_t stream)
{
    IHG_TRY(hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                                     stream),
                      "hipMemcpyAsync"));
    return hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// filter_window / mk_filter_mask (src/setup.c:710-754): the window
// exp(-36 x^36) of x = max_i |k_i| / k_i^max.  It decreases with x, so the
// window of the largest axis fraction is the smallest of the three axis
// windows: F(i, j, k) = min(Fx[i], Fy[j], Fz[k]) exactly.  The three axis
// tables (Nx + Ny + Mz values) are evaluated on the host with the same libm
// calls as the reference; the kernels take the minimum per point instead of
// reading an M-value mask.
static void host_filter_axes(const ihg_ctx *c, double *out)
{
    const int Nx = c->Nx, Ny = c->Ny, Nz = c->Nz, Mz = c->Mz;
    const KGeom &kg = c->kg;
    const double kxmax = (Nx / 2) * kg.kx1;
    const double kymax = (Ny / 2) * kg.ky1;
    const double kzmax = (Nz / 2) * kg.kz1;
    for (int i = 0; i < Nx; ++i) {
        const double kxf = (i > Nx / 2) ? kg.kx1 * (i - Nx) : kg.kx1 * i;
        out[i] = exp(-36.0 * pow(fabs(kxf) / kxmax, 36));
    }
    for (int j = 0; j < Ny; ++j) {
        const double kyf = (j > Ny / 2) ? kg.ky1 * (j - Ny) : kg.ky1 * j;
        out[Nx + j] = exp(-36.0 * pow(fabs(kyf) / kymax, 36));
    }
    for (int k = 0; k < Mz; ++k) {
        const double kzf = kg.kz1 * k;
        out[Nx + Ny + k] = exp(-36.0 * pow(fabs(kzf) / kzmax, 36));
    }
}

static const char *const FAM_NAMES[F_NFAM] = {
    "z_r2c", "z_c2r", "y_pass", "x_pass", "x_fanout", "z_c2r_rho", "means",
    "update_phi_psi", "z_r2c_sources", "y_tt_update_h", "x_roundtrip",
    "observables", "dp_stage", "dp_finish", "rk4", "stream_copy", "dp_dense",
    "gsl_stage", "gsl_finish", "dp_dense_stage"};

static int prof_flush(ihg_ctx *c)
{
    if (c->prof_pending.empty()) return 0;
    IHG_TRY(hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    for (const ProfRec &r : c->prof_pending) {
        float ms = 0.0f;
        IHG_TRY(hip_check(hipEventElapsedTime(&ms, r.a, r.b),
                          "hipEventElapsedTime"));
        c->fam_ms[r.fam] += ms;
        c->fam_bytes[r.fam] += r.bytes;
        c->fam_count[r.fam] += 1;
        c->ev_pool.push_back(r.a);
        c->ev_pool.push_back(r.b);
    }
    c->prof_pending.clear();
    return 0;
}

static int pool_event(ihg_ctx *c, hipEvent_t *e)
{
    if (!c->ev_pool.empty()) {
        *e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return 0;
    }
    return hip_check(hipEventCreate(e), "hipEventCreate");
}

int prof_begin(ihg_ctx *c, int fam, double bytes)
{
    if (c->prof_pending.size() >= 65536) IHG_TRY(prof_flush(c));
    ProfRec r{fam, nullptr, nullptr, bytes};
    IHG_TRY(pool_event(c, &r.a));
    IHG_TRY(pool_event(c, &r.b));
    IHG_TRY(hip_check(hipEventRecord(r.a, c->stream), "hipEventRecord"));
    c->prof_pending.push_back(r);
    return 0;
}

int prof_end(ihg_ctx *c)
{
    return hip_check(hipEventRecord(c->prof_pending.back().b, c->stream),
                     "hipEventRecord");
}

} // namespace ihg

using namespace ihg;

extern "C" const char *ihg_last_error(void) { return g_err; }

extern "C" int ihg_prof_nfam(void) { return F_NFAM; }

extern "C" const char *ihg_prof_family(int fam)
{
    return (fam >= 0 && fam < F_NFAM) ? FAM_NAMES[fam] : "";
}

extern "C" int ihg_prof_enable(ihg_ctx *c, int on)
{
    if (!c) return set_error("NULL context");
    if (!on) IHG_TRY(prof_flush(c));
    c->prof_on = on != 0;
    return 0;
}

extern "C" int ihg_prof_reset(ihg_ctx *c)
{
    if (!c) return set_error("NULL context");
    IHG_TRY(prof_flush(c));
    for (int f = 0; f < F_NFAM; ++f) {
        c->fam_ms[f] = 0.0;
        c->fam_bytes[f] = 0.0;
        c->fam_count[f] = 0;
    }
    return 0;
}

extern "C" int ihg_prof_read(ihg_ctx *c, double *ms, double *bytes,
                             long *count, int nfam)
{
    if (!c || !ms || !bytes || !count) return set_error("NULL argument");
    IHG_TRY(prof_flush(c));
    for (int f = 0; f < nfam && f < F_NFAM; ++f) {
        ms[f] = c->fam_ms[f];
        bytes[f] = c->fam_bytes[f];
        count[f] = c->fam_count[f];
    }
    return 0;
}

extern "C" int ihg_destroy(ihg_ctx *c)
{
    if (!c) return 0;
    (void)hipStreamSynchronize(c->stream);
    for (const ProfRec &r : c->prof_pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : c->ev_pool)
        (void)hipEventDestroy(e);
    void *dev[] = {c->tw, c->twf, c->filter, c->W, c->Wc, c->grad3, c->rho,
                   c->partials, c->dscal, c->dobs, c->gather, c->fin_partials};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    for (void *p : c->user_arrays)
        (void)hipFree(p);
    if (c->hscal) (void)hipHostFree(c->hscal);
    if (c->hobs) (void)hipHostFree(c->hobs);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

static int create_ctx(const ihg_params *p, int nreal, ihg_ctx **out)
{
    if (!p || !out) return set_error("NULL argument");
    *out = nullptr;
    if (nreal < 1 || nreal > 65535)
        return set_error("ensemble of %d realizations (1..65535)", nreal);
    if (!pow2_in_range(p->nx) || !pow2_in_range(p->ny) || !pow2_in_range(p->nz))
        return set_error("grid %dx%dx%d: each size must be a power of two in "
                         "4..512", p->nx, p->ny, p->nz);
    if (!(p->xb > p->xa) || !(p->yb > p->ya) || !(p->zb > p->za))
        return set_error("invalid spatial bounds, need a < b");
    if (p->bins < 1 || p->bins > 2048)
        return set_error("POWER_SPECTRUM_BINS must be in 1..2048");
    ihg_ctx *c = new ihg_ctx();
    c->p = *p;
    c->nreal = nreal;
    c->Nx = p->nx;
    c->Ny = p->ny;
    c->Nz = p->nz;
    c->Mz = p->nz / 2 + 1;
    c->cand_x = ihg::default_candidate_x(c->Nx);
    c->N = (long)c->Nx * c->Ny * c->Nz;
    c->M = (long)c->Nx * c->Ny * c->Mz;
    c->Next = 2 * c->M;
    c->Ntot = 4 * c->N + 4 * c->Next + 1;
    // init_grid_pars, src/setup.c:192-208
    const double TWOPI = 6.283185307179586476925286766559;
    KGeom &kg = c->kg;
    kg.Nx = c->Nx;
    kg.Ny = c->Ny;
    kg.Nz = c->Nz;
    kg.Mz = c->Mz;
    kg.kx1 = TWOPI / (p->xb - p->xa);
    kg.kx2 = TWOPI * TWOPI / ((p->xb - p->xa) * (p->xb - p->xa));
    kg.ky1 = TWOPI / (p->yb - p->ya);
    kg.ky2 = TWOPI * TWOPI / ((p->yb - p->ya) * (p->yb - p->ya));
    kg.kz1 = TWOPI / (p->zb - p->za);
    kg.kz2 = TWOPI * TWOPI / ((p->zb - p->za) * (p->zb - p->za));
    // k_max, src/setup.c:660-664
    const double k2_max = kg.kx2 * (c->Nx / 2) * (c->Nx / 2) +
                          kg.ky2 * (c->Ny / 2) * (c->Ny / 2) +
                          kg.kz2 * (c->Nz / 2) * (c->Nz / 2);
    kg.k_max = sqrt(k2_max);
    kg.log_bins = p->log_bins != 0;
    kg.k_lo = fmin(kg.kx1, fmin(kg.ky1, kg.kz1));
    kg.log_span = log(kg.k_max / kg.k_lo);

    int rc = 0;
    auto fail = [&](int r) {
        ihg_destroy(c);
        return r;
    };
    rc = hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking),
                   "hipStreamCreate");
    if (rc) {
        delete c;
        return rc;
    }
    // twiddles: tw[L + m] = e^{-2 pi i m / L} for L = 4..512
    {
        double2 *h = (double2 *)calloc(1024, sizeof(double2));
        for (int L = 4; L <= 512; L *= 2)
            for (int m = 0; m < L; ++m) {
                long double ang = -2.0L * 3.14159265358979323846264338327950288L *
                                  (long double)m / (long double)L;
                h[L + m] = make_double2((double)cosl(ang), (double)sinl(ang));
            }
        if ((rc = dalloc(&c->tw, 1024, c->stream)) == 0)
            rc = upload_now(c->tw, h, 1024 * sizeof(double2), c->stream);
        if (!rc) {
            float2 *hf = (float2 *)calloc(1024, sizeof(float2));
            for (int i = 0; i < 1024; ++i)
                hf[i] = make_float2((float)h[i].x, (float)h[i].y);
            if ((rc = hip_check(hipMalloc(&c->twf, 1024 * sizeof(float2)),
                                "hipMalloc")) == 0)
                rc = upload_now(c->twf, hf, 1024 * sizeof(float2), c->stream);
            free(hf);
        }
        free(h);
        if (rc) return fail(rc);
    }
    if (p->enable_filter) {
        const long nf = (long)c->Nx + c->Ny + c->Mz;
        double *h = (double *)malloc(nf * sizeof(double));
        if (!h) return fail(set_error("out of host memory"));
        host_filter_axes(c, h);
        if ((rc = dalloc(&c->filter, nf, c->stream)) == 0)
            rc = upload_now(c->filter, h, nf * sizeof(double), c->stream);
        free(h);
        if (rc) return fail(rc);
    }
    // per-realization work arrays of an ensemble, one after the other
    c->npartials = 3L * 512 * (p->bins > 8 ? p->bins : 8) + 2L * 65536 + 4096;
    if (nreal > 1) {
        c->ens.f = c->Ntot + 1; // even: every realization 16-byte aligned
        c->ens.w = 6 * c->M;
        c->ens.g = 3 * c->N;
        c->ens.p = c->npartials;
        c->ens.s = S_NSCAL;
    }
    if ((rc = dalloc(&c->W, 6 * c->M * nreal, c->stream))) return fail(rc);
    if ((rc = dalloc(&c->Wc, 2 * c->M, c->stream))) return fail(rc);
    if (p->enable_gw && (rc = dalloc(&c->grad3, 3 * c->N * nreal, c->stream)))
        return fail(rc);
    // rho | pressure | tmp.f in one allocation (rho and tmp.f 2N apart)
    if ((rc = dalloc(&c->rho, 3 * c->N, c->stream))) return fail(rc);
    c->pressure = c->rho + c->N;
    c->tmpf = c->rho + 2 * c->N;
    if ((rc = dalloc(&c->partials, c->npartials * nreal, c->stream))) return fail(rc);
    if ((rc = dalloc(&c->dscal, S_NSCAL * nreal, c->stream))) return fail(rc);
    c->obs_len = 28 + 4L * p->bins;
    if ((rc = dalloc(&c->dobs, c->obs_len, c->stream))) return fail(rc);
    if ((rc = hip_check(hipHostMalloc((void **)&c->hscal, 16 * sizeof(double)),
                        "hipHostMalloc")))
        return fail(rc);
    if ((rc = hip_check(hipHostMalloc((void **)&c->hobs,
                                      c->obs_len * sizeof(double)),
                        "hipHostMalloc")))
        return fail(rc);
    *out = c;
    return 0;
}

extern "C" int ihg_create(const ihg_params *p, ihg_ctx **out)
{
    return create_ctx(p, 1, out);
}

extern "C" int ihg_create_ensemble(const ihg_params *p, int nreal,
                                   ihg_ctx **out)
{
    return create_ctx(p, nreal, out);
}

extern "C" int ihg_ensemble(const ihg_ctx *c, long out[2])
{
    if (!c || !out) return set_error("NULL argument");
    out[0] = c->nreal;
    out[1] = c->nreal > 1 ? c->ens.f : c->Ntot;
    return 0;
}

extern "C" int ihg_sizes(const ihg_ctx *c, size_t out[4])
{
    if (!c || !out) return set_error("NULL argument");
    out[0] = c->N;
    out[1] = c->M;
    out[2] = c->Next;
    out[3] = c->Ntot;
    return 0;
}

extern "C" int ihg_malloc(ihg_ctx *c, double **ptr, size_t count)
{
    if (!c || !ptr) return set_error("NULL argument");
    IHG_TRY(hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    IHG_TRY(dalloc(ptr, count, c->stream));
    c->user_arrays.push_back(*ptr);
    return 0;
}

extern "C" int ihg_free(ihg_ctx *c, double *ptr)
{
    if (!c) return set_error("NULL context");
    if (!ptr) return 0;
    auto &v = c->user_arrays;
    auto it = std::find(v.begin(), v.end(), (void *)ptr);
    if (it == v.end())
        return set_error("ihg_free: %p was not allocated by this context", (void *)ptr);
    v.erase(it);
    IHG_TRY(hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    return hip_check(hipFree(ptr), "hipFree");
}

extern "C" int ihg_upload(ihg_ctx *c, double *dst, const double *src,
                          size_t count)
{
    if (!c || !dst || !src) return set_error("NULL argument");
    IHG_TRY(hip_check(hipMemcpyAsync(dst, src, count * sizeof(double),
                                     hipMemcpyHostToDevice, c->stream),
                      "hipMemcpyAsync H2D"));
    return hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

extern "C" int ihg_download(ihg_ctx *c, double *dst, const double *src,
                            size_t count)
{
    if (!c || !dst || !src) return set_error("NULL argument");
    IHG_TRY(hip_check(hipMemcpyAsync(dst, src, count * sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream),
                      "hipMemcpyAsync D2H"));
    return hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

extern "C" int ihg_copy(ihg_ctx *c, double *dst, const double *src,
                        size_t count)
{
    if (!c || !dst || !src) return set_error("NULL argument");
    return hip_check(hipMemcpyAsync(dst, src, count * sizeof(double),
                                    hipMemcpyDeviceToDevice, c->stream),
                     "hipMemcpyAsync D2D");
}

// a plain copy kernel (one double2 per thread; an odd last element by
// thread 0 of block 0): state snapshots, and the HBM ceiling and counter
// calibration the pass figures are compared with
template <bool NT>
__global__ __launch_bounds__(256) void k_stream_copy(
    const double *__restrict__ src, double *__restrict__ dst, long count)
{
    const long n2 = count / 2;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n2)
        ihg::st2<NT>(reinterpret_cast<double2 *>(dst) + i,
                     ihg::ld2<NT>(reinterpret_cast<const double2 *>(src) + i));
    if (i == 0 && (count & 1))
        dst[count - 1] = src[count - 1];
}

extern "C" int ihg_stream_copy(ihg_ctx *c, double *dst, const double *src,
                               size_t count)
{
    if (!c || !dst || !src) return set_error("NULL argument");
    if (((uintptr_t)dst | (uintptr_t)src) % 16)
        return set_error("ihg_stream_copy: 16-byte aligned arrays only");
    if (count == 0) return 0;
    const long n2 = (long)(count / 2) > 0 ? (long)(count / 2) : 1;
    return timed(c, F_STREAM_COPY, 16.0 * count, [&]() {
        const bool nt = 8.0 * (double)count >= ihg::NT_BYTES;
        hipLaunchKernelGGL((nt ? k_stream_copy<true> : k_stream_copy<false>),
                           dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0,
                           c->stream, src, dst, (long)count);
        return hip_check(hipGetLastError(), "k_stream_copy");
    });
}

extern "C" int ihg_zero(ihg_ctx *c, double *dst, size_t count)
{
    if (!c || !dst) return set_error("NULL argument");
    return hip_check(hipMemsetAsync(dst, 0, count * sizeof(double), c->stream),
                     "hipMemsetAsync");
}

extern "C" int ihg_sync(ihg_ctx *c)
{
    if (!c) return set_error("NULL context");
    return hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

extern "C" int ihg_fft_r2c(ihg_ctx *c, const double *in, double *out)
{
    if (!c) return set_error("NULL context");
    IHG_TRY(check_ptr(in, "in"));
    IHG_TRY(check_ptr(out, "out"));
    return fft_r2c(c, in, reinterpret_cast<double2 *>(out), 1, 0, 0);
}

extern "C" int ihg_fft_c2r(ihg_ctx *c, const double *in, double *out)
{
    if (!c) return set_error("NULL context");
    IHG_TRY(check_ptr(in, "in"));
    IHG_TRY(check_ptr(out, "out"));
    return fft_c2r(c, reinterpret_cast<const double2 *>(in), out, c->Wc, 1, 0,
                   0);
}

extern "C" int ihg_set_field(ihg_ctx *c, const double *field)
{
    if (!c) return set_error("NULL context");
    c->field = const_cast<double *>(field);
    return 0;
}

extern "C" int ihg_mk_rhs(ihg_ctx *c, double t, double *f, double *result,
                          unsigned flags)
{
    (void)t; // the equations are autonomous (src/toolbox.c:75-88)
    if (!c) return set_error("NULL context");
    IHG_TRY(check_ptr(f, "f"));
    IHG_TRY(check_ptr(result, "result"));
    if (f == result) return set_error("f and result must differ");
    c->calls_rhs += 1;
    const int output = (flags & IHG_OUTPUT) ? 1 : 0;
    // with the filter, the filtered spectra of phi and psi are kept for the
    // derivatives and the output (tmp.phic / tmp.psic, src/toolbox.c:164-171)
    const int kept = (flags & IHG_FILTER) && c->p.enable_filter;
    if (kept) IHG_TRY(rhs_filter(c, f));
    IHG_TRY(rhs_core(c, f, result, output, kept));
    if (c->p.enable_gw) IHG_TRY(rhs_gw(c, f, result));
    if (output) IHG_TRY(rhs_output(c, f, kept));
    return 0;
}

extern "C" int ihg_autotune(ihg_ctx *c, int reps, int choice[2],
                            float *ms, int nms)
{
    if (!c) return set_error("NULL context");
    if (reps < 1) reps = 1;
    float tmp[16] = {0};
    IHG_TRY(autotune(c, reps, choice, tmp));
    for (int i = 0; ms && i < nms && i < 16; ++i) ms[i] = tmp[i];
    return 0;
}

extern "C" int ihg_fft_candidates(ihg_ctx *c, int count[2], int current[2])
{
    if (!c) return set_error("NULL context");
    int n[2];
    ihg::candidate_counts(c, n);
    if (count) count[0] = n[0], count[1] = n[1];
    if (current) current[0] = c->cand_y, current[1] = c->cand_x;
    return 0;
}

extern "C" int ihg_set_fft_candidates(ihg_ctx *c, int y, int x)
{
    if (!c) return set_error("NULL context");
    int n[2];
    ihg::candidate_counts(c, n);
    if (y < 0 || y >= n[0] || x < 0 || x >= n[1])
        return set_error("FFT candidate out of range (y %d of %d, x %d of %d)",
                         y, n[0], x, n[1]);
    c->cand_y = y;
    c->cand_x = x;
    return 0;
}

extern "C" int ihg_set_gw_fp32(ihg_ctx *c, int on)
{
    if (!c) return set_error("NULL context");
    c->p.gw_fp32 = on ? 1 : 0;
    return 0;
}

extern "C" int ihg_features(ihg_ctx *c, const double *f, double out[3])
{
    if (!c || !out) return set_error("NULL argument");
    IHG_TRY(check_ptr(f, "f"));
    return features(c, f, out);
}

extern "C" int ihg_mk_rhs_ensemble(ihg_ctx *c, double t, double *f,
                                   double *result, unsigned flags)
{
    (void)t;
    if (!c) return set_error("NULL context");
    IHG_TRY(check_ptr(f, "f"));
    IHG_TRY(check_ptr(result, "result"));
    if (f == result) return set_error("f and result must differ");
    if (flags & IHG_OUTPUT)
        return set_error("ihg_mk_rhs_ensemble: output calls are per "
                         "realization (ihg_mk_rhs on its slice)");
    const int nr = c->nreal;
    c->calls_rhs += nr;
    const int kept = (flags & IHG_FILTER) && c->p.enable_filter;
    if (kept) IHG_TRY(rhs_filter(c, f, nr));
    IHG_TRY(rhs_core(c, f, result, 0, kept, nr));
    if (c->p.enable_gw) IHG_TRY(rhs_gw(c, f, result, nr));
    return 0;
}

extern "C" int ihg_mk_rhs_rk4(ihg_ctx *c, double t, double *f, double *result,
                              unsigned flags, const ihg_rk4_ep *ep)
{
    if (!c || !ep) return set_error("NULL argument");
    IHG_TRY(check_ptr(f, "f"));
    IHG_TRY(check_ptr(result, "result"));
    IHG_TRY(check_ptr(ep->y, "ep->y"));
    if (f == result) return set_error("f and result must differ");
    if (flags & IHG_OUTPUT)
        return set_error("ihg_mk_rhs_rk4: no output calls");
    Rk4Ep d{};
    d.dt = ep->dt;
    d.y = ep->y;
    if (ep->mode == IHG_RK4_STAGE) {
        IHG_TRY(check_ptr(ep->out, "ep->out"));
        if (ep->out == f || ep->out == result || ep->out == ep->y)
            return set_error("ihg_mk_rhs_rk4: out must differ from f, result and y");
        d.mode = RK4_STAGE;
        d.half = ep->half ? 1 : 0;
        d.out = ep->out;
    } else if (ep->mode == IHG_RK4_FINAL) {
        IHG_TRY(check_ptr(ep->k1, "ep->k1"));
        IHG_TRY(check_ptr(ep->k2, "ep->k2"));
        IHG_TRY(check_ptr(ep->k3, "ep->k3"));
        if (ep->y == f || ep->y == result)
            return set_error("ihg_mk_rhs_rk4: y must differ from f and result");
        if (ep->k1 == result || ep->k2 == result || ep->k3 == result)
            return set_error("ihg_mk_rhs_rk4: k1..k3 must differ from result");
        d.mode = RK4_FINAL;
        d.out = const_cast<double *>(ep->y); // y, in place
        d.k1 = ep->k1;
        d.k2 = ep->k2;
        d.k3 = ep->k3;
    } else {
        return set_error("ihg_mk_rhs_rk4: mode %d", ep->mode);
    }
    const int nr = c->nreal;
    const int kept = (flags & IHG_FILTER) && c->p.enable_filter;
    if (!rk4_fusable(c)) {
        // the same values from the separate stage kernels
        IHG_TRY(ihg_mk_rhs_ensemble(c, t, f, result, flags));
        if (d.mode == RK4_STAGE)
            return ihg_rk4_stage(c, d.out, d.y, result, d.dt, d.half);
        return ihg_rk4_final(c, d.out, d.k1, d.k2, d.k3, result, d.dt);
    }
    c->calls_rhs += nr;
    if (kept) IHG_TRY(rhs_filter(c, f, nr));
    IHG_TRY(rhs_core(c, f, result, 0, kept, nr, &d));
    if (c->p.enable_gw) IHG_TRY(rhs_gw(c, f, result, nr, &d));
    return 0;
}

extern "C" int ihg_means(ihg_ctx *c, double out[2])
{
    if (!c || !out) return set_error("NULL argument");
    IHG_TRY(hip_check(hipMemcpyAsync(c->hscal, c->dscal, 2 * sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream),
                      "hipMemcpyAsync"));
    IHG_TRY(hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    out[0] = c->hscal[0];
    out[1] = c->hscal[1];
    return 0;
}

extern "C" int ihg_observables(ihg_ctx *c, double *out, size_t len)
{
    if (!c || !out) return set_error("NULL argument");
    if (!c->have_obs) return set_error("no output call yet");
    if (len < (size_t)c->obs_len)
        return set_error("observable buffer needs %ld doubles", c->obs_len);
    IHG_TRY(hip_check(hipMemcpyAsync(c->hobs, c->dobs,
                                     c->obs_len * sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream),
                      "hipMemcpyAsync"));
    IHG_TRY(hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    memcpy(out, c->hobs, c->obs_len * sizeof(double));
    return 0;
}

extern "C" unsigned long ihg_calls_rhs(const ihg_ctx *c)
{
    return c ? c->calls_rhs : 0;
}

extern "C" int ihg_mk_initial_psi(ihg_ctx *c, double ti, double *field,
                                  double *dfield)
{
    if (!c) return set_error("NULL context");
    IHG_TRY(check_ptr(field, "field"));
    IHG_TRY(check_ptr(dfield, "dfield"));
    IHG_TRY(hip_check(hipMemsetAsync(field + 2 * c->N, 0,
                                     2 * c->N * sizeof(double), c->stream),
                      "hipMemsetAsync"));
    IHG_TRY(ihg_mk_rhs(c, ti, field, dfield, 0));
    return initial_psi(c, field, dfield);
}

// the stride-th points of a real field (row-major, z fastest) packed in
// the same order: out[(bx oy + by) oz + bz] = src[(bx sx Ny + by sy) Nz + bz sz]
__global__ __launch_bounds__(256) void k_gather_strided(
    const double *__restrict__ src, double *__restrict__ dst, long Ny, long Nz,
    long oy, long oz, long total, int sx, int sy, int sz)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const long bz = i % oz;
    const long t = i / oz;
    const long by = t % oy;
    const long bx = t / oy;
    dst[i] = src[(bx * sx * Ny + by * sy) * Nz + bz * sz];
}

extern "C" int ihg_download_strided(ihg_ctx *c, const double *src,
                                    const int stride[3], double *out)
{
    if (!c || !stride || !out) return set_error("NULL argument");
    if (c->nreal > 1) return set_error("ihg_download_strided: not on an ensemble context");
    const long n[3] = {c->Nx, c->Ny, c->Nz};
    long o[3];
    for (int d = 0; d < 3; ++d) {
        if (stride[d] < 1 || stride[d] > n[d])
            return set_error("stride[%d] = %d outside 1..%ld", d, stride[d], n[d]);
        o[d] = (n[d] + stride[d] - 1) / stride[d];
    }
    if (!src) src = c->rho; // rho of the last mk_rhs
    const long total = o[0] * o[1] * o[2];
    if (total == c->N) return ihg_download(c, out, src, (size_t)c->N);
    if ((size_t)total > c->gather_len) {
        if (c->gather) IHG_HIP(hipFree(c->gather));
        c->gather = nullptr;
        c->gather_len = 0;
        IHG_TRY(dalloc(&c->gather, (size_t)total, c->stream));
        c->gather_len = (size_t)total;
    }
    hipLaunchKernelGGL(k_gather_strided, dim3((unsigned)((total + 255) / 256)),
                       dim3(256), 0, c->stream, src, c->gather, n[1], n[2], o[1],
                       o[2], total, stride[0], stride[1], stride[2]);
    IHG_LAUNCHED("k_gather_strided");
    return ihg_download(c, out, c->gather, (size_t)total);
}

extern "C" int ihg_aux(ihg_ctx *c, int which, double *out)
{
    if (!c || !out) return set_error("NULL argument");
    const double *src = nullptr;
    if (which >= 0 && which <= 2) {
        if (!c->grad3) return set_error("gradients are kept only with GW on");
        if (c->p.gw_fp32) {
            // the precision form keeps them in FP32
            float *tmp = (float *)malloc(c->N * sizeof(float));
            if (!tmp) return set_error("out of host memory");
            int rc = hip_check(hipMemcpyAsync(tmp,
                                              reinterpret_cast<const float *>(c->grad3) + which * c->N,
                                              c->N * sizeof(float),
                                              hipMemcpyDeviceToHost, c->stream),
                               "download gradients");
            if (!rc) rc = hip_check(hipStreamSynchronize(c->stream), "sync");
            for (long i = 0; !rc && i < c->N; ++i) out[i] = tmp[i];
            free(tmp);
            return rc;
        }
        src = c->grad3 + which * c->N;
    } else if (which == 3) {
        src = c->rho;
    } else if (which == 4) {
        src = c->pressure;
    } else if (which == 5) {
        src = c->tmpf;
    } else {
        return set_error("aux index %d out of range", which);
    }
    return ihg_download(c, out, src, c->N);
}
