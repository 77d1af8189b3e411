// ihg_step.hip -- stage arithmetic of run_rk4 (src/rk4.c:78-102) and of
// try_step / error (src/dopri853.c:331-486).
//
// All stage kernels stream the live part of the state: [0, 4N) plus, with
// GW on, the four h blocks; the scale factor a (last element) is handled by
// thread 0.  With GW off the h blocks of every stepper array stay zero (the
// reference zeroes its DOPRI853 buffers, src/dopri853.c:610-624, and never
// writes them), so skipping them changes no value; the error norm still
// divides by Ntot.
#include <cmath>
#include <mutex>
#include <type_traits>

#include "dp853_tableau.h"
#include "gsl_tableau.h"
#include "ihg_internal.h"
#include "mem.h"
#include "reduce.h"
#include "rk4_kernels.h"

namespace ihg {

struct StageIn {
    const double *k[12];
    double a[12];
    int nk;
};

static inline long live_len(const ihg_ctx *c)
{
    return c->p.enable_gw ? c->Ntot - 1 : 4 * c->N;
}

// RK4 vectors of a context: one state (live part + a) or, for an ensemble,
// all realizations' states back to back (nreal * stride, every element)
static inline void rk4_extent(const ihg_ctx *c, long &nl, long &tail)
{
    if (c->nreal > 1) {
        nl = c->nreal * c->ens.f;
        tail = -1;
    } else {
        nl = live_len(c);
        tail = c->Ntot - 1;
    }
}

// whether arrays of n doubles take the streaming hint
static inline bool streaming(long n) { return 8.0 * (double)n >= NT_BYTES; }

static inline unsigned stream_blocks(long npairs)
{
    long b = (npairs + RED_WG - 1) / RED_WG;
    if (b < 1) b = 1;
    return (unsigned)(b > 262144 ? 262144 : b);
}

// out = y + dt * sum_j a_j k_j   (try_step's stage loops); NT: streaming
// loads and stores (arrays far larger than the caches)
template <bool NT>
__global__ __launch_bounds__(RED_WG) void k_dp_stage(
    double *__restrict__ out, const double *__restrict__ y, StageIn s,
    double dt, long nlive, long Ntot)
{
    const long npairs = nlive / 2;
    const long stride = (long)gridDim.x * RED_WG;
    const double2 *y2 = reinterpret_cast<const double2 *>(y);
    double2 *o2 = reinterpret_cast<double2 *>(out);
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < npairs;
         i += stride) {
        const double2 k0 = ld2<NT>(reinterpret_cast<const double2 *>(s.k[0]) + i);
        double ax = s.a[0] * k0.x, ay = s.a[0] * k0.y;
#pragma unroll
        for (int j = 1; j < 12; ++j) {
            if (j < s.nk) {
                const double2 kj =
                    ld2<NT>(reinterpret_cast<const double2 *>(s.k[j]) + i);
                ax += s.a[j] * kj.x;
                ay += s.a[j] * kj.y;
            }
        }
        const double2 yy = ld2<NT>(y2 + i);
        st2<NT>(o2 + i, make_double2(yy.x + dt * ax, yy.y + dt * ay));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const long i = Ntot - 1;
        double acc = s.a[0] * s.k[0][i];
        for (int j = 1; j < s.nk; ++j)
            acc += s.a[j] * s.k[j][i];
        out[i] = y[i] + dt * acc;
    }
}

// dense output: out = y + beta (y_new - y) + dt * sum_j w_j k_j, the
// interpolant of an accepted step at one theta (weights from the host)
template <bool NT>
__global__ __launch_bounds__(RED_WG) void k_dp_dense(
    double *__restrict__ out, const double *__restrict__ y,
    const double *__restrict__ y_new, double beta, StageIn s, double dt,
    long nlive, long Ntot)
{
    const long npairs = nlive / 2;
    const long stride = (long)gridDim.x * RED_WG;
    const double2 *y2 = reinterpret_cast<const double2 *>(y);
    const double2 *n2 = reinterpret_cast<const double2 *>(y_new);
    double2 *o2 = reinterpret_cast<double2 *>(out);
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < npairs;
         i += stride) {
        double ax = 0.0, ay = 0.0;
#pragma unroll
        for (int j = 0; j < 12; ++j) {
            if (j < s.nk) {
                const double2 kj =
                    ld2<NT>(reinterpret_cast<const double2 *>(s.k[j]) + i);
                ax += s.a[j] * kj.x;
                ay += s.a[j] * kj.y;
            }
        }
        const double2 yy = ld2<NT>(y2 + i), yn = ld2<NT>(n2 + i);
        st2<NT>(o2 + i, make_double2(yy.x + beta * (yn.x - yy.x) + dt * ax,
                                     yy.y + beta * (yn.y - yy.y) + dt * ay));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const long i = Ntot - 1;
        double acc = 0.0;
        for (int j = 0; j < s.nk; ++j)
            acc += s.a[j] * s.k[j][i];
        out[i] = y[i] + beta * (y_new[i] - y[i]) + dt * acc;
    }
}

template <bool NT>
__global__ __launch_bounds__(RED_WG) void k_rk4_stage(
    double *__restrict__ out, const double *__restrict__ y,
    const double *__restrict__ k, double dt, int half, long nlive, long tail)
{
    rk4_stage_body<NT>(blockIdx.x, gridDim.x, out, y, k, dt, half, nlive, tail);
}

template <bool NT>
__global__ __launch_bounds__(RED_WG) void k_rk4_final(
    double *__restrict__ y, const double *__restrict__ k1,
    const double *__restrict__ k2, const double *__restrict__ k3,
    const double *__restrict__ k4, double dt, long nlive, long tail)
{
    rk4_final_body<NT>(blockIdx.x, gridDim.x, y, k1, k2, k3, k4, dt, nlive,
                       tail);
}

struct FinishIn {
    const double *k[12];
};

__device__ __forceinline__ void dp_finish_point(const FinishIn &f, long i,
                                                double yv, double dt,
                                                double rtol, double atol,
                                                double &ynew, double &e5,
                                                double &e3)
{
    const double k0 = f.k[0][i], k5 = f.k[5][i], k6 = f.k[6][i],
                 k7 = f.k[7][i], k8 = f.k[8][i], k9 = f.k[9][i],
                 k10 = f.k[10][i], k11 = f.k[11][i];
    const double bs = DP_B[0] * k0 + DP_B[5] * k5 + DP_B[6] * k6 +
                      DP_B[7] * k7 + DP_B[8] * k8 + DP_B[9] * k9 +
                      DP_B[10] * k10 + DP_B[11] * k11;
    ynew = yv + dt * bs;
    const double yerr = bs - DP_BHH[0] * k0 - DP_BHH[1] * k8 - DP_BHH[2] * k11;
    const double yerr2 = DP_ER[0] * k0 + DP_ER[5] * k5 + DP_ER[6] * k6 +
                         DP_ER[7] * k7 + DP_ER[8] * k8 + DP_ER[9] * k9 +
                         DP_ER[10] * k10 + DP_ER[11] * k11;
    const double inv = 1.0 / (atol + rtol * fmax(fabs(yv), fabs(ynew)));
    double t = yerr * inv;
    e3 += t * t;
    t = yerr2 * inv;
    e5 += t * t;
}

// the same for the elements 2i, 2i+1 (their error terms summed in that
// order)
template <bool NT>
__device__ __forceinline__ void dp_finish_pair(const FinishIn &f, long i,
                                               double2 yv, double dt,
                                               double rtol, double atol,
                                               double2 &ynew, double &e5,
                                               double &e3)
{
    const double2 *k[12];
#pragma unroll
    for (int j = 0; j < 12; ++j)
        k[j] = reinterpret_cast<const double2 *>(f.k[j]);
    const double2 k0 = ld2<NT>(k[0] + i), k5 = ld2<NT>(k[5] + i),
                  k6 = ld2<NT>(k[6] + i), k7 = ld2<NT>(k[7] + i),
                  k8 = ld2<NT>(k[8] + i), k9 = ld2<NT>(k[9] + i),
                  k10 = ld2<NT>(k[10] + i), k11 = ld2<NT>(k[11] + i);
    auto one = [&](double y0, double a0, double a5, double a6, double a7,
                   double a8, double a9, double a10, double a11) {
        const double bs = DP_B[0] * a0 + DP_B[5] * a5 + DP_B[6] * a6 +
                          DP_B[7] * a7 + DP_B[8] * a8 + DP_B[9] * a9 +
                          DP_B[10] * a10 + DP_B[11] * a11;
        const double yn = y0 + dt * bs;
        const double yerr = bs - DP_BHH[0] * a0 - DP_BHH[1] * a8 - DP_BHH[2] * a11;
        const double yerr2 = DP_ER[0] * a0 + DP_ER[5] * a5 + DP_ER[6] * a6 +
                             DP_ER[7] * a7 + DP_ER[8] * a8 + DP_ER[9] * a9 +
                             DP_ER[10] * a10 + DP_ER[11] * a11;
        // one division per element: both ratios scale by 1 / sk
        const double inv = 1.0 / (atol + rtol * fmax(fabs(y0), fabs(yn)));
        double t = yerr * inv;
        e3 += t * t;
        t = yerr2 * inv;
        e5 += t * t;
        return yn;
    };
    ynew.x = one(yv.x, k0.x, k5.x, k6.x, k7.x, k8.x, k9.x, k10.x, k11.x);
    ynew.y = one(yv.y, k0.y, k5.y, k6.y, k7.y, k8.y, k9.y, k10.y, k11.y);
}

// b-weighted update, embedded errors and the two sums of error()
template <bool NT>
__global__ __launch_bounds__(RED_WG) void k_dp_finish(
    const double *__restrict__ y, FinishIn f, double *__restrict__ ynew,
    double dt, double rtol, double atol, long nlive, long Ntot,
    double *__restrict__ part)
{
    __shared__ double sh[8];
    double e5 = 0.0, e3 = 0.0;
    const long stride = (long)gridDim.x * RED_WG;
    // pairs of elements (16-byte loads and stores; nlive is even), then the
    // scale factor a
    const long npairs = nlive / 2;
    const double2 *y2 = reinterpret_cast<const double2 *>(y);
    double2 *n2 = reinterpret_cast<double2 *>(ynew);
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < npairs;
         i += stride) {
        double2 yn;
        const double2 yv = ld2<NT>(y2 + i);
        dp_finish_pair<NT>(f, i, yv, dt, rtol, atol, yn, e5, e3);
        st2<NT>(n2 + i, yn);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double yn;
        dp_finish_point(f, Ntot - 1, y[Ntot - 1], dt, rtol, atol, yn, e5, e3);
        ynew[Ntot - 1] = yn;
    }
    e5 = block_reduce<OpSum>(e5, sh, RED_WG);
    e3 = block_reduce<OpSum>(e3, sh, RED_WG);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = e5;
        part[2 * blockIdx.x + 1] = e3;
    }
}

// first level of the sum of many block-partial pairs: work-group g adds the
// pairs [g * chunk, (g + 1) * chunk) (a fixed partition: deterministic)
__global__ __launch_bounds__(RED_WG) void k_sum2_mid(
    const double *__restrict__ part, long nb, long chunk,
    double *__restrict__ out)
{
    __shared__ double sh[8];
    const long b0 = (long)blockIdx.x * chunk;
    const long b1 = b0 + chunk < nb ? b0 + chunk : nb;
    const double2 *p2 = reinterpret_cast<const double2 *>(part);
    double s0 = 0.0, s1 = 0.0;
#pragma unroll 4
    for (long b = b0 + threadIdx.x; b < b1; b += RED_WG) {
        const double2 v = p2[b];
        s0 += v.x;
        s1 += v.y;
    }
    s0 = block_reduce<OpSum>(s0, sh, RED_WG);
    s1 = block_reduce<OpSum>(s1, sh, RED_WG);
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = s0;
        out[2 * blockIdx.x + 1] = s1;
    }
}

__global__ __launch_bounds__(RED_WG) void k_sum2_final(
    const double *__restrict__ part, int nb, double *__restrict__ out)
{
    __shared__ double sh[8];
    double s0 = 0.0, s1 = 0.0;
    for (int b = threadIdx.x; b < nb; b += RED_WG) {
        s0 += part[2 * b];
        s1 += part[2 * b + 1];
    }
    s0 = block_reduce<OpSum>(s0, sh, RED_WG);
    s1 = block_reduce<OpSum>(s1, sh, RED_WG);
    if (threadIdx.x == 0) {
        out[0] = s0;
        out[1] = s1;
    }
}

// sums of (a - b)^2 and (c - d)^2 over n elements (check_for_stiffness,
// src/dopri853.c:550-560), block partials for k_sum2_final
__global__ __launch_bounds__(RED_WG) void k_diff_norms(
    const double *__restrict__ a, const double *__restrict__ b,
    const double *__restrict__ c, const double *__restrict__ d, long n,
    double *__restrict__ part)
{
    __shared__ double sh[8];
    double s0 = 0.0, s1 = 0.0;
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < n; i += stride) {
        const double u = a[i] - b[i], v = c[i] - d[i];
        s0 += u * u;
        s1 += v * v;
    }
    s0 = block_reduce<OpSum>(s0, sh, RED_WG);
    s1 = block_reduce<OpSum>(s1, sh, RED_WG);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0;
        part[2 * blockIdx.x + 1] = s1;
    }
}

// the GSL steppers' final sum and the standard control's error ratio
// (gsl_odeiv2 rkf45 / rkck / rk8pd step and control y_new):
//   y_new = y + dt sum_j b_j k_j
//   yerr  = dt sum_j w_j k_j            (err_mode 0, w = e)
//         = dt (sum_j w_j k_j - sum_j b_j k_j)   (err_mode 1, w = bhat)
//   r     = max_i |yerr_i| / |eps_rel |y_new_i| + eps_abs|
// A NaN ratio never replaces the running maximum (the control's `>`).
struct GslFinishIn {
    const double *k[GSL_MAX_STAGES];
    double b[GSL_MAX_STAGES];
    double w[GSL_MAX_STAGES];
    int nk;
    int err_mode;
};

// y_new and the error ratio of one element from its two weighted sums
__device__ __forceinline__ double gsl_ratio(const GslFinishIn &f, double yv,
                                            double sb, double sw, double dt,
                                            double eps_abs, double eps_rel,
                                            double &rm)
{
    const double yn = yv + dt * sb;
    const double yerr = f.err_mode ? dt * (sw - sb) : dt * sw;
    const double r = fabs(yerr) / fabs(eps_rel * fabs(yn) + eps_abs);
    rm = r > rm ? r : rm;
    return yn;
}

__device__ __forceinline__ double gsl_point(const GslFinishIn &f, double yv,
                                            const double *kv, double dt,
                                            double eps_abs, double eps_rel,
                                            double &rm)
{
    double sb = 0.0, sw = 0.0;
#pragma unroll
    for (int j = 0; j < GSL_MAX_STAGES; ++j) {
        if (j < f.nk) {
            sb += f.b[j] * kv[j];
            sw += f.w[j] * kv[j];
        }
    }
    return gsl_ratio(f, yv, sb, sw, dt, eps_abs, eps_rel, rm);
}

// The final sum and the error ratio with the number of weighted stages NK
// fixed at compile time (the kernel is instantiated per count), so that
// every stage load of a pair is issued before the sums need it; one element
// pair per thread.  The block maxima go to part[0, gridDim.x).
template <bool NT, int NK>
__global__ __launch_bounds__(RED_WG) void k_gsl_finish(
    const double *__restrict__ y, GslFinishIn f, double *__restrict__ ynew,
    double dt, double eps_abs, double eps_rel, long nlive, long Ntot,
    double *__restrict__ part)
{
    __shared__ double sh[8];
    double rm = 0.0;
    const long stride = (long)gridDim.x * RED_WG;
    const long npairs = nlive / 2;
    const double2 *y2 = reinterpret_cast<const double2 *>(y);
    double2 *n2 = reinterpret_cast<double2 *>(ynew);
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < npairs;
         i += stride) {
        double2 kv[NK];
#pragma unroll
        for (int j = 0; j < NK; ++j)
            kv[j] = ld2<NT>(reinterpret_cast<const double2 *>(f.k[j]) + i);
        const double2 yv = ld2<NT>(y2 + i);
        double2 yn;
        double sb = 0.0, sw = 0.0;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            sb += f.b[j] * kv[j].x;
            sw += f.w[j] * kv[j].x;
        }
        yn.x = gsl_ratio(f, yv.x, sb, sw, dt, eps_abs, eps_rel, rm);
        sb = 0.0;
        sw = 0.0;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            sb += f.b[j] * kv[j].y;
            sw += f.w[j] * kv[j].y;
        }
        yn.y = gsl_ratio(f, yv.y, sb, sw, dt, eps_abs, eps_rel, rm);
        st2<NT>(n2 + i, yn);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const long i = Ntot - 1;
        double kv[GSL_MAX_STAGES];
        for (int j = 0; j < f.nk; ++j)
            kv[j] = f.k[j][i];
        ynew[i] = gsl_point(f, y[i], kv, dt, eps_abs, eps_rel, rm);
    }
    rm = block_reduce<OpMax>(rm, sh, RED_WG);
    if (threadIdx.x == 0) part[blockIdx.x] = rm;
}

// first level of the maximum of many block maxima: work-group g takes
// part[g * chunk, (g + 1) * chunk)
__global__ __launch_bounds__(RED_WG) void k_max_mid(
    const double *__restrict__ part, long nb, long chunk,
    double *__restrict__ out)
{
    __shared__ double sh[8];
    const long b0 = (long)blockIdx.x * chunk;
    const long b1 = b0 + chunk < nb ? b0 + chunk : nb;
    double m = 0.0;
    for (long b = b0 + threadIdx.x; b < b1; b += RED_WG)
        m = part[b] > m ? part[b] : m;
    m = block_reduce<OpMax>(m, sh, RED_WG);
    if (threadIdx.x == 0) out[blockIdx.x] = m;
}

__global__ __launch_bounds__(RED_WG) void k_max_final(
    const double *__restrict__ part, int nb, double *__restrict__ out)
{
    __shared__ double sh[8];
    double m = 0.0;
    for (int b = threadIdx.x; b < nb; b += RED_WG)
        m = part[b] > m ? part[b] : m;
    m = block_reduce<OpMax>(m, sh, RED_WG);
    if (threadIdx.x == 0) out[0] = m;
}

static const GslTableau *gsl_tableau_of(int method)
{
    switch (method) {
    case 1: return &GSL_RKF45;
    case 2: return &GSL_RKCK;
    case 3: return &GSL_RK8PD;
    default: return nullptr;
    }
}

} // namespace ihg

using namespace ihg;

static int check_ptr(const void *p, const char *name)
{
    if (!p) return set_error("%s is NULL", name);
    if (((size_t)p) & 15)
        return set_error("%s is not 16-byte aligned", name);
    return 0;
}

extern "C" int ihg_rk4_stage(ihg_ctx *c, double *out, const double *y,
                             const double *k, double dt, int half)
{
    if (!c) return set_error("NULL context");
    IHG_TRY(check_ptr(out, "out"));
    IHG_TRY(check_ptr(y, "y"));
    IHG_TRY(check_ptr(k, "k"));
    long nl, tail;
    rk4_extent(c, nl, tail);
    return timed(c, F_RK4, 3.0 * 8.0 * nl, [&]() {
        hipLaunchKernelGGL((streaming(nl) ? k_rk4_stage<true>
                                          : k_rk4_stage<false>),
                           dim3(stream_blocks(nl / 2)),
                           dim3(RED_WG), 0, c->stream, out, y, k, dt,
                           half ? 1 : 0, nl, tail);
        return hip_check(hipGetLastError(), "k_rk4_stage");
    });
}

extern "C" int ihg_rk4_final(ihg_ctx *c, double *y, const double *k1,
                             const double *k2, const double *k3,
                             const double *k4, double dt)
{
    if (!c) return set_error("NULL context");
    IHG_TRY(check_ptr(y, "y"));
    IHG_TRY(check_ptr(k1, "k1"));
    IHG_TRY(check_ptr(k2, "k2"));
    IHG_TRY(check_ptr(k3, "k3"));
    IHG_TRY(check_ptr(k4, "k4"));
    long nl, tail;
    rk4_extent(c, nl, tail);
    return timed(c, F_RK4, 6.0 * 8.0 * nl, [&]() {
        hipLaunchKernelGGL((streaming(nl) ? k_rk4_final<true>
                                          : k_rk4_final<false>),
                           dim3(stream_blocks(nl / 2)),
                           dim3(RED_WG), 0, c->stream, y, k1, k2, k3, k4, dt,
                           nl, tail);
        return hip_check(hipGetLastError(), "k_rk4_final");
    });
}

extern "C" int ihg_dp853_stage(ihg_ctx *c, int s, double dt, const double *y,
                               const double *const *k, double *out)
{
    if (!c || !k) return set_error("NULL argument");
    if (s < 2 || s > 12) return set_error("stage %d out of range 2..12", s);
    IHG_TRY(check_ptr(y, "y"));
    IHG_TRY(check_ptr(out, "out"));
    StageIn in{};
    int n = 0;
    for (int j = 0; j < s - 1; ++j) {
        if (DP_A[s - 1][j] == 0.0) continue;
        IHG_TRY(check_ptr(k[j], "k[j]"));
        in.k[n] = k[j];
        in.a[n] = DP_A[s - 1][j];
        ++n;
    }
    in.nk = n;
    const long nl = live_len(c);
    return timed(c, F_DP_STAGE, (n + 2.0) * 8.0 * nl, [&]() {
        hipLaunchKernelGGL((streaming(nl) ? k_dp_stage<true>
                                          : k_dp_stage<false>),
                           dim3(stream_blocks(nl / 2)),
                           dim3(RED_WG), 0, c->stream, out, y, in, dt, nl,
                           c->Ntot);
        return hip_check(hipGetLastError(), "k_dp_stage");
    });
}

extern "C" int ihg_dp853_dense_stage(ihg_ctx *c, int s, double dt,
                                     const double *y, const double *const *k,
                                     double *out)
{
    if (!c || !k) return set_error("NULL argument");
    if (s < 14 || s > 16) return set_error("dense stage %d out of range 14..16", s);
    IHG_TRY(check_ptr(y, "y"));
    IHG_TRY(check_ptr(out, "out"));
    StageIn in{};
    int n = 0;
    for (int j = 0; j < s - 1; ++j) {
        const double a = DP_A_EXTRA[s - 14][j];
        if (a == 0.0) continue;
        IHG_TRY(check_ptr(k[j], "k[j]"));
        in.k[n] = k[j];
        in.a[n] = a;
        ++n;
    }
    in.nk = n;
    const long nl = live_len(c);
    return timed(c, F_DP_DENSE_STAGE, (n + 2.0) * 8.0 * nl, [&]() {
        hipLaunchKernelGGL((streaming(nl) ? k_dp_stage<true>
                                          : k_dp_stage<false>),
                           dim3(stream_blocks(nl / 2)),
                           dim3(RED_WG), 0, c->stream, out, y, in, dt, nl,
                           c->Ntot);
        return hip_check(hipGetLastError(), "k_dp_stage");
    });
}

// The interpolant of DOP853's dense output at theta in [0, 1]:
//   y(theta) = y + theta (F0 + (1-theta) (F1 + theta (F2 + (1-theta) (F3 +
//              theta (F4 + (1-theta) (F5 + theta F6))))))
// with F0 = y_new - y, F1 = dt k0 - F0, F2 = 2 F0 - dt (k12 + k0) and
// F3..F6 = dt sum_j DP_D[r][j] k_j, written as y + beta (y_new - y) +
// dt sum_j w_j k_j with the weights of the nested form's products
// p_m = theta^ceil((m+1)/2) (1-theta)^floor((m+1)/2), m = 0..6.
extern "C" int ihg_dp853_dense(ihg_ctx *c, double theta, double dt,
                               const double *y, const double *y_new,
                               const double *const *k, double *out)
{
    if (!c || !k) return set_error("NULL argument");
    if (!(theta >= 0.0 && theta <= 1.0))
        return set_error("dense output at theta = %g outside [0, 1]", theta);
    IHG_TRY(check_ptr(y, "y"));
    IHG_TRY(check_ptr(y_new, "y_new"));
    IHG_TRY(check_ptr(out, "out"));
    const double u = 1.0 - theta;
    double p[7];
    p[0] = theta;
    for (int m = 1; m < 7; ++m)
        p[m] = p[m - 1] * ((m & 1) ? u : theta);
    double w[16] = {0.0};
    w[0] = p[1] - p[2];
    w[12] = -p[2];
    for (int r = 0; r < 4; ++r)
        for (int j = 0; j < 16; ++j)
            w[j] += p[3 + r] * DP_D[r][j];
    const double beta = p[0] - p[1] + 2.0 * p[2];
    StageIn in{};
    int n = 0;
    for (int j = 0; j < 16; ++j) {
        if (w[j] == 0.0) continue;
        if (n == 12) return set_error("dense output: more than 12 stages");
        IHG_TRY(check_ptr(k[j], "k[j]"));
        in.k[n] = k[j];
        in.a[n] = w[j];
        ++n;
    }
    in.nk = n;
    const long nl = live_len(c);
    return timed(c, F_DP_DENSE, (n + 3.0) * 8.0 * nl, [&]() {
        hipLaunchKernelGGL((streaming(nl) ? k_dp_dense<true>
                                          : k_dp_dense<false>),
                           dim3(stream_blocks(nl / 2)),
                           dim3(RED_WG), 0, c->stream, out, y, y_new, beta, in,
                           dt, nl, c->Ntot);
        return hip_check(hipGetLastError(), "k_dp_dense");
    });
}

extern "C" int ihg_dp853_dense_tableau(double *c_extra, double *a_extra,
                                       double *d)
{
    if (!c_extra || !a_extra || !d) return set_error("NULL argument");
    for (int i = 0; i < 3; ++i) {
        c_extra[i] = DP_C_EXTRA[i];
        for (int j = 0; j < 16; ++j)
            a_extra[i * 16 + j] = DP_A_EXTRA[i][j];
    }
    for (int r = 0; r < 4; ++r)
        for (int j = 0; j < 16; ++j)
            d[r * 16 + j] = DP_D[r][j];
    return 0;
}

// the stage arrays the final sum reads: stages 2..5 carry no b / er weight,
// so their slots repeat k[0] (never dereferenced for a weight of 0)
static int finish_inputs(const double *y, const double *const *k,
                         const double *y_new, FinishIn &in)
{
    if (!k) return set_error("NULL argument");
    IHG_TRY(check_ptr(y, "y"));
    IHG_TRY(check_ptr(y_new, "y_new"));
    for (int j = 0; j < 12; ++j) {
        if (j >= 1 && j <= 4) continue;
        IHG_TRY(check_ptr(k[j], "k[j]"));
        in.k[j] = k[j];
    }
    for (int j = 1; j <= 4; ++j)
        in.k[j] = k[0];
    return 0;
}

// work-groups of the allocation-free finish beyond which its kernel
// strides over the rest (its fixed array holds that many partials)
constexpr long FINISH_MAX_GROUPS = 32768;

// The context's finish takes one element pair per thread (up to
// FINISH_WIDE_GROUPS work-groups; 263 K at 256^3): streaming without the
// grid-stride loop runs at 5.3 against 4.9 TB/s.  Its block partials are
// summed in two levels, FINISH_MID work-groups over fixed chunks and then
// one, so the order of the sums is fixed.
constexpr long FINISH_WIDE_GROUPS = 1L << 20;
constexpr long FINISH_MID = 1024;
static long finish_groups_wide(long nl)
{
    long nb = (nl / 2 + RED_WG - 1) / RED_WG;
    if (nb > FINISH_WIDE_GROUPS) nb = FINISH_WIDE_GROUPS;
    return nb < 1 ? 1 : nb;
}

// work-groups of a finish and of the first level of its sum
struct FinishPlan {
    long nb, nmid, chunk;
    long doubles() const { return 2 * nb + 2 * nmid + 2; } // + the sums
};

static FinishPlan finish_plan(long nb)
{
    FinishPlan p;
    p.nb = nb;
    p.nmid = nb < FINISH_MID ? 1 : FINISH_MID;
    p.chunk = (nb + p.nmid - 1) / p.nmid;
    return p;
}

// k_dp_finish over nlive with its partials in part[0, 2 nb)
static int launch_finish(hipStream_t s, const FinishPlan &fp, const double *y,
                         const FinishIn &in, double *y_new, double dt,
                         double rtol, double atol, long nlive, long ntot,
                         double *part)
{
    hipLaunchKernelGGL((streaming(nlive) ? k_dp_finish<true>
                                         : k_dp_finish<false>),
                       dim3((unsigned)fp.nb), dim3(RED_WG), 0, s, y, in, y_new,
                       dt, rtol, atol, nlive, ntot, part);
    IHG_LAUNCHED("k_dp_finish");
    return 0;
}

// the two sums of those partials: the first level in part[2 nb, 2 nb +
// 2 nmid), the sums in the last two doubles of the plan
static int launch_finish_sums(hipStream_t s, const FinishPlan &fp, double *part)
{
    double *mid = part + 2 * fp.nb, *sums = mid + 2 * fp.nmid;
    if (fp.nmid > 1) {
        hipLaunchKernelGGL(k_sum2_mid, dim3((unsigned)fp.nmid), dim3(RED_WG),
                           0, s, part, fp.nb, fp.chunk, mid);
        IHG_LAUNCHED("k_sum2_mid");
    }
    hipLaunchKernelGGL(k_sum2_final, dim3(1), dim3(RED_WG), 0, s,
                       fp.nmid > 1 ? mid : part,
                       (int)(fp.nmid > 1 ? fp.nmid : fp.nb), sums);
    IHG_LAUNCHED("k_sum2_final");
    return 0;
}

// the context's finish partials hold at least n doubles (allocated by the
// first finish that needs them, once per context and size)
static int ensure_fin_partials(ihg_ctx *c, long n)
{
    if (n <= c->fin_npartials) return 0;
    IHG_HIP(hipStreamSynchronize(c->stream));
    if (c->fin_partials) IHG_HIP(hipFree(c->fin_partials));
    c->fin_partials = nullptr;
    c->fin_npartials = 0;
    IHG_HIP(hipMalloc((void **)&c->fin_partials, n * sizeof(double)));
    c->fin_npartials = n;
    return 0;
}

extern "C" int ihg_dp853_finish(ihg_ctx *c, double dt, const double *y,
                                const double *const *k, double *y_new,
                                double rtol, double atol, double sums[2])
{
    if (!c || !sums) return set_error("NULL argument");
    FinishIn in{};
    IHG_TRY(finish_inputs(y, k, y_new, in));
    const long nl = live_len(c);
    const FinishPlan fp = finish_plan(finish_groups_wide(nl));
    IHG_TRY(ensure_fin_partials(c, fp.doubles()));
    double *sums_dev = c->fin_partials + fp.doubles() - 2;
    // y and the 8 weighted stages in, y_new out
    IHG_TRY(timed(c, F_DP_FINISH, 10.0 * 8.0 * nl, [&]() {
        return launch_finish(c->stream, fp, y, in, y_new, dt, rtol, atol, nl,
                             c->Ntot, c->fin_partials);
    }));
    IHG_TRY(launch_finish_sums(c->stream, fp, c->fin_partials));
    IHG_HIP(hipMemcpyAsync(c->hscal + 8, sums_dev, 2 * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    IHG_HIP(hipStreamSynchronize(c->stream));
    sums[0] = c->hscal[8];
    sums[1] = c->hscal[9];
    return 0;
}

// ---- ihg_dp853_finish without a context and without allocation --------
// The block partials and the two sums live in `scratch`, one copy per GPU
// for the life of the process.  Host code never passes `scratch` itself:
// its device address, looked up once per device, is kept in scratch_dev.
// the partials of FINISH_MAX_GROUPS work-groups, the first level of their
// sum and the two sums
constexpr long NOALLOC_DOUBLES = 2 * FINISH_MAX_GROUPS + 2 * FINISH_MID + 2;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static double *scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

static int scratch_of_current_device(double **out)
{
    int dev = -1;
    IHG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return set_error("device %d beyond the %d devices the scratch table "
                         "holds", dev, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        IHG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error("hipGetSymbolAddress gave no address");
        scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = scratch_dev[dev];
    return 0;
}

extern "C" int ihg_dp853_finish_noalloc(void *stream, long nlive, long ntot,
                                        double dt, const double *y,
                                        const double *const *k, double *y_new,
                                        double rtol, double atol,
                                        double sums[2])
{
    if (!sums) return set_error("NULL argument");
    if (nlive < 2 || (nlive & 1) || ntot < nlive + 1)
        return set_error("nlive = %ld must be even, at least 2 and below "
                         "ntot = %ld", nlive, ntot);
    if (nlive > IHG_NOALLOC_MAX_LIVE)
        return set_error("nlive = %ld beyond the maximum %ld of "
                         "ihg_dp853_finish_noalloc", nlive,
                         (long)IHG_NOALLOC_MAX_LIVE);
    FinishIn in{};
    IHG_TRY(finish_inputs(y, k, y_new, in));
    double *part = nullptr;
    IHG_TRY(scratch_of_current_device(&part));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the context's work-groups (one element pair per thread) when the
    // scratch holds their partials -- the same partition of the sums, so the
    // same values -- and FINISH_MAX_GROUPS striding work-groups beyond
    const long wide = finish_groups_wide(nlive);
    const FinishPlan fp =
        finish_plan(wide <= FINISH_MAX_GROUPS ? wide : FINISH_MAX_GROUPS);
    // every partial the sums read is written by the kernels before, so the
    // scratch needs no clearing between calls
    IHG_TRY(launch_finish(s, fp, y, in, y_new, dt, rtol, atol, nlive, ntot,
                          part));
    IHG_TRY(launch_finish_sums(s, fp, part));
    IHG_HIP(hipMemcpyAsync(sums, part + fp.doubles() - 2, 2 * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    IHG_HIP(hipStreamSynchronize(s));
    return 0;
}

extern "C" int ihg_diff_norms(ihg_ctx *c, const double *a, const double *b,
                              const double *cc, const double *d, size_t n,
                              double out[2])
{
    if (!c || !a || !b || !cc || !d || !out) return set_error("NULL argument");
    if (n < 1) return set_error("empty arrays");
    long nb = ((long)n + RED_WG * 4 - 1) / (RED_WG * 4);
    if (nb > 4096) nb = 4096;
    if (2 * nb + 2 > c->npartials) return set_error("partials buffer too small");
    hipLaunchKernelGGL(k_diff_norms, dim3((unsigned)nb), dim3(RED_WG), 0,
                       c->stream, a, b, cc, d, (long)n, c->partials);
    IHG_LAUNCHED("k_diff_norms");
    hipLaunchKernelGGL(k_sum2_final, dim3(1), dim3(RED_WG), 0, c->stream,
                       c->partials, (int)nb, c->dscal + 8);
    IHG_LAUNCHED("k_sum2_final");
    IHG_HIP(hipMemcpyAsync(c->hscal + 8, c->dscal + 8, 2 * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    IHG_HIP(hipStreamSynchronize(c->stream));
    out[0] = c->hscal[8];
    out[1] = c->hscal[9];
    return 0;
}

extern "C" int ihg_dp853_tableau(double *a, double *cc, double *b, double *bhh,
                                 double *er)
{
    if (!a || !cc || !b || !bhh || !er) return set_error("NULL argument");
    for (int s = 0; s < 13; ++s)
        for (int j = 0; j < 12; ++j)
            a[s * 12 + j] = (s < 12) ? DP_A[s][j] : DP_B[j];
    for (int s = 0; s < 12; ++s)
        cc[s] = DP_C[s];
    cc[12] = 1.0;
    for (int j = 0; j < 12; ++j) {
        b[j] = DP_B[j];
        er[j] = DP_ER[j];
    }
    for (int j = 0; j < 3; ++j)
        bhh[j] = DP_BHH[j];
    return 0;
}

extern "C" int ihg_gsl_tableau(int method, int info[3], double *c, double *a,
                               double *b, double *w)
{
    const GslTableau *t = gsl_tableau_of(method);
    if (!t) return set_error("GSL stepper %d is not 1 (RKF45), 2 (RKCK) or 3 (DOPRI89)",
                             method);
    if (!info || !c || !a || !b || !w) return set_error("NULL argument");
    info[0] = t->stages;
    info[1] = t->order;
    info[2] = t->err_mode;
    for (int s = 0; s < GSL_MAX_STAGES; ++s) {
        c[s] = s < t->stages ? t->c[s] : 0.0;
        b[s] = s < t->stages ? t->b[s] : 0.0;
        w[s] = s < t->stages ? t->w[s] : 0.0;
        for (int j = 0; j < GSL_MAX_STAGES - 1; ++j)
            a[s * (GSL_MAX_STAGES - 1) + j] = s < t->stages ? t->a[s][j] : 0.0;
    }
    return 0;
}

extern "C" int ihg_gsl_stage(ihg_ctx *c, int method, int s, double dt,
                             const double *y, const double *const *k,
                             double *out)
{
    const GslTableau *t = gsl_tableau_of(method);
    if (!c || !k) return set_error("NULL argument");
    if (!t) return set_error("GSL stepper %d is not 1, 2 or 3", method);
    if (s < 2 || s > t->stages)
        return set_error("stage %d out of range 2..%d", s, t->stages);
    IHG_TRY(check_ptr(y, "y"));
    IHG_TRY(check_ptr(out, "out"));
    StageIn in{};
    int n = 0;
    for (int j = 0; j < s - 1; ++j) {
        if (t->a[s - 1][j] == 0.0) continue;
        IHG_TRY(check_ptr(k[j], "k[j]"));
        in.k[n] = k[j];
        in.a[n] = t->a[s - 1][j];
        ++n;
    }
    if (n < 1) return set_error("stage %d has no weights", s);
    in.nk = n;
    const long nl = live_len(c);
    return timed(c, F_GSL_STAGE, (n + 2.0) * 8.0 * nl, [&]() {
        hipLaunchKernelGGL((streaming(nl) ? k_dp_stage<true>
                                          : k_dp_stage<false>),
                           dim3(stream_blocks(nl / 2)),
                           dim3(RED_WG), 0, c->stream, out, y, in, dt, nl,
                           c->Ntot);
        return hip_check(hipGetLastError(), "k_dp_stage (gsl)");
    });
}

// f(std::integral_constant<int, n>) for the weighted-stage counts of the
// GSL steppers' final sums, n = 1..GSL_MAX_STAGES
static_assert(GSL_MAX_STAGES == 13, "one case per stage count");
template <class F>
static int dispatch_gsl_stages(int n, F &&f)
{
    switch (n) {
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 5: return f(std::integral_constant<int, 5>{});
    case 6: return f(std::integral_constant<int, 6>{});
    case 7: return f(std::integral_constant<int, 7>{});
    case 8: return f(std::integral_constant<int, 8>{});
    case 9: return f(std::integral_constant<int, 9>{});
    case 10: return f(std::integral_constant<int, 10>{});
    case 11: return f(std::integral_constant<int, 11>{});
    case 12: return f(std::integral_constant<int, 12>{});
    case 13: return f(std::integral_constant<int, 13>{});
    default: return set_error("no GSL final sum of %d stages", n);
    }
}

extern "C" int ihg_gsl_finish(ihg_ctx *c, int method, double dt,
                              const double *y, const double *const *k,
                              double *y_new, double eps_abs, double eps_rel,
                              double *rmax)
{
    const GslTableau *t = gsl_tableau_of(method);
    if (!c || !k || !rmax) return set_error("NULL argument");
    if (!t) return set_error("GSL stepper %d is not 1, 2 or 3", method);
    IHG_TRY(check_ptr(y, "y"));
    IHG_TRY(check_ptr(y_new, "y_new"));
    GslFinishIn in{};
    int n = 0;
    for (int j = 0; j < t->stages; ++j) {
        if (t->b[j] == 0.0 && t->w[j] == 0.0) continue;
        IHG_TRY(check_ptr(k[j], "k[j]"));
        in.k[n] = k[j];
        in.b[n] = t->b[j];
        in.w[n] = t->w[j];
        ++n;
    }
    in.nk = n;
    in.err_mode = t->err_mode;
    const long nl = live_len(c);
    // one element pair per thread; the block maxima in part[0, nb), their
    // first level in part[nb, nb + nmid), the maximum at part[nb + nmid]
    const FinishPlan fp = finish_plan(finish_groups_wide(nl));
    IHG_TRY(ensure_fin_partials(c, fp.nb + fp.nmid + 1));
    double *part = c->fin_partials, *mid = part + fp.nb,
           *dst = mid + fp.nmid;
    // y and the weighted stages in, y_new out
    IHG_TRY(timed(c, F_GSL_FINISH, (n + 2.0) * 8.0 * nl, [&]() {
        return dispatch_gsl_stages(n, [&](auto nk) {
            constexpr int NK = decltype(nk)::value;
            hipLaunchKernelGGL((streaming(nl) ? k_gsl_finish<true, NK>
                                              : k_gsl_finish<false, NK>),
                               dim3((unsigned)fp.nb), dim3(RED_WG), 0,
                               c->stream, y, in, y_new, dt, eps_abs, eps_rel,
                               nl, c->Ntot, part);
            return hip_check(hipGetLastError(), "k_gsl_finish");
        });
    }));
    if (fp.nmid > 1) {
        hipLaunchKernelGGL(k_max_mid, dim3((unsigned)fp.nmid), dim3(RED_WG), 0,
                           c->stream, part, fp.nb, fp.chunk, mid);
        IHG_LAUNCHED("k_max_mid");
    }
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(RED_WG), 0, c->stream,
                       fp.nmid > 1 ? mid : part,
                       (int)(fp.nmid > 1 ? fp.nmid : fp.nb), dst);
    IHG_LAUNCHED("k_max_final");
    IHG_HIP(hipMemcpyAsync(c->hscal + 8, dst, sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    IHG_HIP(hipStreamSynchronize(c->stream));
    // the control's running maximum starts at DBL_MIN
    const double dbl_min = 2.2250738585072014e-308;
    *rmax = c->hscal[8] > dbl_min ? c->hscal[8] : dbl_min;
    return 0;
}
