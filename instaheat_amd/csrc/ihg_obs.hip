// ihg_obs.hip -- output_all (src/toolbox.c:251-270) and mk_initial_psi
// (src/setup.c:825-903) on the device.
//
// Observables of an output call land in ctx->dobs, packed as
//   smry[6][4]  phi, dphi, psi, dpsi (of the summary field), rho, pressure:
//               mean, variance, min, max          (mk_summary, :634-690)
//   cstr[4]     <ham^2>, max|ham|, <mom^2>, max|mom|  (mk_constraints, :524)
//   phi_ps, psi_ps, rho_ps, gw [bins each]        (:567-582, :435-484)
#include <cfloat>
#include <cmath>

#include "dispatch.h"
#include "ihg_internal.h"
#include "reduce.h"

namespace ihg {



constexpr int MAX_BINS = 2048;
constexpr unsigned RED_BLOCKS = 512;

struct Ptrs8 {
    const double *p[8];
};

// slot of a reduction result (nullptr: not stored)
struct Dst8 {
    double *mean[8], *min[8], *max[8];
};

static unsigned red_blocks(long n)
{
    long b = (n + RED_WG * 4 - 1) / (RED_WG * 4);
    if (b < 1) b = 1;
    return (unsigned)(b > (long)RED_BLOCKS ? RED_BLOCKS : b);
}

// sum, min, max of each array (mean_var_min_max / mean, src/toolbox.c:676-734)
__global__ __launch_bounds__(RED_WG) void k_stats1(Ptrs8 arrs, long n,
                                                   double *__restrict__ part)
{
    __shared__ double sh[8];
    const double *x = arrs.p[blockIdx.y];
    double s = 0.0, mn = x[0], mx = x[0];
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < n; i += stride) {
        const double v = x[i];
        s += v;
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    s = block_reduce<OpSum>(s, sh, RED_WG);
    mn = block_reduce<OpMin>(mn, sh, RED_WG);
    mx = block_reduce<OpMax>(mx, sh, RED_WG);
    if (threadIdx.x == 0) {
        double *o = part + 3 * ((long)blockIdx.y * gridDim.x + blockIdx.x);
        o[0] = s;
        o[1] = mn;
        o[2] = mx;
    }
}

__global__ __launch_bounds__(RED_WG) void k_stats1_final(
    const double *__restrict__ part, int nb, long n, Dst8 dst)
{
    __shared__ double sh[8];
    const int y = blockIdx.x;
    const double *pp = part + 3L * y * nb;
    double s = 0.0, mn = pp[1], mx = pp[2];
    for (int b = threadIdx.x; b < nb; b += RED_WG) {
        s += pp[3 * b];
        mn = pp[3 * b + 1] < mn ? pp[3 * b + 1] : mn;
        mx = pp[3 * b + 2] > mx ? pp[3 * b + 2] : mx;
    }
    s = block_reduce<OpSum>(s, sh, RED_WG);
    mn = block_reduce<OpMin>(mn, sh, RED_WG);
    mx = block_reduce<OpMax>(mx, sh, RED_WG);
    if (threadIdx.x == 0) {
        if (dst.mean[y]) *dst.mean[y] = s / n;
        if (dst.min[y]) *dst.min[y] = mn;
        if (dst.max[y]) *dst.max[y] = mx;
    }
}

// variance (src/toolbox.c:744-754): sums of d^2 and d, d = f - mean, with the
// mean read from obs[4 y]
__global__ __launch_bounds__(RED_WG) void k_stats2(Ptrs8 arrs, long n,
                                                   const double *__restrict__ obs,
                                                   double *__restrict__ part)
{
    __shared__ double sh[8];
    const double *x = arrs.p[blockIdx.y];
    const double mean = obs[4 * blockIdx.y];
    double s1 = 0.0, s2 = 0.0;
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < n; i += stride) {
        const double d = x[i] - mean;
        s1 += d * d;
        s2 += d;
    }
    s1 = block_reduce<OpSum>(s1, sh, RED_WG);
    s2 = block_reduce<OpSum>(s2, sh, RED_WG);
    if (threadIdx.x == 0) {
        double *o = part + 2 * ((long)blockIdx.y * gridDim.x + blockIdx.x);
        o[0] = s1;
        o[1] = s2;
    }
}

__global__ __launch_bounds__(RED_WG) void k_stats2_final(
    const double *__restrict__ part, int nb, long n, double *__restrict__ obs)
{
    __shared__ double sh[8];
    const int y = blockIdx.x;
    const double *pp = part + 2L * y * nb;
    double s1 = 0.0, s2 = 0.0;
    for (int b = threadIdx.x; b < nb; b += RED_WG) {
        s1 += pp[2 * b];
        s2 += pp[2 * b + 1];
    }
    s1 = block_reduce<OpSum>(s1, sh, RED_WG);
    s2 = block_reduce<OpSum>(s2, sh, RED_WG);
    if (threadIdx.x == 0)
        obs[4 * y + 1] = (s1 - s2 * s2 / n) / (n - 1);
}

// mk_constraints (src/toolbox.c:524-551)
__global__ __launch_bounds__(RED_WG) void k_constraints(
    const double *__restrict__ f, const double *__restrict__ rho,
    const double *__restrict__ lpsi, long N, long Ntot,
    const double *__restrict__ dscal, double *__restrict__ part)
{
    __shared__ double sh[8];
    const double aa = f[Ntot - 1];
    const double a2 = aa * aa;
    const double rho_mean = dscal[S_RHO_MEAN];
    const double hubble = sqrt(rho_mean / 3.0);
    const double h3 = 3.0 * hubble;
    const double phi_mean = dscal[S_PHI_MEAN];
    const double dphi_mean = dscal[S_DPHI_MEAN];
    double hl2 = 0.0, hmx = 0.0, ml2 = 0.0, mmx = 0.0;
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < N; i += stride) {
        const double tmp1 = hubble * f[2 * N + i] + f[3 * N + i];
        const double ham = lpsi[i] / a2 - h3 * tmp1 - 0.5 * (rho[i] - rho_mean);
        const double mom = tmp1 - 0.5 * dphi_mean * (f[i] - phi_mean);
        hl2 += ham * ham;
        ml2 += mom * mom;
        hmx = fmax(hmx, fabs(ham));
        mmx = fmax(mmx, fabs(mom));
    }
    hl2 = block_reduce<OpSum>(hl2, sh, RED_WG);
    hmx = block_reduce<OpMax>(hmx, sh, RED_WG);
    ml2 = block_reduce<OpSum>(ml2, sh, RED_WG);
    mmx = block_reduce<OpMax>(mmx, sh, RED_WG);
    if (threadIdx.x == 0) {
        double *o = part + 4 * blockIdx.x;
        o[0] = hl2;
        o[1] = hmx;
        o[2] = ml2;
        o[3] = mmx;
    }
}

__global__ __launch_bounds__(RED_WG) void k_constraints_final(
    const double *__restrict__ part, int nb, long N, double *__restrict__ cstr)
{
    __shared__ double sh[8];
    double hl2 = 0.0, hmx = 0.0, ml2 = 0.0, mmx = 0.0;
    for (int b = threadIdx.x; b < nb; b += RED_WG) {
        hl2 += part[4 * b];
        hmx = fmax(hmx, part[4 * b + 1]);
        ml2 += part[4 * b + 2];
        mmx = fmax(mmx, part[4 * b + 3]);
    }
    hl2 = block_reduce<OpSum>(hl2, sh, RED_WG);
    hmx = block_reduce<OpMax>(hmx, sh, RED_WG);
    ml2 = block_reduce<OpSum>(ml2, sh, RED_WG);
    mmx = block_reduce<OpMax>(mmx, sh, RED_WG);
    if (threadIdx.x == 0) {
        cstr[0] = hl2 / N;
        cstr[1] = hmx;
        cstr[2] = ml2 / N;
        cstr[3] = mmx;
    }
}

struct Spec3 {
    const double2 *x[3];
};

__device__ __forceinline__ void mode_index(const KGeom &kg, long i, int &ix,
                                           int &iy, int &iz)
{
    const long plane = (long)kg.Ny * kg.Mz;
    ix = (int)(i / plane);
    const long rem = i - (long)ix * plane;
    iy = (int)(rem / kg.Mz);
    iz = (int)(rem - (long)iy * kg.Mz);
}

// bin index of src/toolbox.c:579 / :477, clamped to the histogram (with
// many bins, rounding at the Nyquist corner could otherwise give `bins`);
// with log bins, bins of equal width in log |k| from the fundamental |k|
// to k_max: trunc(bins log(k / k_lo) / log(k_max / k_lo) - 1e-14)
__device__ __forceinline__ int bin_of(double k, const KGeom &kg, int bins)
{
    const double x = kg.log_bins ? bins * log(k / kg.k_lo) / kg.log_span
                                 : bins * k / kg.k_max;
    const int b = (int)trunc(x - 1.0e-14);
    return b < 0 ? 0 : (b >= bins ? bins - 1 : b);
}

// mk_power_spectrum (src/toolbox.c:567-582), per-block histograms in LDS
__global__ __launch_bounds__(RED_WG) void k_power_spectrum(
    Spec3 sp, long M, KGeom kg, int bins, double Nd, double *__restrict__ part)
{
    __shared__ double hist[MAX_BINS];
    for (int b = threadIdx.x; b < bins; b += RED_WG)
        hist[b] = 0.0;
    __syncthreads();
    const double2 *X = sp.x[blockIdx.y];
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < M; i += stride) {
        if (i == 0) continue;
        int ix, iy, iz;
        mode_index(kg, i, ix, iy, iz);
        const double2 v = X[i];
        double p2 = v.x * v.x + v.y * v.y;
        if (fabs(kg.kzd(iz)) > DBL_EPSILON) p2 *= 2.0;
        const int idx = bin_of(sqrt(kg.ksq(ix, iy, iz)), kg, bins);
        atomicAdd(&hist[idx], p2 / Nd);
    }
    __syncthreads();
    double *o = part + ((long)blockIdx.y * gridDim.x + blockIdx.x) * bins;
    for (int b = threadIdx.x; b < bins; b += RED_WG)
        o[b] = hist[b];
}

// mk_gw_spectrum (src/toolbox.c:435-484).  T = double: the reference's
// arithmetic; T = float: the precision form (per-mode power and per-block
// bin sums in FP32; bins from the FP64 |k|, cross-block sums in FP64)
template <class T>
__global__ __launch_bounds__(RED_WG) void k_gw_spectrum(
    const double *__restrict__ f, long N, long Next, long M, KGeom kg, int bins,
    double lxyz, const double *__restrict__ dscal, double *__restrict__ part)
{
    __shared__ T hist[MAX_BINS];
    for (int b = threadIdx.x; b < bins; b += RED_WG)
        hist[b] = 0;
    __syncthreads();
    const long Ndh1 = 4 * N + 2 * Next, Ndh2 = Ndh1 + Next;
    const double hubble = sqrt(dscal[S_RHO_MEAN] / 3.0);
    const double rat = pow(0.01, 1.0 / 3.0);
    const T fac = (T)(M_PI * rat / (3.0 * hubble * hubble * lxyz));
    const T two = 2;
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < M; i += stride) {
        if (i == 0) continue;
        int ix, iy, iz;
        mode_index(kg, i, ix, iy, iz);
        const double kxd = kg.kxf(ix), kyd = kg.kyf(iy), kzd = kg.kzf(iz);
        const double k2d = kg.ksq(ix, iy, iz), kd = sqrt(k2d);
        const T kx = (T)kxd, ky = (T)kyd, kz = (T)kzd;
        const T kx2 = kx * kx, ky2 = ky * ky, kz2 = kz * kz;
        const T k2 = (T)k2d, k = (T)kd;
        const T dh1r = (T)f[Ndh1 + 2 * i], dh2r = (T)f[Ndh2 + 2 * i];
        const T dh1i = (T)f[Ndh1 + 2 * i + 1], dh2i = (T)f[Ndh2 + 2 * i + 1];
        T pw;
        if (fabs(kzd) > DBL_EPSILON) {
            pw = two * k2 / (kz2 * (ky2 + kz2)) *
                 ((kx2 + kz2) * (dh1r * dh1r + dh1i * dh1i) +
                  two * kx * ky * (dh1r * dh2r + dh1i * dh2i) +
                  (ky2 + kz2) * (dh2r * dh2r + dh2i * dh2i));
        } else if (fabs(kyd) > DBL_EPSILON) {
            pw = two * (kx2 + ky2) / (ky2 * ky2) *
                 ((kx2 + ky2) * (dh1r * dh1r + dh1i * dh1i) +
                  ky2 * (dh2r * dh2r + dh2i * dh2i));
        } else {
            pw = two * (dh1r * dh1r + dh1i * dh1i + dh2r * dh2r + dh2i * dh2i);
        }
        if (fabs(kg.kzd(iz)) > DBL_EPSILON) pw *= two;
        atomicAdd(&hist[bin_of(kd, kg, bins)], fac * k2 * k * pw);
    }
    __syncthreads();
    double *o = part + (long)blockIdx.x * bins;
    for (int b = threadIdx.x; b < bins; b += RED_WG)
        o[b] = (double)hist[b];
}

// sum of the per-block histograms in block order; out = sum / div
__global__ __launch_bounds__(RED_WG) void k_hist_final(
    const double *__restrict__ part, int nb, int bins, double div,
    double *__restrict__ out)
{
    const double *pp = part + (long)blockIdx.y * nb * bins;
    double *o = out + (long)blockIdx.y * bins;
    for (int b = blockIdx.x * RED_WG + threadIdx.x; b < bins;
         b += gridDim.x * RED_WG) {
        double s = 0.0;
        for (int k = 0; k < nb; ++k)
            s += pp[(long)k * bins + b];
        o[b] = s / div;
    }
}

// tmp.fc = -psi_hat * k^2 / N (src/toolbox.c:180-182)
__global__ __launch_bounds__(RED_WG) void k_neg_ksq(
    const double2 *__restrict__ in, double2 *__restrict__ out, long M, KGeom kg,
    double Nd)
{
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < M; i += stride) {
        int ix, iy, iz;
        mode_index(kg, i, ix, iy, iz);
        const double ks = kg.ksq(ix, iy, iz);
        const double2 v = in[i];
        out[i] = make_double2(-v.x * ks / Nd, -v.y * ks / Nd);
    }
}

int rhs_output(ihg_ctx *c, const double *f, int kept)
{
    const long N = c->N, M = c->M, Ntot = c->Ntot;
    const int bins = c->p.bins;
    const double *field = c->field ? c->field : f;
    double2 *W = c->W, *Wc = c->Wc;
    double *obs = c->dobs;
    // phi_hat, psi_hat (kept by the filter pass, else of f) and rho_hat
    if (!kept)
        IHG_TRY(fft_r2c(c, f, Wc, 2, 2 * N, M));
    IHG_TRY(fft_r2c(c, c->rho, W, 1, 0, 0));
    // power spectra
    {
        const unsigned nb = red_blocks(M);
        Spec3 sp{{Wc, Wc + M, W}};
        hipLaunchKernelGGL(k_power_spectrum, dim3(nb, 3), dim3(RED_WG), 0,
                           c->stream, sp, M, c->kg, bins, (double)N,
                           c->partials);
        IHG_LAUNCHED("k_power_spectrum");
        hipLaunchKernelGGL(k_hist_final, dim3((bins + RED_WG - 1) / RED_WG, 3),
                           dim3(RED_WG), 0, c->stream, c->partials, (int)nb,
                           bins, 1.0, obs + 28);
        IHG_LAUNCHED("k_hist_final");
    }
    // Laplacian of psi (tmp.f)
    hipLaunchKernelGGL(k_neg_ksq, dim3(red_blocks(M)), dim3(RED_WG), 0,
                       c->stream, Wc + M, W + M, M, c->kg, (double)N);
    IHG_LAUNCHED("k_neg_ksq");
    IHG_TRY(fft_c2r(c, W + M, c->tmpf, W + 2 * M, 1, 0, 0));
    // means / min / max: field phi, dphi, psi, dpsi, rho, pressure, f phi, dphi
    const unsigned nb = red_blocks(N);
    Ptrs8 a8{{field, field + N, field + 2 * N, field + 3 * N, c->rho,
              c->pressure, f, f + N}};
    hipLaunchKernelGGL(k_stats1, dim3(nb, 8), dim3(RED_WG), 0, c->stream, a8,
                       N, c->partials);
    IHG_LAUNCHED("k_stats1");
    Dst8 d{};
    for (int y = 0; y < 6; ++y) {
        d.mean[y] = obs + 4 * y;
        d.min[y] = obs + 4 * y + 2;
        d.max[y] = obs + 4 * y + 3;
    }
    d.mean[6] = c->dscal + S_PHI_MEAN;
    d.mean[7] = c->dscal + S_DPHI_MEAN;
    hipLaunchKernelGGL(k_stats1_final, dim3(8), dim3(RED_WG), 0, c->stream,
                       c->partials, (int)nb, N, d);
    IHG_LAUNCHED("k_stats1_final");
    hipLaunchKernelGGL(k_stats2, dim3(nb, 6), dim3(RED_WG), 0, c->stream, a8,
                       N, obs, c->partials);
    IHG_LAUNCHED("k_stats2");
    hipLaunchKernelGGL(k_stats2_final, dim3(6), dim3(RED_WG), 0, c->stream,
                       c->partials, (int)nb, N, obs);
    IHG_LAUNCHED("k_stats2_final");
    // constraints
    hipLaunchKernelGGL(k_constraints, dim3(nb), dim3(RED_WG), 0, c->stream, f,
                       c->rho, c->tmpf, N, Ntot, c->dscal, c->partials);
    IHG_LAUNCHED("k_constraints");
    hipLaunchKernelGGL(k_constraints_final, dim3(1), dim3(RED_WG), 0,
                       c->stream, c->partials, (int)nb, N, obs + 24);
    IHG_LAUNCHED("k_constraints_final");
    // gravitational wave spectrum
    if (c->p.enable_gw) {
        const unsigned nbm = red_blocks(M);
        const double lxyz = (c->p.xb - c->p.xa) * (c->p.yb - c->p.ya) *
                            (c->p.zb - c->p.za);
        if (c->p.gw_fp32)
            hipLaunchKernelGGL(k_gw_spectrum<float>, dim3(nbm), dim3(RED_WG), 0,
                               c->stream, f, N, c->Next, M, c->kg, bins, lxyz,
                               c->dscal, c->partials);
        else
            hipLaunchKernelGGL(k_gw_spectrum<double>, dim3(nbm), dim3(RED_WG),
                               0, c->stream, f, N, c->Next, M, c->kg, bins,
                               lxyz, c->dscal, c->partials);
        IHG_LAUNCHED("k_gw_spectrum");
        hipLaunchKernelGGL(k_hist_final, dim3((bins + RED_WG - 1) / RED_WG, 1),
                           dim3(RED_WG), 0, c->stream, c->partials, (int)nbm,
                           bins, (double)N, obs + 28 + 3 * bins);
        IHG_LAUNCHED("k_hist_final");
    } else {
        IHG_HIP(hipMemsetAsync(obs + 28 + 3 * bins, 0, bins * sizeof(double),
                               c->stream));
    }
    c->have_obs = 1;
    return 0;
}

// ---------------------------------------------------------------------------
// mk_initial_psi (src/setup.c:825-903)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(RED_WG) void k_ipsi_prep(
    const double *__restrict__ field, double *__restrict__ drho,
    double *__restrict__ fm, long N, const double *__restrict__ dscal)
{
    const double phi_mean = dscal[S_PHI_MEAN], dphi_mean = dscal[S_DPHI_MEAN];
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < N; i += stride) {
        drho[i] = field[N + i] - dphi_mean;
        fm[i] = field[i] - phi_mean;
    }
}

__global__ __launch_bounds__(RED_WG) void k_ipsi_spec(
    const double2 *__restrict__ drc, const double2 *__restrict__ fc,
    double2 *__restrict__ out, long M, KGeom kg, double Nd,
    const double *__restrict__ field, long Ntot,
    const double *__restrict__ dscal)
{
    const double aa = field[Ntot - 1];
    const double a2 = aa * aa;
    const double dphi_mean = dscal[S_DPHI_MEAN], ddphi_mean = dscal[S_DDPHI_MEAN];
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < M; i += stride) {
        if (i == 0) {
            out[0] = make_double2(0.0, 0.0);
            continue;
        }
        int ix, iy, iz;
        mode_index(kg, i, ix, iy, iz);
        const double2 d = drc[i], q = fc[i];
        const double den =
            (-dphi_mean * dphi_mean + 2.0 * kg.ksq(ix, iy, iz) / a2) * Nd;
        out[i] = make_double2((-dphi_mean * d.x + ddphi_mean * q.x) / den,
                              (-dphi_mean * d.y + ddphi_mean * q.y) / den);
    }
}

__global__ __launch_bounds__(RED_WG) void k_ipsi_dpsi(
    double *__restrict__ field, const double *__restrict__ fm, long N,
    const double *__restrict__ dscal)
{
    const double hubble = sqrt(dscal[S_RHO_MEAN] / 3.0);
    const double dphi_mean = dscal[S_DPHI_MEAN];
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < N; i += stride)
        field[3 * N + i] = 0.5 * dphi_mean * fm[i] - hubble * field[2 * N + i];
}

// Step-proposer features (no counterpart in the reference): per-block sums
// of phi, phi^2 and psi^2
__global__ __launch_bounds__(RED_WG) void k_features(const double *__restrict__ f,
                                                     long N,
                                                     double *__restrict__ part)
{
    __shared__ double sh[8];
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < N; i += stride) {
        const double p = f[i], q = f[2 * N + i];
        s1 += p;
        s2 += p * p;
        s3 += q * q;
    }
    s1 = block_reduce<OpSum>(s1, sh, RED_WG);
    s2 = block_reduce<OpSum>(s2, sh, RED_WG);
    s3 = block_reduce<OpSum>(s3, sh, RED_WG);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s1;
        part[3 * blockIdx.x + 1] = s2;
        part[3 * blockIdx.x + 2] = s3;
    }
}

// out = {a, variance of phi, mean of psi^2}
__global__ __launch_bounds__(RED_WG) void k_features_final(
    const double *__restrict__ part, int nb, long N, long Ntot,
    const double *__restrict__ f, double *__restrict__ out)
{
    __shared__ double sh[8];
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int b = threadIdx.x; b < nb; b += RED_WG) {
        s1 += part[3 * b];
        s2 += part[3 * b + 1];
        s3 += part[3 * b + 2];
    }
    s1 = block_reduce<OpSum>(s1, sh, RED_WG);
    s2 = block_reduce<OpSum>(s2, sh, RED_WG);
    s3 = block_reduce<OpSum>(s3, sh, RED_WG);
    if (threadIdx.x == 0) {
        const double var = (s2 - s1 * s1 / (double)N) / (double)(N - 1);
        out[0] = f[Ntot - 1];
        out[1] = var > 0.0 ? var : 0.0;
        out[2] = s3 / (double)N;
    }
}

int features(ihg_ctx *c, const double *f, double out[3])
{
    const unsigned nb = red_blocks(c->N);
    hipLaunchKernelGGL(k_features, dim3(nb), dim3(RED_WG), 0, c->stream, f,
                       c->N, c->partials);
    IHG_LAUNCHED("k_features");
    hipLaunchKernelGGL(k_features_final, dim3(1), dim3(RED_WG), 0, c->stream,
                       c->partials, (int)nb, c->N, c->Ntot, f,
                       c->dscal + S_FEATURES);
    IHG_LAUNCHED("k_features_final");
    IHG_TRY(hip_check(hipMemcpyAsync(c->hscal + S_FEATURES,
                                     c->dscal + S_FEATURES, 3 * sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream),
                      "download features"));
    IHG_TRY(hip_check(hipStreamSynchronize(c->stream), "sync"));
    for (int i = 0; i < 3; ++i) out[i] = c->hscal[S_FEATURES + i];
    return 0;
}

__global__ void k_set_scalar(double *__restrict__ p, double v) { *p = v; }

// mean / variance of one device array with the summaries' reductions
// (mean and variance, src/toolbox.c:726-754); the results land in the
// scalar slots S_STAT (mean) and S_STAT + 1 (variance)
static int array_stat(ihg_ctx *c, const double *d, long n, int var,
                      double mean, double *out)
{
    const unsigned nb = red_blocks(n);
    Ptrs8 a8{{d, d, d, d, d, d, d, d}};
    double *slot = c->dscal + S_STAT;
    if (!var) {
        hipLaunchKernelGGL(k_stats1, dim3(nb, 1), dim3(RED_WG), 0, c->stream,
                           a8, n, c->partials);
        IHG_LAUNCHED("k_stats1");
        Dst8 dst{};
        dst.mean[0] = slot;
        hipLaunchKernelGGL(k_stats1_final, dim3(1), dim3(RED_WG), 0,
                           c->stream, c->partials, (int)nb, n, dst);
        IHG_LAUNCHED("k_stats1_final");
    } else {
        // k_stats2 reads the mean from slot[0] and writes slot[1]
        hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c->stream, slot,
                           mean);
        IHG_LAUNCHED("k_set_scalar");
        hipLaunchKernelGGL(k_stats2, dim3(nb, 1), dim3(RED_WG), 0, c->stream,
                           a8, n, slot, c->partials);
        IHG_LAUNCHED("k_stats2");
        hipLaunchKernelGGL(k_stats2_final, dim3(1), dim3(RED_WG), 0,
                           c->stream, c->partials, (int)nb, n, slot);
        IHG_LAUNCHED("k_stats2_final");
    }
    IHG_TRY(hip_check(hipMemcpyAsync(c->hscal + S_STAT, slot,
                                     2 * sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream),
                      "download statistic"));
    IHG_TRY(hip_check(hipStreamSynchronize(c->stream), "sync"));
    *out = c->hscal[S_STAT + (var ? 1 : 0)];
    return 0;
}

int initial_psi(ihg_ctx *c, double *field, double *dfield)
{
    const long N = c->N, M = c->M;
    // phi_mean, dphi_mean, ddphi_mean
    const unsigned nb = red_blocks(N);
    Ptrs8 a8{{field, field + N, dfield + N, field, field, field, field, field}};
    hipLaunchKernelGGL(k_stats1, dim3(nb, 3), dim3(RED_WG), 0, c->stream, a8,
                       N, c->partials);
    IHG_LAUNCHED("k_stats1");
    Dst8 d{};
    d.mean[0] = c->dscal + S_PHI_MEAN;
    d.mean[1] = c->dscal + S_DPHI_MEAN;
    d.mean[2] = c->dscal + S_DDPHI_MEAN;
    hipLaunchKernelGGL(k_stats1_final, dim3(3), dim3(RED_WG), 0, c->stream,
                       c->partials, (int)nb, N, d);
    IHG_LAUNCHED("k_stats1_final");
    // rho buffer <- dphi - <dphi>, tmpf <- phi - <phi> (rho and tmpf are 2N
    // apart in one allocation, so both go through one batched r2c)
    hipLaunchKernelGGL(k_ipsi_prep, dim3(red_blocks(N)), dim3(RED_WG), 0,
                       c->stream, field, c->rho, c->tmpf, N, c->dscal);
    IHG_LAUNCHED("k_ipsi_prep");
    IHG_TRY(fft_r2c(c, c->rho, c->W, 2, 2 * N, M));
    hipLaunchKernelGGL(k_ipsi_spec, dim3(red_blocks(M)), dim3(RED_WG), 0,
                       c->stream, c->W, c->W + M, c->W + 2 * M, M, c->kg,
                       (double)N, field, c->Ntot, c->dscal);
    IHG_LAUNCHED("k_ipsi_spec");
    IHG_TRY(fft_c2r(c, c->W + 2 * M, field + 2 * N, c->W + 3 * M, 1, 0, 0));
    hipLaunchKernelGGL(k_ipsi_dpsi, dim3(red_blocks(N)), dim3(RED_WG), 0,
                       c->stream, field, c->tmpf, N, c->dscal);
    IHG_LAUNCHED("k_ipsi_dpsi");
    return 0;
}

} // namespace ihg

extern "C" int ihg_mean(ihg_ctx *c, const double *d, size_t n, double *out)
{
    if (!c || !d || !out) return ihg::set_error("NULL argument");
    if (n < 1) return ihg::set_error("mean of an empty array");
    return ihg::array_stat(c, d, (long)n, 0, 0.0, out);
}

extern "C" int ihg_variance(ihg_ctx *c, double mean, const double *d,
                            size_t n, double *out)
{
    if (!c || !d || !out) return ihg::set_error("NULL argument");
    if (n < 2) return ihg::set_error("variance of fewer than two values");
    return ihg::array_stat(c, d, (long)n, 1, mean, out);
}
