// ihg_ic.hip -- seeded Gaussian initial data generated on the device.
//
// The reference fills `field` on the host (src/setup.c:771-790, e.g.
// init_from_bunch_davies with GSL's generator).  Here the recipe is
// counter-based, so every mode's random numbers depend on (seed, mode index)
// only and the whole spectrum is one launch:
//   key    = splitmix64(seed ^ 0x5DEECE66D)
//   u1     = ((splitmix64(key + 2 id)     >> 11) + 1) 2^-53       in (0, 1]
//   u2     = ((splitmix64(key + 2 id + 1) >> 11) + 1) 2^-53 - 2^-53 in [0, 1)
//   omega  = sqrt(k^2 + m^2),  sigma = A sqrt(1 / (2 a0^3 omega))
//   phi_k  = sqrt(-2 ln u1) sigma / sqrt(2) (cos 2 pi u2, sin 2 pi u2)
// for 0 < |k| <= k_cut (zero elsewhere), id = (i Ny + j) Mz + k the index
// of the mode in the r2c half spectrum.  delta phi is the unnormalised c2r
// of that half spectrum (FFTW semantics); phi = phi0 + delta phi,
// dphi = dphi0, psi, dpsi and every h block zero, a = a0.  Products and sums
// are rounded one by one (no contraction), so the host restatement used as
// the test oracle computes the same k^2 and cut decisions.
#include <cmath>
#include <cstdint>

#include "ihg_internal.h"

namespace ihg {

__host__ __device__ inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ inline double unit_open0(uint64_t r)
{
    return __dmul_rn((double)(r >> 11) + 1.0, 1.0 / 9007199254740992.0);
}

struct IcSpec {
    int Nx, Ny, Mz;
    double kx1, ky1, kz1;
    uint64_t key;
    double amplitude, a0, m2, kcut2;
};

__global__ __launch_bounds__(256) void k_ic_spectrum(double2 *spec, long M,
                                                     IcSpec s)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= M) return;
    const int k = (int)(id % s.Mz);
    const long ij = id / s.Mz;
    const int j = (int)(ij % s.Ny), i = (int)(ij / s.Ny);
    const double kx = __dmul_rn(s.kx1, (double)(i > s.Nx / 2 ? i - s.Nx : i));
    const double ky = __dmul_rn(s.ky1, (double)(j > s.Ny / 2 ? j - s.Ny : j));
    const double kz = __dmul_rn(s.kz1, (double)k);
    const double k2 = __dadd_rn(__dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky)),
                                __dmul_rn(kz, kz));
    double2 v = make_double2(0.0, 0.0);
    if (k2 != 0.0 && !(k2 > s.kcut2)) {
        const uint64_t base = s.key + 2 * (uint64_t)id;
        const double u1 = unit_open0(splitmix64(base));
        const double u2 = __dadd_rn(unit_open0(splitmix64(base + 1)), -0x1p-53);
        const double omega = sqrt(__dadd_rn(k2, s.m2));
        const double den = __dmul_rn(__dmul_rn(__dmul_rn(__dmul_rn(2.0, s.a0), s.a0),
                                               s.a0), omega);
        const double sigma = __dmul_rn(s.amplitude, sqrt(1.0 / den));
        const double r = __dmul_rn(sqrt(__dmul_rn(-2.0, log(u1))), sigma) / sqrt(2.0);
        const double ph = __dmul_rn(6.283185307179586476925286766559, u2);
        v = make_double2(__dmul_rn(r, cos(ph)), __dmul_rn(r, sin(ph)));
    }
    spec[id] = v;
}

// phi += phi0 over [0, N), dphi = dphi0, every later element zero, a = a0
__global__ __launch_bounds__(256) void k_ic_finish(double *f, long N, long Ntot,
                                                   double phi0, double dphi0,
                                                   double a0)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= Ntot) return;
    double v;
    if (i == Ntot - 1)
        v = a0;
    else if (i < N)
        v = __dadd_rn(f[i], phi0);
    else if (i < 2 * N)
        v = dphi0;
    else
        v = 0.0;
    f[i] = v;
}

} // namespace ihg

using namespace ihg;

extern "C" int ihg_init_gaussian(ihg_ctx *c, double *field, uint64_t seed,
                                 double amplitude, double phi0, double dphi0,
                                 double a0, double k_cut)
{
    if (!c || !field) return set_error("NULL argument");
    if (((size_t)field) & 15) return set_error("field is not 16-byte aligned");
    if (!(amplitude >= 0.0) || !(a0 > 0.0) || !(k_cut >= 0.0))
        return set_error("ihg_init_gaussian: amplitude and k_cut must be >= 0, a0 > 0");
    const ihg_params &p = c->p;
    const double TWOPI = 6.283185307179586476925286766559;
    IcSpec s;
    s.Nx = c->Nx;
    s.Ny = c->Ny;
    s.Mz = c->Mz;
    s.kx1 = TWOPI / (p.xb - p.xa);
    s.ky1 = TWOPI / (p.yb - p.ya);
    s.kz1 = TWOPI / (p.zb - p.za);
    s.key = splitmix64(seed ^ 0x5DEECE66Dull);
    s.amplitude = amplitude;
    s.a0 = a0;
    s.m2 = p.mass * p.mass;
    s.kcut2 = k_cut * k_cut;
    double2 *spec = c->W, *work = c->W + c->M;
    hipLaunchKernelGGL(k_ic_spectrum, dim3((unsigned)((c->M + 255) / 256)),
                       dim3(256), 0, c->stream, spec, c->M, s);
    IHG_TRY(hip_check(hipGetLastError(), "k_ic_spectrum"));
    IHG_TRY(fft_c2r(c, spec, field, work, 1, 0, 0));
    hipLaunchKernelGGL(k_ic_finish, dim3((unsigned)((c->Ntot + 255) / 256)),
                       dim3(256), 0, c->stream, field, c->N, c->Ntot, phi0,
                       dphi0, a0);
    return hip_check(hipGetLastError(), "k_ic_finish");
}
