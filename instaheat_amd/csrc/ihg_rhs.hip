// ihg_rhs.hip -- the right-hand side mk_rhs (src/toolbox.c:75-626) as fused
// HIP kernels.
//
// Data flow of one call (no filter, no output, GW off):
//   z r2c pass      phi                  -> W4            (real -> half rows)
//   y pass          W4                   -> W4            (forward)
//   x fan-out       W4 -> W0..W3: forward x, times i k_x/N, i k_y/N, i k_z/N,
//                   -k^2/N, backward x                    (toolbox.c:132-139)
//   y pass          W0..W3 backward
//   z c2r + RHS     W0..W3 -> d_i phi, lap phi in LDS; rho, p, block sums;
//                   writes lap -> result[N..2N), p -> result[3N..4N)
//                                                         (toolbox.c:140-243)
//   means           block sums -> rho_mean, pressure_mean, da   (:87, :241)
//   update          result[0..4N) from f, lap, p, means   (toolbox.c:282-299)
// GW on adds: z r2c of the six products d_i phi d_j phi (pairs in one complex
// row), y and x passes, and one pass that does the TT projection and the h
// update (toolbox.c:311-424).
#include <cfloat>
#include <cmath>

#include "dispatch.h"
#include "fft_kernels.h"
#include "ihg_internal.h"
#include "reduce.h"

namespace ihg {

// scalar slots in ctx->dscal
enum {
    S_RHO_MEAN = 0,
    S_P_MEAN = 1,
    S_PHI_MEAN = 2,  // mean of f[0..N) (constraints, psi init)
    S_DPHI_MEAN = 3, // mean of f[N..2N)
    S_DDPHI_MEAN = 4,
    S_NSCAL = 16
};

static inline unsigned nblocks(long n, int per_block, unsigned cap)
{
    long b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    return (unsigned)(b > (long)cap ? cap : b);
}

// ---------------------------------------------------------------------------
// spectral derivative factors (mk_gradient_squared_and_laplacian,
// src/toolbox.c:132-139): pre = phi_hat / N; outputs pre * i * k_x,
// pre * i * k_y, pre * i * k_z (Nyquist-zeroed) and -pre * k^2
// ---------------------------------------------------------------------------
struct DerivFac {
    KGeom kg;
    double Nd;
    __device__ double2 operator()(int q, long line, int n, double2 v) const
    {
        const int iy = (int)(line / kg.Mz), iz = (int)(line - (long)iy * kg.Mz);
        const double2 pre = make_double2(v.x / Nd, v.y / Nd);
        if (q == 3) {
            const double ks = kg.ksq(n, iy, iz);
            return make_double2(-(pre.x * ks), -(pre.y * ks));
        }
        const double k = (q == 0) ? kg.kxd(n) : ((q == 1) ? kg.kyd(iy) : kg.kzd(iz));
        return make_double2(-pre.y * k, pre.x * k);
    }
};

// capply_filter (src/toolbox.c:616-626): in *= filter; out = in / N
struct FilterFac {
    const double *__restrict__ filt;
    long plane; // Ny * Mz
    double Nd;
    __device__ double2 operator()(int, long line, int n, double2 v) const
    {
        // padding lines of the last work-group (line >= plane) read entry 0
        const double F = filt[(long)n * plane + (line < plane ? line : 0)];
        const double2 w = make_double2(v.x * F, v.y * F);
        return make_double2(w.x / Nd, w.y / Nd);
    }
};

// ---------------------------------------------------------------------------
// z c2r of the four derivative spectra fused with assemble_gradient_squared
// and mk_rho_and_p (src/toolbox.c:196-243).  Line 2r holds (d_x, d_y) of row
// r, line 2r+1 holds (d_z, lap).
// ---------------------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(FFT_WG) void k_z_c2r_rho(
    const double2 *__restrict__ W, long M, const double *__restrict__ f,
    double *__restrict__ result, double *__restrict__ grad3,
    double *__restrict__ rho, double *__restrict__ pressure,
    double *__restrict__ partials, long nrows, long N, double m2,
    const double2 *__restrict__ tw)
{
    using TL = Tile<L>;
    constexpr int MZ = L / 2 + 1;
    constexpr int RPB = TL::B / 2; // rows per block
    __shared__ double2 lds[TL::B * TL::PITCH];
    const int tid = threadIdx.x;
    const int bc = tid / TL::T;
    const long line = (long)blockIdx.x * TL::B + bc;
    const bool act = line < 2 * nrows;
    const long row = act ? (line >> 1) : 0;
    const int p = (int)(line & 1);
    const double2 *A = W + (2 * p) * M + row * MZ;
    const double2 *Bq = W + (2 * p + 1) * M + row * MZ;
    auto rd = [&](int, int n) {
        const double2 z = herm_merge<L>(A, Bq, n); // in bounds for every lane
        return act ? z : make_double2(0.0, 0.0);
    };
    fft_line<L, +1, true>(lds + bc * TL::PITCH, tid % TL::T, tw, rd);

    const long Ntot = 4 * N + 4 * (2 * M) + 1;
    const double aa = f[Ntot - 1];
    const double a2 = aa * aa;
    double srho = 0.0, sp = 0.0;
    const long row0 = (long)blockIdx.x * RPB;
    for (int idx = tid; idx < RPB * L; idx += FFT_WG) {
        const int rr = idx / L, n = idx - rr * L;
        const long r = row0 + rr;
        if (r >= nrows) continue;
        const double2 z1 = lds[(2 * rr) * TL::PITCH + Lds<L>::idx(n)];
        const double2 z2 = lds[(2 * rr + 1) * TL::PITCH + Lds<L>::idx(n)];
        const double dx = z1.x, dy = z1.y, dz = z2.x, lap = z2.y;
        double g = dx * dx;
        g += dy * dy;
        g += dz * dz;
        const long i = r * L + n;
        const double ph = f[i], dph = f[N + i], ps = f[2 * N + i];
        const double t1 = (0.5 - ps) * dph * dph;
        const double t2 = (0.5 + ps) * g / a2;
        const double V = m2 * ph * ph / 2.0;
        const double rv = t1 + t2 + V;
        const double pv = t1 - t2 / 3.0 - V;
        result[N + i] = lap;
        result[3 * N + i] = pv;
        if (grad3) {
            grad3[i] = dx;
            grad3[N + i] = dy;
            grad3[2 * N + i] = dz;
        }
        if (rho) {
            rho[i] = rv;
            pressure[i] = pv;
        }
        srho += rv;
        sp += pv;
    }
    double *sh = reinterpret_cast<double *>(lds);
    srho = block_reduce<OpSum>(srho, sh, FFT_WG);
    sp = block_reduce<OpSum>(sp, sh + 8, FFT_WG);
    if (tid == 0) {
        partials[2 * blockIdx.x] = srho;
        partials[2 * blockIdx.x + 1] = sp;
    }
}

// rho_mean, pressure_mean (src/toolbox.c:241-242) and da (src/toolbox.c:87)
__global__ __launch_bounds__(RED_WG) void k_means(
    const double *__restrict__ partials, int nparts, long N,
    const double *__restrict__ f, double *__restrict__ result, long Ntot,
    double *__restrict__ dscal)
{
    __shared__ double sh[8];
    double s0 = 0.0, s1 = 0.0;
    for (int i = threadIdx.x; i < nparts; i += RED_WG) {
        s0 += partials[2 * i];
        s1 += partials[2 * i + 1];
    }
    s0 = block_reduce<OpSum>(s0, sh, RED_WG);
    s1 = block_reduce<OpSum>(s1, sh, RED_WG);
    if (threadIdx.x == 0) {
        const double rm = s0 / N, pm = s1 / N;
        dscal[S_RHO_MEAN] = rm;
        dscal[S_P_MEAN] = pm;
        result[Ntot - 1] = f[Ntot - 1] * sqrt(rm / 3.0);
    }
}

// update_phi_psi (src/toolbox.c:282-299); lap and p were parked in
// result[N..2N) and result[3N..4N) by k_z_c2r_rho
__global__ __launch_bounds__(RED_WG) void k_update_phi_psi(
    const double *__restrict__ f, double *__restrict__ result, long N,
    long Ntot, double m2, const double *__restrict__ dscal)
{
    const double aa = f[Ntot - 1];
    const double a2 = aa * aa;
    const double hubble = sqrt(dscal[S_RHO_MEAN] / 3.0);
    const double h3 = 3.0 * hubble;
    const double pm = dscal[S_P_MEAN];
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < N; i += stride) {
        const double ph = f[i], dph = f[N + i], ps = f[2 * N + i],
                     dps = f[3 * N + i];
        const double lap = result[N + i], pr = result[3 * N + i];
        result[i] = dph;
        result[N + i] = (1.0 + 4.0 * ps) * lap / a2 - (h3 - 4.0 * dps) * dph -
                        (1.0 + 2.0 * ps) * (m2 * ph);
        result[2 * N + i] = dps;
        result[3 * N + i] = 0.5 * pr + (ps - 0.5) * pm - 4.0 * hubble * dps;
    }
}

// ---------------------------------------------------------------------------
// GW sources: z r2c of the six products d_i phi d_j phi (mk_gw_sources,
// src/toolbox.c:410-424).  A work-group owns B rows: every thread loads its
// first-stage points of d_x phi, d_y phi, d_z phi once into registers, then
// the three complex rows (S11, S12), (S13, S22), (S23, S33) of each row are
// transformed one after the other; component c = 2q + s goes to W + c M.
// ---------------------------------------------------------------------------
template <int L, int Q, class Rows>
__device__ __forceinline__ void sources_pair(double2 *lds, int tid,
                                             const Rows &g, bool act,
                                             double2 *__restrict__ W, long M,
                                             long row0, long nrows,
                                             const double2 *__restrict__ tw)
{
    using TL = Tile<L>;
    constexpr int MZ = L / 2 + 1;
    const int bc = tid / TL::T;
    auto rd = [&](int s, int) {
        double u, v;
        if constexpr (Q == 0) {
            u = g.x[s] * g.x[s];
            v = g.x[s] * g.y[s];
        } else if constexpr (Q == 1) {
            u = g.x[s] * g.z[s];
            v = g.y[s] * g.y[s];
        } else {
            u = g.y[s] * g.z[s];
            v = g.z[s] * g.z[s];
        }
        return act ? make_double2(u, v) : make_double2(0.0, 0.0);
    };
    fft_line<L, -1, true>(lds + bc * TL::PITCH, tid % TL::T, tw, rd);
    for (int idx = tid; idx < TL::B * 2 * MZ; idx += FFT_WG) {
        const int lb = idx / (2 * MZ);
        const int rem = idx - lb * 2 * MZ;
        const int s = rem / MZ, k = rem - s * MZ;
        const long r = row0 + lb;
        if (r < nrows)
            W[(2 * Q + s) * M + r * MZ + k] =
                herm_split<L>(lds + lb * TL::PITCH, k, s);
    }
    __syncthreads(); // split reads done before the next transform
}

template <int L>
struct GradRegs {
    double x[Plan<L>::E], y[Plan<L>::E], z[Plan<L>::E];
};

template <int L>
__global__ __launch_bounds__(FFT_WG) void k_z_r2c_sources(
    const double *__restrict__ g3, double2 *__restrict__ W, long M,
    long nrows, long N, const double2 *__restrict__ tw)
{
    using TL = Tile<L>;
    __shared__ double2 lds[TL::B * TL::PITCH];
    const int tid = threadIdx.x;
    const int bc = tid / TL::T, j = tid % TL::T;
    const long row0 = (long)blockIdx.x * TL::B;
    const long row = row0 + bc;
    const bool act = row < nrows;
    const double *gx = g3 + (act ? row : 0) * L;
    GradRegs<L> g;
#pragma unroll
    for (int s = 0; s < TL::E; ++s) {
        const int n = stage1_point<L>(j, s);
        g.x[s] = gx[n];
        g.y[s] = gx[N + n];
        g.z[s] = gx[2 * N + n];
    }
    sources_pair<L, 0>(lds, tid, g, act, W, M, row0, nrows, tw);
    sources_pair<L, 1>(lds, tid, g, act, W, M, row0, nrows, tw);
    sources_pair<L, 2>(lds, tid, g, act, W, M, row0, nrows, tw);
}

__device__ __forceinline__ double2 rc(double a, double2 z)
{
    return make_double2(a * z.x, a * z.y);
}

// mk_gw_sources_tt (src/toolbox.c:363-397) fused with update_h
// (src/toolbox.c:311-346)
__global__ __launch_bounds__(RED_WG) void k_tt_update_h(
    const double2 *__restrict__ W, long M, const double *__restrict__ f,
    double *__restrict__ result, long N, long Next, long Ntot, KGeom kg,
    const double *__restrict__ dscal)
{
    const double aa = f[Ntot - 1];
    const double a2 = aa * aa;
    const double hubble = sqrt(dscal[S_RHO_MEAN] / 3.0);
    const double h3 = 3.0 * hubble;
    const double pm = dscal[S_P_MEAN];
    const long Nh1 = 4 * N, Nh2 = Nh1 + Next, Ndh1 = Nh2 + Next,
               Ndh2 = Ndh1 + Next;
    const double2 *f2 = reinterpret_cast<const double2 *>(f);
    double2 *r2 = reinterpret_cast<double2 *>(result);
    const long plane = (long)kg.Ny * kg.Mz;
    const long stride = (long)gridDim.x * RED_WG;
    for (long i = (long)blockIdx.x * RED_WG + threadIdx.x; i < M; i += stride) {
        const int ix = (int)(i / plane);
        const long rem = i - (long)ix * plane;
        const int iy = (int)(rem / kg.Mz), iz = (int)(rem - (long)iy * kg.Mz);
        double2 o0 = make_double2(0.0, 0.0), o1 = make_double2(0.0, 0.0);
        const double ksq = kg.ksq(ix, iy, iz);
        if (i != 0) {
            double2 s[6];
#pragma unroll
            for (int cc = 0; cc < 6; ++cc)
                s[cc] = W[cc * M + i];
            const double kx = kg.kxf(ix), ky = kg.kyf(iy), kz = kg.kzf(iz);
            const double fx = kx / ksq, fy = ky / ksq, fz = kz / ksq;
            const double2 t1 = cadd(
                cadd(cadd(rc(kx * fx, s[0]), rc(ky * fy, s[3])), rc(kz * fz, s[5])),
                rc(2.0, cadd(cadd(rc(kx * fy, s[1]), rc(kx * fz, s[2])),
                             rc(ky * fz, s[4]))));
            const double2 t2 = cadd(cadd(s[0], s[3]), s[5]);
            const double2 s1 = cadd(t1, t2), s2 = csub(t1, t2);
            if (fabs(kz) > DBL_EPSILON) {
                const double2 k1 = cadd(cadd(rc(kx, s[0]), rc(ky, s[1])), rc(kz, s[2]));
                const double2 k2 = cadd(cadd(rc(kx, s[1]), rc(ky, s[3])), rc(kz, s[4]));
                o0 = cadd(csub(s[0], rc(2.0 * fx, k1)),
                          rc(0.5, cadd(rc(fx * kx, s1), s2)));
                o1 = cadd(csub(csub(s[1], rc(fx, k2)), rc(fy, k1)),
                          rc(0.5 * fx * ky, s1));
            } else if (fabs(ky) > DBL_EPSILON) {
                const double2 k1 = cadd(cadd(rc(kx, s[0]), rc(ky, s[1])), rc(kz, s[2]));
                const double2 k3 = cadd(cadd(rc(kx, s[2]), rc(ky, s[4])), rc(kz, s[5]));
                o0 = cadd(csub(s[0], rc(2.0 * fx, k1)),
                          rc(0.5, cadd(rc(fx * kx, s1), s2)));
                o1 = cadd(csub(csub(s[2], rc(fx, k3)), rc(fz, k1)),
                          rc(0.5 * fx * kz, s1));
            } else {
                const double2 k2 = cadd(cadd(rc(kx, s[1]), rc(ky, s[3])), rc(kz, s[4]));
                const double2 k3 = cadd(cadd(rc(kx, s[2]), rc(ky, s[4])), rc(kz, s[5]));
                o0 = cadd(csub(s[3], rc(2.0 * fy, k2)),
                          rc(0.5, cadd(rc(fy * ky, s1), s2)));
                o1 = cadd(csub(csub(s[4], rc(fy, k3)), rc(fz, k2)),
                          rc(0.5 * fy * kz, s1));
            }
        }
        const double2 h1 = f2[Nh1 / 2 + i], h2 = f2[Nh2 / 2 + i];
        const double2 d1 = f2[Ndh1 / 2 + i], d2 = f2[Ndh2 / 2 + i];
        r2[Nh1 / 2 + i] = d1;
        r2[Nh2 / 2 + i] = d2;
        const double tmp = -(2.0 * pm + ksq / a2);
        r2[Ndh1 / 2 + i] = make_double2(tmp * h1.x - h3 * d1.x + 2.0 * o0.x / a2,
                                        tmp * h1.y - h3 * d1.y + 2.0 * o0.y / a2);
        r2[Ndh2 / 2 + i] = make_double2(tmp * h2.x - h3 * d2.x + 2.0 * o1.x / a2,
                                        tmp * h2.y - h3 * d2.y + 2.0 * o1.y / a2);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static LineSet x_lineset(const ihg_ctx *c)
{
    const long plane = (long)c->Ny * c->Mz;
    return LineSet{plane, plane, 0, plane};
}

int rhs_core(ihg_ctx *c, double *f, double *result, int output)
{
    const long N = c->N, M = c->M;
    const double Nd = (double)N;
    const double m2 = c->p.mass * c->p.mass;
    double2 *W = c->W;
    // phi_hat (partially: z and y passes) into W4
    IHG_TRY(z_r2c_pass(c, f, W + 4 * M, 1, 0, 0));
    IHG_TRY(y_pass(c, W + 4 * M, -1, 1, 0));
    // x fan-out into the four derivative spectra W0..W3
    IHG_TRY(dispatch_len(c->Nx, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        using TL = Tile<L>;
        const LineSet ls = x_lineset(c);
        DerivFac fac{c->kg, Nd};
        return timed(c, F_X_FANOUT, 5.0 * 16.0 * M, [&]() {
            hipLaunchKernelGGL((k_x_fanout<L, 4, DerivFac>),
                               dim3((unsigned)((ls.nlines + TL::B - 1) / TL::B)),
                               dim3(FFT_WG), 0, c->stream, W + 4 * M, W, M, ls,
                               0L, c->tw, fac);
            return hip_check(hipGetLastError(), "k_x_fanout");
        });
    }));
    IHG_TRY(y_pass(c, W, +1, 4, M));
    const long nrows = (long)c->Nx * c->Ny;
    const bool gw = c->p.enable_gw != 0;
    int nparts = 0;
    IHG_TRY(dispatch_len(c->Nz, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        using TL = Tile<L>;
        const unsigned nb = (unsigned)((2 * nrows + TL::B - 1) / TL::B);
        nparts = (int)nb;
        if (2L * nb > c->npartials)
            return set_error("partials buffer too small");
        // 4 spectra in; phi, dphi, psi in; lap, p out (+ gradients, rho, p)
        const double bytes = 4.0 * 16.0 * M +
                             8.0 * N * (5 + (gw ? 3 : 0) + (output ? 2 : 0));
        return timed(c, F_Z_C2R_RHO, bytes, [&]() {
            hipLaunchKernelGGL((k_z_c2r_rho<L>), dim3(nb), dim3(FFT_WG), 0,
                               c->stream, W, M, f, result,
                               gw ? c->grad3 : nullptr,
                               output ? c->rho : nullptr,
                               output ? c->pressure : nullptr, c->partials,
                               nrows, N, m2, c->tw);
            return hip_check(hipGetLastError(), "k_z_c2r_rho");
        });
    }));
    IHG_TRY(timed(c, F_MEANS, 16.0 * nparts, [&]() {
        hipLaunchKernelGGL(k_means, dim3(1), dim3(RED_WG), 0, c->stream,
                           c->partials, nparts, N, f, result, c->Ntot,
                           c->dscal);
        return hip_check(hipGetLastError(), "k_means");
    }));
    return timed(c, F_UPDATE, 10.0 * 8.0 * N, [&]() {
        hipLaunchKernelGGL(k_update_phi_psi,
                           dim3(nblocks(N, RED_WG * 4, 65536)), dim3(RED_WG),
                           0, c->stream, f, result, N, c->Ntot, m2, c->dscal);
        return hip_check(hipGetLastError(), "k_update_phi_psi");
    });
}

int rhs_gw(ihg_ctx *c, const double *f, double *result)
{
    const long N = c->N, M = c->M;
    const long nrows = (long)c->Nx * c->Ny;
    double2 *W = c->W;
    IHG_TRY(dispatch_len(c->Nz, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        using TL = Tile<L>;
        const unsigned nb = (unsigned)((nrows + TL::B - 1) / TL::B);
        return timed(c, F_Z_R2C_SRC, 3.0 * 8.0 * N + 6.0 * 16.0 * M, [&]() {
            hipLaunchKernelGGL((k_z_r2c_sources<L>), dim3(nb), dim3(FFT_WG),
                               0, c->stream, c->grad3, W, M, nrows, N, c->tw);
            return hip_check(hipGetLastError(), "k_z_r2c_sources");
        });
    }));
    IHG_TRY(y_pass(c, W, -1, 6, M));
    IHG_TRY(x_pass(c, W, -1, 6, M));
    // six sources and h1, h2, dh1, dh2 in; the four h blocks of result out
    return timed(c, F_TT_UPDATE_H, 14.0 * 16.0 * M, [&]() {
        hipLaunchKernelGGL(k_tt_update_h, dim3(nblocks(M, RED_WG * 2, 65536)),
                           dim3(RED_WG), 0, c->stream, W, M, f, result, N,
                           c->Next, c->Ntot, c->kg, c->dscal);
        return hip_check(hipGetLastError(), "k_tt_update_h");
    });
}

// mk_ffts_and_filter's filter branch (src/toolbox.c:164-171): phi, dphi,
// psi, dpsi (contiguous in f) filtered in place, all four in one batch
int rhs_filter(ihg_ctx *c, double *f)
{
    const long N = c->N, M = c->M;
    double2 *W = c->W;
    IHG_TRY(z_r2c_pass(c, f, W, 4, N, M));
    IHG_TRY(y_pass(c, W, -1, 4, M));
    IHG_TRY(dispatch_len(c->Nx, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        using TL = Tile<L>;
        const LineSet ls = x_lineset(c);
        FilterFac fac{c->filter, (long)c->Ny * c->Mz, (double)N};
        return timed(c, F_X_ROUNDTRIP, 4.0 * 2.0 * 16.0 * M, [&]() {
            hipLaunchKernelGGL((k_x_fanout<L, 1, FilterFac>),
                               dim3((unsigned)((ls.nlines + TL::B - 1) / TL::B), 4),
                               dim3(FFT_WG), 0, c->stream, W, W, 0L, ls, M,
                               c->tw, fac);
            return hip_check(hipGetLastError(), "k_x_fanout (filter)");
        });
    }));
    IHG_TRY(y_pass(c, W, +1, 4, M));
    return z_c2r_pass(c, W, f, 4, M, N);
}

} // namespace ihg
