// rhs_kernels.h -- the kernels of the right-hand side mk_rhs
// (src/toolbox.c:75-626), each a device body that takes its work-group
// coordinates and its LDS from the caller, and a thin kernel that runs it
// on the hardware's coordinates.  The grouped launches (ihg_group.hip) run
// the same bodies over a table of segments, so a grouped launch computes
// what the per-size launches compute, value for value.
#pragma once

#include <cfloat>
#include <cmath>
#include <type_traits>

#include "fft_kernels.h"
#include "ihg_internal.h"
#include "mem.h"
#include "reduce.h"
#include "rk4_kernels.h"

namespace ihg {

// ---------------------------------------------------------------------------
// spectral derivative factors (mk_gradient_squared_and_laplacian,
// src/toolbox.c:132-139): pre = phi_hat / N; outputs pre * i * k_x,
// pre * i * k_y, pre * i * k_z (Nyquist-zeroed) and -pre * k^2
// ---------------------------------------------------------------------------
struct DerivFac {
    KGeom kg;
    double Nd;
    __device__ double2 operator()(Blk, int q, long line, int n,
                                  double2 v) const
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

// capply_filter (src/toolbox.c:616-626): in *= filter; out = in / N.
// The filtered spectra in *= filter of phi (batch 0) and psi (batch 2) are
// kept, as mk_ffts_and_filter keeps tmp.phic / tmp.psic
// (src/toolbox.c:164-171): phi's in keep_phi (and keep_phi2), psi's in
// keep_psi; null pointers keep nothing.
struct FilterFac {
    const double *__restrict__ filt; // Fx[Nx] | Fy[Ny] | Fz[Mz]
    long plane; // Ny * Mz
    int Nx, Ny, Mz;
    double Nd;
    double2 *keep_phi, *keep_phi2, *keep_psi;
    long rkeep; // realization stride of keep_phi (ensembles)
    __device__ double2 operator()(Blk bk, int, long line, int n,
                                  double2 v) const
    {
        const bool in = line < plane; // padding lines read entry 0
        const long lc = in ? line : 0;
        const long i = (long)n * plane + lc;
        // the mask at (n, j, k) of line j Mz + k: the smallest axis window
        const int j = (int)(lc / Mz), k = (int)(lc - (long)j * Mz);
        const double F = fmin(fmin(filt[n], filt[Nx + j]), filt[Nx + Ny + k]);
        const double2 w = make_double2(v.x * F, v.y * F);
        if (in && bk.y == 0 && keep_phi) {
            keep_phi[bk.z * rkeep + i] = w;
            if (keep_phi2) keep_phi2[i] = w;
        }
        if (in && bk.y == 2 && keep_psi)
            keep_psi[i] = w;
        return make_double2(w.x / Nd, w.y / Nd);
    }
};

// LDS of the z c2r + rho pass in complex values: the tile's lines, then the
// L stage twiddles
template <int L>
__host__ __device__ constexpr int z_rho_lines()
{
    return Tile<L>::B * Tile<L>::PITCH;
}
template <int L>
__host__ __device__ constexpr int z_rho_lds()
{
    return z_rho_lines<L>() + L;
}

// ---------------------------------------------------------------------------
// z c2r of the four derivative spectra fused with assemble_gradient_squared
// and mk_rho_and_p (src/toolbox.c:196-243).  Line 2r holds (d_x, d_y) of row
// r, line 2r+1 holds (d_z, lap).
// ---------------------------------------------------------------------------
template <int L, class G>
__device__ __forceinline__ void z_c2r_rho_body(
    const Blk &bk, double2 *lds, const double2 *__restrict__ W, long M,
    const double *__restrict__ f, double *__restrict__ result,
    G *__restrict__ grad3, double *__restrict__ rho,
    double *__restrict__ pressure, double *__restrict__ partials, long nrows,
    long N, double m2, const double2 *__restrict__ tw, const Ens &e)
{
    // realization z of an ensemble
    W += bk.z * e.w;
    f += bk.z * e.f;
    result += bk.z * e.f;
    partials += bk.z * e.p;
    if (grad3) grad3 += bk.z * e.g;
    using TL = Tile<L>;
    constexpr int MZ = L / 2 + 1;
    constexpr int RPB = TL::B / 2; // rows per block
    const int tid = threadIdx.x;
    const int bc = tid / TL::T;
    const long line = (long)bk.x * TL::B + bc;
    const bool act = line < 2 * nrows;
    const long row = act ? (line >> 1) : 0;
    const int p = (int)(line & 1);
    const double2 *A = W + (2 * p) * M + row * MZ;
    const double2 *Bq = W + (2 * p + 1) * M + row * MZ;
    // stage twiddles from LDS behind the tile's lines, so that the loads of
    // phi, dphi, psi below stay in flight through the whole transform
    tw = twiddles_to_lds_wave<L>(lds + z_rho_lines<L>(), tw);
    auto rd = [&](int, int n) {
        const double2 z = herm_merge<L>(A, Bq, n); // in bounds for every lane
        return act ? z : make_double2(0.0, 0.0);
    };
    // Each wavefront finishes the RPW = LPW / 2 rows of its own lines (line
    // 2 rr holds (d_x, d_y) of row rr, line 2 rr + 1 (d_z, lap)), K = E / 2
    // points per lane.  phi, dphi, psi of those points are loaded before
    // the transform, so their latency hides behind it; rows past the end
    // read row 0 (in bounds, unused)
    constexpr int RPW = TL::LPW / 2;
    static_assert(RPW >= 1, "both lines of a row in one wavefront");
    constexpr int K = RPW * L / 64;
    static_assert(K * 64 == RPW * L, "whole rows per wavefront sweep");
    const int w = tid / 64, lane = tid % 64;
    const long row0 = (long)bk.x * RPB + w * RPW;
    const double2 *wl = lds + w * TL::LPW * TL::PITCH;
    double php[K], dphp[K], psp[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int idx = lane + k * 64;
        const int rr = idx / L, n = idx - rr * L;
        const long r = row0 + rr;
        const long i = (r < nrows ? r : 0) * L + n;
        php[k] = f[i];
        dphp[k] = f[N + i];
        psp[k] = f[2 * N + i];
    }
    fft_line_wave<L, +1, typename TL::P>(lds + bc * TL::PITCH, tid % TL::T, tw, rd);

    const long Ntot = 4 * N + 4 * (2 * M) + 1;
    const double aa = f[Ntot - 1];
    const double a2 = aa * aa;
    double srho = 0.0, sp = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int idx = lane + k * 64;
        const int rr = idx / L, n = idx - rr * L;
        const long r = row0 + rr;
        if (r >= nrows) continue;
        const double2 z1 = wl[(2 * rr) * TL::PITCH + Lds<L>::idx(n)];
        const double2 z2 = wl[(2 * rr + 1) * TL::PITCH + Lds<L>::idx(n)];
        const double dx = z1.x, dy = z1.y, dz = z2.x, lap = z2.y;
        double g = dx * dx;
        g += dy * dy;
        g += dz * dz;
        const long i = r * L + n;
        const double ph = php[k], dph = dphp[k], ps = psp[k];
        const double t1 = (0.5 - ps) * dph * dph;
        const double t2 = (0.5 + ps) * g / a2;
        const double V = m2 * ph * ph / 2.0;
        const double rv = t1 + t2 + V;
        const double pv = t1 - t2 / 3.0 - V;
        result[N + i] = lap;
        result[3 * N + i] = pv;
        if (grad3) {
            grad3[i] = (G)dx;
            grad3[N + i] = (G)dy;
            grad3[2 * N + i] = (G)dz;
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
        partials[2 * bk.x] = srho;
        partials[2 * bk.x + 1] = sp;
    }
}

template <int L, class G>
__global__ __launch_bounds__(FFT_WG) void k_z_c2r_rho(
    const double2 *__restrict__ W, long M, const double *__restrict__ f,
    double *__restrict__ result, G *__restrict__ grad3,
    double *__restrict__ rho, double *__restrict__ pressure,
    double *__restrict__ partials, long nrows, long N, double m2,
    const double2 *__restrict__ tw, Ens e)
{
    __shared__ double2 lds[z_rho_lds<L>()];
    z_c2r_rho_body<L, G>(hw_blk(), lds, W, M, f, result, grad3, rho, pressure,
                         partials, nrows, N, m2, tw, e);
}

// rho_mean, pressure_mean (src/toolbox.c:241-242) and da (src/toolbox.c:87)
// (one work-group per realization r; sh: 8 doubles of LDS)
// EP: the RK4 epilogue of the scale factor's slope
template <bool EP = false>
__device__ __forceinline__ void means_body(
    unsigned r, double *sh, const double *__restrict__ partials, int nparts,
    long N, const double *__restrict__ f, double *__restrict__ result,
    long Ntot, double *__restrict__ dscal, const Ens &e,
    const Rk4Ep &ep = Rk4Ep{})
{
    partials += r * e.p;
    f += r * e.f;
    result += r * e.f;
    dscal += r * e.s;
    // (rho, p) pairs as 16-byte loads, four in flight per thread
    const double2 *p2 = reinterpret_cast<const double2 *>(partials);
    double s0 = 0.0, s1 = 0.0;
#pragma unroll 4
    for (int i = threadIdx.x; i < nparts; i += RED_WG) {
        const double2 v = p2[i];
        s0 += v.x;
        s1 += v.y;
    }
    s0 = block_reduce<OpSum>(s0, sh, RED_WG);
    s1 = block_reduce<OpSum>(s1, sh, RED_WG);
    if (threadIdx.x == 0) {
        const double rm = s0 / N, pm = s1 / N;
        dscal[S_RHO_MEAN] = rm;
        dscal[S_P_MEAN] = pm;
        const double da = f[Ntot - 1] * sqrt(rm / 3.0);
        if constexpr (EP)
            rk4_ep_one(ep.shifted(r * e.f), Ntot - 1, da, result + Ntot - 1);
        else
            result[Ntot - 1] = da;
    }
}

// update_phi_psi (src/toolbox.c:282-299); lap and p were parked in
// result[N..2N) and result[3N..4N) by k_z_c2r_rho
// (work-group bx of gx on realization z; EP: the RK4 epilogue of the four
// blocks' slopes)
template <bool NT, bool EP = false>
__device__ __forceinline__ void update_phi_psi_body(
    unsigned bx, unsigned gx, unsigned z, const double *__restrict__ f,
    double *__restrict__ result, long N, long Ntot, double m2,
    const double *__restrict__ dscal, const Ens &e, const Rk4Ep &ep0 = Rk4Ep{})
{
    f += z * e.f;
    result += z * e.f;
    dscal += z * e.s;
    [[maybe_unused]] const Rk4Ep ep = EP ? ep0.shifted(z * e.f) : ep0;
    const double aa = f[Ntot - 1];
    const double a2 = aa * aa;
    const double hubble = sqrt(dscal[S_RHO_MEAN] / 3.0);
    const double h3 = 3.0 * hubble;
    const double pm = dscal[S_P_MEAN];
    // pairs of points (16-byte loads and stores; N is even and every block
    // starts 16-byte aligned)
    const long stride = (long)gx * RED_WG;
    const double2 *F = reinterpret_cast<const double2 *>(f);
    double2 *Rs = reinterpret_cast<double2 *>(result);
    const long H = N / 2;
    for (long i = (long)bx * RED_WG + threadIdx.x; i < H; i += stride) {
        const double2 ph = ld2<NT>(F + i), dph = ld2<NT>(F + H + i),
                      ps = ld2<NT>(F + 2 * H + i), dps = ld2<NT>(F + 3 * H + i);
        const double2 lap = ld2<NT>(Rs + H + i), pr = ld2<NT>(Rs + 3 * H + i);
        const double2 r1 = make_double2(
            (1.0 + 4.0 * ps.x) * lap.x / a2 - (h3 - 4.0 * dps.x) * dph.x -
                (1.0 + 2.0 * ps.x) * (m2 * ph.x),
            (1.0 + 4.0 * ps.y) * lap.y / a2 - (h3 - 4.0 * dps.y) * dph.y -
                (1.0 + 2.0 * ps.y) * (m2 * ph.y));
        const double2 r3 = make_double2(
            0.5 * pr.x + (ps.x - 0.5) * pm - 4.0 * hubble * dps.x,
            0.5 * pr.y + (ps.y - 0.5) * pm - 4.0 * hubble * dps.y);
        if constexpr (EP) {
            rk4_ep_pair<NT>(ep, i, dph, Rs + i);
            rk4_ep_pair<NT>(ep, H + i, r1, Rs + H + i);
            rk4_ep_pair<NT>(ep, 2 * H + i, dps, Rs + 2 * H + i);
            rk4_ep_pair<NT>(ep, 3 * H + i, r3, Rs + 3 * H + i);
        } else {
            st2<NT>(Rs + i, dph);
            st2<NT>(Rs + H + i, r1);
            st2<NT>(Rs + 2 * H + i, dps);
            st2<NT>(Rs + 3 * H + i, r3);
        }
    }
}

template <bool NT>
__global__ __launch_bounds__(RED_WG) void k_update_phi_psi(
    const double *__restrict__ f, double *__restrict__ result, long N,
    long Ntot, double m2, const double *__restrict__ dscal, Ens e)
{
    update_phi_psi_body<NT>(blockIdx.x, gridDim.x, blockIdx.z, f, result, N,
                            Ntot, m2, dscal, e);
}

// ---------------------------------------------------------------------------
// GW sources: z r2c of the six products d_i phi d_j phi (mk_gw_sources,
// src/toolbox.c:410-424).  A work-group owns B rows: every thread loads its
// first-stage points of d_x phi, d_y phi, d_z phi once into registers, then
// the three complex rows (S11, S12), (S13, S22), (S23, S33) of each row are
// transformed one after the other; component c = 2q + s goes to W + c M.
// ---------------------------------------------------------------------------
template <int L, int Q, class C, class Rows>
__device__ __forceinline__ void sources_pair(C *lds, int tid, const Rows &g,
                                             bool act, C *__restrict__ W,
                                             long M, long row0, long nrows,
                                             const C *__restrict__ tw)
{
    using TL = Tile<L>;
    using R = typename CT<C>::R;
    constexpr int MZ = L / 2 + 1;
    const int bc = tid / TL::T;
    auto rd = [&](int s, int) {
        R u, v;
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
        return act ? CT<C>::make(u, v) : CT<C>::make(0, 0);
    };
    fft_line_wave<L, -1, typename TL::P>(lds + bc * TL::PITCH, tid % TL::T, tw, rd);
    // each wavefront splits and stores the rows of its own lines
    const int w = tid / 64, lane = tid % 64;
    for (int idx = lane; idx < TL::LPW * 2 * MZ; idx += 64) {
        const int lb = w * TL::LPW + idx / (2 * MZ);
        const int rem = idx - (lb - w * TL::LPW) * 2 * MZ;
        const int s = rem / MZ, k = rem - s * MZ;
        const long r = row0 + lb;
        if (r < nrows)
            W[(2 * Q + s) * M + r * MZ + k] =
                herm_split<L>(lds + lb * TL::PITCH, k, s);
    }
    line_sync<true>(); // split reads done before the next transform
}

template <int L, class R>
struct GradRegs {
    R x[Tile<L>::E], y[Tile<L>::E], z[Tile<L>::E];
};

// C = double2: the FP64 path; C = float2: the precision form (gradients
// stored in FP32, products, FFTs and output in FP32)
template <int L, class C>
__device__ __forceinline__ void z_r2c_sources_body(
    const Blk &bk, C *lds, const typename CT<C>::R *__restrict__ g3,
    C *__restrict__ W, long M, long nrows, long N, const C *__restrict__ tw,
    long rg, long rw)
{
    g3 += bk.z * rg;
    W += bk.z * rw;
    using TL = Tile<L>;
    using R = typename CT<C>::R;
    const int tid = threadIdx.x;
    const int bc = tid / TL::T, j = tid % TL::T;
    const long row0 = (long)bk.x * TL::B;
    const long row = row0 + bc;
    const bool act = row < nrows;
    const R *gx = g3 + (act ? row : 0) * L;
    GradRegs<L, R> g;
#pragma unroll
    for (int s = 0; s < TL::E; ++s) {
        const int n = stage1_point<L, typename TL::P>(j, s);
        g.x[s] = gx[n];
        g.y[s] = gx[N + n];
        g.z[s] = gx[2 * N + n];
    }
    sources_pair<L, 0>(lds, tid, g, act, W, M, row0, nrows, tw);
    sources_pair<L, 1>(lds, tid, g, act, W, M, row0, nrows, tw);
    sources_pair<L, 2>(lds, tid, g, act, W, M, row0, nrows, tw);
}

template <int L, class C>
__global__ __launch_bounds__(FFT_WG) void k_z_r2c_sources(
    const typename CT<C>::R *__restrict__ g3, C *__restrict__ W, long M,
    long nrows, long N, const C *__restrict__ tw, long rg, long rw)
{
    __shared__ C lds[Tile<L>::B * Tile<L>::PITCH];
    z_r2c_sources_body<L, C>(hw_blk(), lds, g3, W, M, nrows, N, tw, rg, rw);
}

// ---------------------------------------------------------------------------
// TT projection (mk_gw_sources_tt, src/toolbox.c:363-397) as two linear forms
// of the six source components: o0 = sum_c alpha_c S_c, o1 = sum_c beta_c S_c.
// With f = k / k^2 (full k), t1 = sum_ij k_i f_j S_ij, t2 = tr S, s1 = t1 + t2,
// s2 = t1 - t2 and the kept axis pair (p, q) = (x, y) if |kz| > eps, (x, z)
// if |ky| > eps, (y, z) otherwise, the reference computes
//   o0 = S_pp - 2 f_p (sum_j k_j S_pj) + (f_p k_p s1 + s2) / 2
//   o1 = S_pq - f_p (sum_j k_j S_qj) - f_q (sum_j k_j S_pj) + f_p k_q s1 / 2
// alpha_C, beta_C are the derivatives of these with respect to component
// C = 0..5 = (xx, xy, xz, yy, yz, zz); the k = 0 mode projects to 0.
// ---------------------------------------------------------------------------
template <int C, class R>
__device__ __forceinline__ void tt_coef(R kx, R ky, R kz, R inv_ksq,
                                        bool b1, bool zero, R &al, R &be)
{
    const R fx = kx * inv_ksq, fy = ky * inv_ksq, fz = kz * inv_ksq;
    const bool b2 = !b1 && fabs((double)ky) > DBL_EPSILON;
    const bool py = !b1 && !b2; // p = y, else p = x
    const bool qy = b1;         // q = y, else q = z
    const R kp = py ? ky : kx, fp = py ? fy : fx;
    const R kq = qy ? ky : kz, fq = qy ? fy : fz;
    R t1, t2, rowp, rowq, dpp, dpq;
    const R O = 0, I = 1, D = 2;
    if constexpr (C == 0) {
        t1 = kx * fx; t2 = I; rowp = py ? O : kx; rowq = O;
        dpp = py ? O : I; dpq = O;
    } else if constexpr (C == 1) {
        t1 = D * (kx * fy); t2 = O; rowp = py ? kx : ky;
        rowq = qy ? kx : O; dpp = O; dpq = (!py && qy) ? I : O;
    } else if constexpr (C == 2) {
        t1 = D * (kx * fz); t2 = O; rowp = py ? O : kz;
        rowq = qy ? O : kx; dpp = O; dpq = (!py && !qy) ? I : O;
    } else if constexpr (C == 3) {
        t1 = ky * fy; t2 = I; rowp = py ? ky : O; rowq = qy ? ky : O;
        dpp = py ? I : O; dpq = O;
    } else if constexpr (C == 4) {
        t1 = D * (ky * fz); t2 = O; rowp = py ? kz : O;
        rowq = qy ? kz : ky; dpp = O; dpq = py ? I : O;
    } else {
        t1 = kz * fz; t2 = I; rowp = O; rowq = qy ? O : kz;
        dpp = O; dpq = O;
    }
    const R s1 = t1 + t2, s2 = t1 - t2;
    const R half = (R)0.5, two = (R)2.0;
    al = zero ? (R)0 : dpp - two * fp * rowp + half * (fp * kp * s1 + s2);
    be = zero ? (R)0 : dpq - fp * rowq - fq * rowp + half * fp * kq * s1;
}

// y transform of source component K (its z and x passes are done) with the
// TT accumulation folded into the last Stockham stage; `cur` holds the
// component's first-stage points (loaded ahead).  Two layouts: the
// work-group's threads, lanes over the B lines as the loads are (thread j
// of line `lb`), run the first stage from `cur` into the component's LDS
// buffer, then one work-group barrier; then each wavefront runs the other
// stages on its own lines (thread jw of line `lw`) with wavefront-local
// exchanges.  Components alternate between two LDS buffers: the barrier of
// component K also ends every wavefront's use of the buffer component K + 1
// writes.  kx, kz, b1 (|k_z| > eps) and z0 (ix = iz = 0) are those of line
// lw, inv_ksq the 1/k^2 of its points.
template <int L, class P, int K, class C>
__device__ __forceinline__ void tt_component(
    C *lb, C *lw, int j, int jw, const C *__restrict__ tw, const C *cur,
    const KGeom &kg, typename CT<C>::R kx, typename CT<C>::R kz, bool b1,
    bool z0, const typename CT<C>::R *inv_ksq, C *o0, C *o1)
{
    using R = typename CT<C>::R;
    // Register pressure control (empty asm statements, no instructions):
    // an opaque zero offset of the twiddle pointer per component (the
    // pointer itself stays visibly an LDS one) stops the compiler from
    // loading every stage twiddle once and holding it across all six
    // transforms, and pinning the accumulators after each component stops it
    // from deferring the sums (and so keeping all six spectra live).
    int zo = 0;
    asm volatile("" : "+s"(zo));
    const C *twc = tw + zo;
    auto rd = [&](int s, int) { return cur[s]; };
    NoWrite nw;
    stockham_stage<L, P::R0, 1, -1, true, false, P, C, false>(lb, lb, j, twc,
                                                              rd, nw);
    auto acc = [&](int s, int n, C v) {
        // recomputed per component from 1/k^2 alone: an empty asm keeps the
        // compiler from holding every derived coefficient of every point
        // live across the six transforms
        R iv = inv_ksq[s];
        asm volatile("" : "+v"(iv));
        const bool zero = z0 && (n == 0);
        R al, be;
        tt_coef<K>(kx, (R)kg.kyf(n), kz, iv, b1, zero, al, be);
        if constexpr (K == 0) {
            o0[s] = CT<C>::make(al * v.x, al * v.y);
            o1[s] = CT<C>::make(be * v.x, be * v.y);
        } else {
            o0[s] = CT<C>::make(o0[s].x + al * v.x, o0[s].y + al * v.y);
            o1[s] = CT<C>::make(o1[s].x + be * v.x, o1[s].y + be * v.y);
        }
    };
    fft_stages_wave<L, -1, P, 1, C>(lw, jw, twc, acc);
#pragma unroll
    for (int s = 0; s < P::E; ++s)
        asm volatile("" : "+v"(o0[s].x), "+v"(o0[s].y), "+v"(o1[s].x),
                     "+v"(o1[s].y));
}

// Line plans of the fused y pass: few values per thread (radix-4 stages),
// so that the two TT accumulators, 1/k^2 and the transform of every point a
// thread owns stay in registers.
template <int L> struct TTPlan;
template <> struct TTPlan<4>   { using P = PlanT<2, 2, 2>; static constexpr int WG = 256; };
template <> struct TTPlan<8>   { using P = PlanT<4, 4, 2>; static constexpr int WG = 256; };
template <> struct TTPlan<16>  { using P = PlanT<4, 4, 4>; static constexpr int WG = 256; };
template <> struct TTPlan<32>  { using P = PlanT<4, 4, 4, 2>; static constexpr int WG = 256; };
template <> struct TTPlan<64>  { using P = PlanT<4, 4, 4, 4>; static constexpr int WG = 256; };
template <> struct TTPlan<128> { using P = PlanT<4, 4, 4, 4, 2>; static constexpr int WG = 512; };
template <> struct TTPlan<256> { using P = PlanT<4, 4, 4, 4, 4>; static constexpr int WG = 512; };
template <> struct TTPlan<512> { using P = PlanT<8, 8, 4, 4, 4>; static constexpr int WG = 512; };

// LDS of the fused pass in complex values: two lines per tile line and the
// L stage twiddles
template <int L, int WG, class P>
__host__ __device__ constexpr int tt_lds()
{
    return 2 * (WG / (L / P::E)) * Lds<L>::PITCH + L;
}

// Last pass of the GW sources: y transforms of the six components (z and x
// passes done), the TT projection, and update_h (src/toolbox.c:311-346):
//   dh = hdot,  dhdot = -(2 p_mean + k^2/a^2) h - 3 H hdot + 2 S^TT / a^2
// (EP: the RK4 epilogue of the four h blocks' slopes)
template <int L, int WG, class P, class C, bool EP = false>
__device__ __forceinline__ void y_tt_update_h_body(
    const Blk &bk, C *lds, const C *W, long M, const double *__restrict__ f,
    double *__restrict__ result, long N, long Next, long Ntot,
    const KGeom &kg, const double *__restrict__ dscal, const LineSet &ls,
    const C *__restrict__ tw, const Ens &e, long rw, const Rk4Ep &ep0 = Rk4Ep{})
{
    // realization z of an ensemble (rw: its W stride in units of C)
    W += bk.z * rw;
    f += bk.z * e.f;
    result += bk.z * e.f;
    dscal += bk.z * e.s;
    using R = typename CT<C>::R;
    constexpr int E = P::E, T = L / E, B = WG / T, PITCH = Lds<L>::PITCH;
    const int tid = threadIdx.x, b = tid % B, j = tid / B;
    const long line = (long)tile_x<true>(bk) * B + b;
    const bool act = line < ls.nlines;
    const long lc = act ? line : 0;
    const long base = ls.base(lc);
    const int ix = (int)(lc / kg.Mz), iz = (int)(lc - (long)ix * kg.Mz);
    const long st = ls.stride;
    // The stage twiddles tw[L, 2L) are staged in LDS behind the lines and
    // read from there: LDS reads are counted apart from the vector-memory
    // loads, which return in order, so a twiddle read never waits for the
    // source and h loads issued ahead of it (from global memory it would
    // wait for all of them).  The first stage, whose barrier makes them
    // visible, takes no twiddle.
    {
        C *const ltw = lds + 2 * B * PITCH;
        for (int i = tid; i < L; i += WG)
            ltw[i] = tw[L + i];
        tw = ltw - L; // indices L + [0, L) of the table
    }
    // The wavefront layout of the stages after the first: wavefront w owns
    // lines w LPW + [0, LPW) of the tile, thread jw of line bw.  (The
    // wavefront index is uniform: with one line per wavefront its
    // coordinates and k_x, k_z live in scalar registers.)
    constexpr int LPW = 64 / T;
    const int wv = __builtin_amdgcn_readfirstlane(tid / 64);
    const int bw = wv * LPW + (tid % 64) / T, jw = tid % T;
    const long linew = (long)tile_x<true>(bk) * B + bw;
    const long lcw = linew < ls.nlines ? linew : 0;
    const int ixw = (int)(lcw / kg.Mz), izw = (int)(lcw - (long)ixw * kg.Mz);
    const R kxw = (R)kg.kxf(ixw), kzw = (R)kg.kzf(izw);
    const bool b1w = fabs(kg.kzf(izw)) > DBL_EPSILON; // branch on the FP64 k
    const bool z0w = (ixw == 0) && (izw == 0);
    R inv[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int n = last_point<L, P>(jw, s);
        inv[s] = (z0w && n == 0) ? (R)0 : (R)(1.0 / kg.ksq(ixw, n, izw));
    }
    // the two LDS buffers, at this thread's line of each layout
    C *const lb0 = lds + b * PITCH, *const lb1 = lds + (B + b) * PITCH;
    C *const lw0 = lds + bw * PITCH, *const lw1 = lds + (B + bw) * PITCH;
    // first-stage points of component c, in bounds for every lane (an idle
    // lane loads line 0's points and its results are never stored); the
    // loads of component c + 1 are issued before the transform of c, so
    // their latency hides behind it.  The loaded values are used as they
    // are: a select on them right after the load would wait for it there.
    // (Loading two components ahead measured no faster.)
    auto load = [&](int c, C *buf) {
        const C *src = W + c * M + base;
#pragma unroll
        for (int s = 0; s < E; ++s)
            buf[s] = src[(long)stage1_point<L, P>(j, s) * st];
    };
    // update_h's inputs (h and hdot at the points this thread ends with;
    // line 0's points for an idle lane, in bounds): with few points per
    // thread they are loaded before the last transform, so their latency
    // hides behind it (with E = 8 the registers are not there)
    const long Nh1 = 4 * N, Nh2 = Nh1 + Next, Ndh1 = Nh2 + Next,
               Ndh2 = Ndh1 + Next;
    const double2 *f2 = reinterpret_cast<const double2 *>(f);
    constexpr bool EARLY = E <= 4;
    double2 h1v[E], h2v[E], d1v[E], d2v[E];
    auto load_h = [&]() {
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const long i = base + last_point<L, P>(j, s) * st;
            h1v[s] = f2[Nh1 / 2 + i];
            h2v[s] = f2[Nh2 / 2 + i];
            d1v[s] = f2[Ndh1 / 2 + i];
            d2v[s] = f2[Ndh2 / 2 + i];
        }
    };
    C o0[E], o1[E], bufs[2][E];
    load(0, bufs[0]);
    load(1, bufs[1]);
    auto step = [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        if constexpr (K == 5 && EARLY) load_h();
        tt_component<L, P, K>(K % 2 ? lb1 : lb0, K % 2 ? lw1 : lw0, j, jw, tw,
                              bufs[K % 2], kg, kxw, kzw, b1w, z0w, inv, o0,
                              o1);
        if constexpr (K + 2 < 6) load(K + 2, bufs[K % 2]);
    };
    step(std::integral_constant<int, 0>{});
    step(std::integral_constant<int, 1>{});
    step(std::integral_constant<int, 2>{});
    step(std::integral_constant<int, 3>{});
    step(std::integral_constant<int, 4>{});
    step(std::integral_constant<int, 5>{});
    // the two sums back to the load layout through LDS: o0 to buffer 0 (no
    // wavefront reads it after component 5's barrier), o1 to this
    // wavefront's own line of buffer 1 once its last reads of it are done
#pragma unroll
    for (int s = 0; s < E; ++s)
        lw0[Lds<L>::idx(last_point<L, P>(jw, s))] = o0[s];
    line_sync<true>();
#pragma unroll
    for (int s = 0; s < E; ++s)
        lw1[Lds<L>::idx(last_point<L, P>(jw, s))] = o1[s];
    lds_barrier();
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int n = Lds<L>::idx(last_point<L, P>(j, s));
        o0[s] = lb0[n];
        o1[s] = lb1[n];
    }
    if constexpr (!EARLY) load_h();
    if (!act)
        return;
    const double aa = f[Ntot - 1];
    const double a2 = aa * aa;
    const double hubble = sqrt(dscal[S_RHO_MEAN] / 3.0);
    const double h3 = 3.0 * hubble;
    const double pm = dscal[S_P_MEAN];
    double2 *r2 = reinterpret_cast<double2 *>(result);
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int n = last_point<L, P>(j, s);
        const long i = base + n * st;
        const double ksq = kg.ksq(ix, n, iz);
        const double2 h1 = h1v[s], h2 = h2v[s], d1 = d1v[s], d2 = d2v[s];
        if constexpr (!EP) {
            r2[Nh1 / 2 + i] = d1;
            r2[Nh2 / 2 + i] = d2;
        }
        const double tmp = -(2.0 * pm + ksq / a2);
        const double s0x = o0[s].x, s0y = o0[s].y, s1x = o1[s].x, s1y = o1[s].y;
        const double2 v1 = make_double2(tmp * h1.x - h3 * d1.x + 2.0 * s0x / a2,
                                        tmp * h1.y - h3 * d1.y + 2.0 * s0y / a2);
        const double2 v2 = make_double2(tmp * h2.x - h3 * d2.x + 2.0 * s1x / a2,
                                        tmp * h2.y - h3 * d2.y + 2.0 * s1y / a2);
        if constexpr (EP) {
            const Rk4Ep ep = ep0.shifted(bk.z * e.f);
            rk4_ep_pair<false>(ep, Nh1 / 2 + i, d1, r2 + Nh1 / 2 + i);
            rk4_ep_pair<false>(ep, Nh2 / 2 + i, d2, r2 + Nh2 / 2 + i);
            rk4_ep_pair<false>(ep, Ndh1 / 2 + i, v1, r2 + Ndh1 / 2 + i);
            rk4_ep_pair<false>(ep, Ndh2 / 2 + i, v2, r2 + Ndh2 / 2 + i);
        } else {
            r2[Ndh1 / 2 + i] = v1;
            r2[Ndh2 / 2 + i] = v2;
        }
    }
}

template <int L, int WG, class P, class C>
__global__ __launch_bounds__(WG) void k_y_tt_update_h(
    const C *W, long M, const double *__restrict__ f,
    double *__restrict__ result, long N, long Next, long Ntot, KGeom kg,
    const double *__restrict__ dscal, LineSet ls, const C *__restrict__ tw,
    Ens e, long rw)
{
    __shared__ C lds[tt_lds<L, WG, P>()];
    y_tt_update_h_body<L, WG, P, C>(hw_blk(), lds, W, M, f, result, N, Next,
                                    Ntot, kg, dscal, ls, tw, e, rw);
}


} // namespace ihg
