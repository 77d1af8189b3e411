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
// row), an x pass, and one y pass that also does the TT projection and the h
// update (toolbox.c:311-424).
#include <cfloat>
#include <cmath>
#include <type_traits>

#include "dispatch.h"
#include "fft_kernels.h"
#include "ihg_internal.h"
#include "mem.h"
#include "reduce.h"
#include "rhs_kernels.h"

namespace ihg {

static inline unsigned nblocks(long n, int per_block, unsigned cap)
{
    long b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    return (unsigned)(b > (long)cap ? cap : b);
}

// rho_mean, pressure_mean and da of every realization (one work-group each)
__global__ __launch_bounds__(RED_WG) void k_means(
    const double *__restrict__ partials, int nparts, long N,
    const double *__restrict__ f, double *__restrict__ result, long Ntot,
    double *__restrict__ dscal, Ens e)
{
    __shared__ double sh[8];
    means_body(blockIdx.x, sh, partials, nparts, N, f, result, Ntot, dscal, e);
}

// the same three kernels with the RK4 epilogue of their slopes
__global__ __launch_bounds__(RED_WG) void k_means_rk4(
    const double *__restrict__ partials, int nparts, long N,
    const double *__restrict__ f, double *__restrict__ result, long Ntot,
    double *__restrict__ dscal, Ens e, Rk4Ep ep)
{
    __shared__ double sh[8];
    means_body<true>(blockIdx.x, sh, partials, nparts, N, f, result, Ntot,
                     dscal, e, ep);
}

template <bool NT>
__global__ __launch_bounds__(RED_WG) void k_update_phi_psi_rk4(
    const double *__restrict__ f, double *__restrict__ result, long N,
    long Ntot, double m2, const double *__restrict__ dscal, Ens e, Rk4Ep ep)
{
    update_phi_psi_body<NT, true>(blockIdx.x, gridDim.x, blockIdx.z, f, result,
                                  N, Ntot, m2, dscal, e, ep);
}

template <int L, int WG, class P>
__global__ __launch_bounds__(WG) void k_y_tt_update_h_rk4(
    const double2 *W, long M, const double *__restrict__ f,
    double *__restrict__ result, long N, long Next, long Ntot, KGeom kg,
    const double *__restrict__ dscal, LineSet ls,
    const double2 *__restrict__ tw, Ens e, long rw, Rk4Ep ep)
{
    __shared__ double2 lds[tt_lds<L, WG, P>()];
    y_tt_update_h_body<L, WG, P, double2, true>(hw_blk(), lds, W, M, f, result,
                                                N, Next, Ntot, kg, dscal, ls,
                                                tw, e, rw, ep);
}

// the y line lengths whose TT pass has the RK4 epilogue compiled in
template <class F>
static int dispatch_rk4_len(int L, F &&f)
{
    switch (L) {
    case 16: return f(std::integral_constant<int, 16>{});
    case 32: return f(std::integral_constant<int, 32>{});
    case 64: return f(std::integral_constant<int, 64>{});
    case 128: return f(std::integral_constant<int, 128>{});
    default: return set_error("no RK4 epilogue for %d-point lines", L);
    }
}

bool rk4_fusable(const ihg_ctx *c)
{
    if (!c->p.enable_gw) return true;
    const int L = c->Ny;
    return !c->p.gw_fp32 && (L == 16 || L == 32 || L == 64 || L == 128);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static LineSet x_lineset(const ihg_ctx *c)
{
    const long plane = (long)c->Ny * c->Mz;
    return LineSet{plane, plane, 0, plane};
}

int rhs_core(ihg_ctx *c, double *f, double *result, int output, int kept,
             int nr, const Rk4Ep *ep)
{
    const long N = c->N, M = c->M;
    const double Nd = (double)N;
    const double m2 = c->p.mass * c->p.mass;
    double2 *W = c->W;
    const Ens e = nr > 1 ? c->ens : Ens{};
    // phi_hat into W4: kept by the filter pass (full spectrum), or here by
    // the z and y passes (the x pass is the fan-out's first half)
    if (!kept) {
        IHG_TRY(z_r2c_pass(c, f, W + 4 * M, 1, 0, 0, nr, e.f, e.w));
        IHG_TRY(y_pass(c, W + 4 * M, -1, 1, 0, nr, e.w));
    }
    // x fan-out into the four derivative spectra W0..W3
    IHG_TRY(dispatch_len(c->Nx, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        using FP = FanPlan<L>;
        using P = typename FP::P;
        constexpr int B = FP::WG / (L / P::E);
        const LineSet ls = x_lineset(c);
        DerivFac fac{c->kg, Nd};
        const dim3 grid((unsigned)((ls.nlines + B - 1) / B), 1, nr);
        return timed(c, F_X_FANOUT, 5.0 * 16.0 * M * nr, [&]() {
            if (kept)
                hipLaunchKernelGGL((k_x_fanout<L, 4, false, DerivFac, P, FP::WG>),
                                   grid, dim3(FP::WG), 0, c->stream, W + 4 * M,
                                   W, M, ls, 0L, c->tw, fac, e.w);
            else
                hipLaunchKernelGGL((k_x_fanout<L, 4, true, DerivFac, P, FP::WG>),
                                   grid, dim3(FP::WG), 0, c->stream, W + 4 * M,
                                   W, M, ls, 0L, c->tw, fac, e.w);
            return hip_check(hipGetLastError(), "k_x_fanout");
        });
    }));
    IHG_TRY(y_pass(c, W, +1, 4, M, nr, e.w));
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
        const double gbytes = (gw && c->p.gw_fp32) ? 4.0 : 8.0;
        const double bytes = (4.0 * 16.0 * M + 8.0 * N * (5 + (output ? 2 : 0)) +
                              (gw ? 3.0 * gbytes * N : 0.0)) * nr;
        const dim3 grid(nb, 1, nr);
        return timed(c, F_Z_C2R_RHO, bytes, [&]() {
            if (gw && c->p.gw_fp32)
                hipLaunchKernelGGL((k_z_c2r_rho<L, float>), grid,
                                   dim3(FFT_WG), 0, c->stream, W, M, f, result,
                                   reinterpret_cast<float *>(c->grad3),
                                   output ? c->rho : nullptr,
                                   output ? c->pressure : nullptr, c->partials,
                                   nrows, N, m2, c->tw, e);
            else
                hipLaunchKernelGGL((k_z_c2r_rho<L, double>), grid,
                                   dim3(FFT_WG), 0, c->stream, W, M, f, result,
                                   gw ? c->grad3 : nullptr,
                                   output ? c->rho : nullptr,
                                   output ? c->pressure : nullptr, c->partials,
                                   nrows, N, m2, c->tw, e);
            return hip_check(hipGetLastError(), "k_z_c2r_rho");
        });
    }));
    IHG_TRY(timed(c, F_MEANS, 16.0 * nparts * nr, [&]() {
        if (ep)
            hipLaunchKernelGGL(k_means_rk4, dim3(nr), dim3(RED_WG), 0,
                               c->stream, c->partials, nparts, N, f, result,
                               c->Ntot, c->dscal, e, *ep);
        else
            hipLaunchKernelGGL(k_means, dim3(nr), dim3(RED_WG), 0, c->stream,
                               c->partials, nparts, N, f, result, c->Ntot,
                               c->dscal, e);
        return hip_check(hipGetLastError(), "k_means");
    }));
    // the epilogue reads y (and k1..k3) and writes the stage output beside
    // the slope: 2 (stage) or 4 (final, no slope written) more blocks
    const double ep_blocks = !ep ? 0.0 : (ep->mode == RK4_FINAL ? 4.0 : 8.0);
    return timed(c, F_UPDATE, (10.0 + ep_blocks) * 8.0 * N * nr, [&]() {
        // the streaming hint when the four blocks of all realizations
        // outgrow the caches
        const bool nt = 4.0 * 8.0 * (double)N * nr >= NT_BYTES;
        const dim3 grid(nblocks(N, RED_WG * 4, 65536), 1, nr);
        if (ep)
            hipLaunchKernelGGL((nt ? k_update_phi_psi_rk4<true>
                                   : k_update_phi_psi_rk4<false>),
                               grid, dim3(RED_WG), 0, c->stream, f, result, N,
                               c->Ntot, m2, c->dscal, e, *ep);
        else
            hipLaunchKernelGGL((nt ? k_update_phi_psi<true>
                                   : k_update_phi_psi<false>),
                               grid, dim3(RED_WG), 0, c->stream, f, result, N,
                               c->Ntot, m2, c->dscal, e);
        return hip_check(hipGetLastError(), "k_update_phi_psi");
    });
}

// GW source pipeline on complex type C: z r2c of the six products, x pass,
// fused y pass + TT + update_h.  Wk holds 6 M values of C, g3 the gradients;
// realization r of an ensemble finds them r * rw and r * rg further on.
template <class C>
static int rhs_gw_t(ihg_ctx *c, const typename CT<C>::R *g3, C *Wk,
                    const C *tw, const double *f, double *result, int nr,
                    long rg, long rw, const Rk4Ep *ep)
{
    const long N = c->N, M = c->M;
    const long nrows = (long)c->Nx * c->Ny;
    const double cs = sizeof(C), rs = sizeof(typename CT<C>::R);
    const Ens e = nr > 1 ? c->ens : Ens{};
    IHG_TRY(dispatch_len(c->Nz, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        using TL = Tile<L>;
        const unsigned nb = (unsigned)((nrows + TL::B - 1) / TL::B);
        return timed(c, F_Z_R2C_SRC, (3.0 * rs * N + 6.0 * cs * M) * nr, [&]() {
            hipLaunchKernelGGL((k_z_r2c_sources<L, C>), dim3(nb, 1, nr),
                               dim3(FFT_WG), 0, c->stream, g3, Wk, M, nrows, N,
                               tw, rg, rw);
            return hip_check(hipGetLastError(), "k_z_r2c_sources");
        });
    }));
    if constexpr (std::is_same<C, float2>::value)
        IHG_TRY(x_pass_f32(c, Wk, -1, 6, M, nr, rw));
    else
        IHG_TRY(x_pass(c, Wk, -1, 6, M, nr, rw));
    // y pass of the six sources, TT projection and update_h: six sources
    // and h1, h2, dh1, dh2 in; the four h blocks of result out
    if constexpr (std::is_same<C, double2>::value) {
        if (ep)
            return dispatch_rk4_len(c->Ny, [&](auto Lc) {
                constexpr int L = decltype(Lc)::value;
                using TP = TTPlan<L>;
                constexpr int B = TP::WG / (L / TP::P::E);
                const LineSet ls{(long)c->Nx * c->Mz, c->Mz, (long)c->Ny * c->Mz, c->Mz};
                const unsigned nb = (unsigned)((ls.nlines + B - 1) / B);
                const double epb = ep->mode == RK4_FINAL ? 8.0 : 16.0;
                return timed(c, F_TT_UPDATE_H, (6.0 * 16.0 + (8.0 + epb) * 16.0) * M * nr, [&]() {
                    hipLaunchKernelGGL((k_y_tt_update_h_rk4<L, TP::WG, typename TP::P>),
                                       dim3(nb, 1, nr), dim3(TP::WG), 0, c->stream, Wk,
                                       M, f, result, N, c->Next, c->Ntot, c->kg,
                                       c->dscal, ls, tw, e, rw, *ep);
                    return hip_check(hipGetLastError(), "k_y_tt_update_h_rk4");
                });
            });
    } else if (ep) {
        return set_error("no RK4 epilogue for the FP32 GW source");
    }
    return dispatch_len(c->Ny, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        using TP = TTPlan<L>;
        constexpr int B = TP::WG / (L / TP::P::E);
        // lines along y: one per (x, kz); points Mz apart
        const LineSet ls{(long)c->Nx * c->Mz, c->Mz, (long)c->Ny * c->Mz, c->Mz};
        const unsigned nb = (unsigned)((ls.nlines + B - 1) / B);
        return timed(c, F_TT_UPDATE_H, (6.0 * cs + 8.0 * 16.0) * M * nr, [&]() {
            hipLaunchKernelGGL((k_y_tt_update_h<L, TP::WG, typename TP::P, C>),
                               dim3(nb, 1, nr), dim3(TP::WG), 0, c->stream, Wk,
                               M, f, result, N, c->Next, c->Ntot, c->kg,
                               c->dscal, ls, tw, e, rw);
            return hip_check(hipGetLastError(), "k_y_tt_update_h");
        });
    });
}

int rhs_gw(ihg_ctx *c, const double *f, double *result, int nr,
           const Rk4Ep *ep)
{
    const Ens e = nr > 1 ? c->ens : Ens{};
    if (c->p.gw_fp32)
        // FP32 values in the FP64 buffers: 6 M float2 fit in W, 3 N floats
        // in grad3 (the same element strides per realization)
        return rhs_gw_t<float2>(c, reinterpret_cast<const float *>(c->grad3),
                                reinterpret_cast<float2 *>(c->W), c->twf, f,
                                result, nr, e.g, e.w, ep);
    return rhs_gw_t<double2>(c, c->grad3, c->W, c->tw, f, result, nr, e.g,
                             e.w, ep);
}

// mk_ffts_and_filter's filter branch (src/toolbox.c:164-171): phi, dphi,
// psi, dpsi (contiguous in f) filtered in place, all four in one batch.
// The filtered spectra of phi (-> W4 for the derivatives and Wc for the
// output) and psi (-> Wc + M) are kept, as the reference keeps them.
int rhs_filter(ihg_ctx *c, double *f, int nr)
{
    const long N = c->N, M = c->M;
    double2 *W = c->W;
    const Ens e = nr > 1 ? c->ens : Ens{};
    IHG_TRY(z_r2c_pass(c, f, W, 4, N, M, nr, e.f, e.w));
    IHG_TRY(y_pass(c, W, -1, 4, M, nr, e.w));
    IHG_TRY(dispatch_len(c->Nx, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        using SS = StridedShape<L>;
        const LineSet ls = x_lineset(c);
        // one realization: keep phi_hat in W4 and Wc, psi_hat in Wc + M;
        // an ensemble (no output calls) keeps phi_hat in each W4
        FilterFac fac{c->filter, (long)c->Ny * c->Mz, c->Nx, c->Ny, c->Mz,
                      (double)N, W + 4 * M, nr > 1 ? nullptr : c->Wc,
                      nr > 1 ? nullptr : c->Wc + M, e.w};
        // four fields in and out, three kept spectra out
        return timed(c, F_X_ROUNDTRIP, 11.0 * 16.0 * M * nr, [&]() {
            hipLaunchKernelGGL((k_x_fanout<L, 1, true, FilterFac>),
                               dim3((unsigned)((ls.nlines + SS::B - 1) / SS::B), 4, nr),
                               dim3(FFT_WG), 0, c->stream, W, W, 0L, ls, M,
                               c->tw, fac, e.w);
            return hip_check(hipGetLastError(), "k_x_fanout (filter)");
        });
    }));
    IHG_TRY(y_pass(c, W, +1, 4, M, nr, e.w));
    return z_c2r_pass(c, W, f, 4, M, N, nr, e.w, e.f);
}

} // namespace ihg
