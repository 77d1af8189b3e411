// ihg_internal.h -- library-private declarations shared by the .hip units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "../../include/ihg.h"

namespace ihg {

// k-space geometry, passed by value to kernels.  The wave numbers are
// derived from the grid indices exactly as mk_k_grid does
// (src/setup.c:609-666), so no k tables are read from HBM.
struct KGeom {
    int Nx, Ny, Nz, Mz;
    double kx1, ky1, kz1; // 2 pi / L_i
    double kx2, ky2, kz2; // (2 pi / L_i)^2
    double k_max;
    // log-spaced spectrum bins: fundamental |k| and log(k_max / k_lo)
    int log_bins;
    double k_lo, log_span;

    // derivative wave numbers: Nyquist entry zeroed (kvec.x/y/z)
    __device__ __forceinline__ double kxd(int i) const
    {
        return (i > Nx / 2) ? kx1 * (i - Nx) : ((2 * i == Nx) ? 0.0 : kx1 * i);
    }
    __device__ __forceinline__ double kyd(int j) const
    {
        return (j > Ny / 2) ? ky1 * (j - Ny) : ((2 * j == Ny) ? 0.0 : ky1 * j);
    }
    __device__ __forceinline__ double kzd(int k) const
    {
        return (2 * k == Nz) ? 0.0 : kz1 * k;
    }
    // full wave numbers (kvec.xf/yf/zf)
    __device__ __forceinline__ double kxf(int i) const
    {
        return (i > Nx / 2) ? kx1 * (i - Nx) : kx1 * i;
    }
    __device__ __forceinline__ double kyf(int j) const
    {
        return (j > Ny / 2) ? ky1 * (j - Ny) : ky1 * j;
    }
    __device__ __forceinline__ double kzf(int k) const { return kz1 * k; }
    // |k|^2 (kvec.sq), same operation order as the reference, no contraction
    __device__ __forceinline__ double ksq(int i, int j, int k) const
    {
#pragma clang fp contract(off)
        double k2 = kz2 * (double)k * (double)k;
        const double ii = (i > Nx / 2) ? (double)(Nx - i) : (double)i;
        const double jj = (j > Ny / 2) ? (double)(Ny - j) : (double)j;
        k2 += kx2 * ii * ii;
        k2 += ky2 * jj * jj;
        return k2;
    }
};

} // namespace ihg

namespace ihg {
// kernel families timed by the live profiler (ihg_prof_*)
enum Fam {
    F_Z_R2C, F_Z_C2R, F_Y_PASS, F_X_PASS, F_X_FANOUT, F_Z_C2R_RHO, F_MEANS,
    F_UPDATE, F_Z_R2C_SRC, F_TT_UPDATE_H, F_X_ROUNDTRIP, F_OBS, F_DP_STAGE,
    F_DP_FINISH, F_RK4, F_STREAM_COPY, F_DP_DENSE, F_GSL_STAGE, F_GSL_FINISH,
    F_DP_DENSE_STAGE, F_NFAM
};
struct ProfRec {
    int fam;
    hipEvent_t a, b;
    double bytes;
};
} // namespace ihg

namespace ihg {
// Realization strides of an ensemble context: realization r of a batched
// launch (blockIdx.z = r) finds its arrays r strides apart.  All zero for a
// single realization.
struct Ens {
    long f = 0; // doubles between realizations' states (f, result)
    long w = 0; // complex values between realizations' W arrays
    long g = 0; // elements between realizations' gradients (grad3)
    long p = 0; // doubles between realizations' reduction partials
    long s = 0; // doubles between realizations' device scalars (dscal)
};
} // namespace ihg

struct ihg_ctx {
    ihg_params p;
    int Nx, Ny, Nz, Mz;
    long N, M, Next, Ntot;
    ihg::KGeom kg;
    hipStream_t stream;
    double2 *tw;       // twiddle tables, tw[L + m] = e^{-2 pi i m / L}
    float2 *twf;       // the same rounded to FP32 (FP32 GW source path)
    double *filter;    // filter window per axis: Fx[Nx] | Fy[Ny] | Fz[Mz]
                       // (the mask `filter` of the reference is their min)
    double2 *W;        // 6 spectral work arrays (M complex each)
    double2 *Wc;       // 2 more spectral arrays (phic / psic for output)
    double *grad3;     // d_x phi, d_y phi, d_z phi (N each) for GW sources
    double *rho;       // N
    double *pressure;  // N
    double *tmpf;      // N, Laplacian of psi on output steps (tmp.f)
    double *partials;  // per-block reduction partials
    long npartials;
    double *fin_partials = nullptr; // k_dp_finish's block partials and the
    long fin_npartials = 0;         // second level of their sum (on demand)
    double *dscal;     // device scalars: [0] rho_mean, [1] pressure_mean, ...
    double *hscal;     // pinned host mirror of dscal
    double *dobs;      // device observables (packed)
    double *hobs;      // pinned host copy of the observables
    long obs_len;
    double *field;     // array mk_summary reads (the reference's `field`)
    int nreal;         // realizations of an ensemble context (1 otherwise)
    int cand_y = 0, cand_x = 0; // strided-pass candidates (autotuner)
    std::vector<void *> user_arrays; // from ihg_malloc, freed by ihg_destroy
    double *gather = nullptr; // strided-output staging (ihg_download_strided)
    size_t gather_len = 0;
    ihg::Ens ens;      // their strides
    unsigned long calls_rhs;
    int have_obs;
    // live profiler: events around every launch of the timed families
    bool prof_on = false;
    std::vector<ihg::ProfRec> prof_pending;
    std::vector<hipEvent_t> ev_pool;
    double fam_ms[ihg::F_NFAM] = {};
    double fam_bytes[ihg::F_NFAM] = {};
    long fam_count[ihg::F_NFAM] = {};
};

namespace ihg {

// FFT launchers (ihg_fft.hip)
int fft_r2c(ihg_ctx *c, const double *in, double2 *out, int batch,
            long rbatch, long cbatch);
int fft_c2r(ihg_ctx *c, const double2 *in, double *out, double2 *work,
            int batch, long cbatch, long rbatch);
int y_pass(ihg_ctx *c, double2 *data, int dir, int batch, long cbatch,
           int nr = 1, long rr = 0);
int x_pass(ihg_ctx *c, double2 *data, int dir, int batch, long cbatch,
           int nr = 1, long rr = 0);
int x_pass_f32(ihg_ctx *c, float2 *data, int dir, int batch, long cbatch,
               int nr = 1, long rr = 0);
int z_r2c_pass(ihg_ctx *c, const double *in, double2 *out, int batch,
               long rbatch, long cbatch, int nr = 1, long rin = 0,
               long rout = 0);
int z_c2r_pass(ihg_ctx *c, const double2 *in, double *out, int batch,
               long cbatch, long rbatch, int nr = 1, long rin = 0,
               long rout = 0);

// right-hand side launchers (ihg_rhs.hip)
// scalar slots in ctx->dscal (16 doubles)
enum {
    S_RHO_MEAN = 0,
    S_P_MEAN = 1,
    S_PHI_MEAN = 2,  // mean of f[0..N) (constraints, psi init)
    S_DPHI_MEAN = 3, // mean of f[N..2N)
    S_DDPHI_MEAN = 4,
    S_FEATURES = 8, // a, variance of phi, mean of psi^2 (step proposer)
    S_STAT = 12,    // mean and variance of ihg_mean / ihg_variance
    S_NSCAL = 16
};

struct Rk4Ep;

// nr = 1: one realization (f, result, the work arrays of realization 0);
// nr = c->nreal: every realization of an ensemble context at once.  ep: the
// RK4 epilogue of the kernels that write the slope (rk4_fusable contexts)
int rhs_core(ihg_ctx *c, double *f, double *result, int output, int kept,
             int nr = 1, const Rk4Ep *ep = nullptr);
// whether the RK4 epilogue is compiled for the context's kernels (FP64, and
// with GW on y lines of 16 to 128 points)
bool rk4_fusable(const ihg_ctx *c);
int features(ihg_ctx *c, const double *f, double out[3]);
int default_candidate_x(int len);
void candidate_counts(const ihg_ctx *c, int n[2]);
int autotune(ihg_ctx *c, int reps, int out[2], float ms_out[16]);
int rhs_filter(ihg_ctx *c, double *f, int nr = 1);
int rhs_gw(ihg_ctx *c, const double *f, double *result, int nr = 1,
           const Rk4Ep *ep = nullptr);
int rhs_output(ihg_ctx *c, const double *f, int kept);
int initial_psi(ihg_ctx *c, double *field, double *dfield);

int set_error(const char *fmt, ...);
int hip_check(hipError_t e, const char *what);

// profiler hooks (ihg_api.hip): bracket one launch on the context stream
int prof_begin(ihg_ctx *c, int fam, double bytes);
int prof_end(ihg_ctx *c);

// launch(F) between two events when the profiler is on
template <class F>
int timed(ihg_ctx *c, int fam, double bytes, F &&launch)
{
    if (c->prof_on) {
        int rc = prof_begin(c, fam, bytes);
        if (rc) return rc;
    }
    int rc = launch();
    if (rc) return rc;
    return c->prof_on ? prof_end(c) : 0;
}

} // namespace ihg

#define IHG_TRY(expr)                                                         \
    do {                                                                      \
        int _rc = (expr);                                                     \
        if (_rc) return _rc;                                                  \
    } while (0)

#define IHG_HIP(expr) IHG_TRY(::ihg::hip_check((expr), #expr))
#define IHG_LAUNCHED(name) IHG_HIP(hipGetLastError())
