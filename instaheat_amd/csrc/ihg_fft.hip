// ihg_fft.hip -- launchers of the 3D r2c / c2r passes (fft / ifft of
// src/toolbox.c:806-827).
#include <type_traits>

#include "dispatch.h"
#include "fft_kernels.h"
#include "ihg_internal.h"

namespace ihg {

static LineSet y_lines(const ihg_ctx *c)
{
    // lines along y: one per (x, kz); points Mz apart
    return LineSet{(long)c->Nx * c->Mz, c->Mz, (long)c->Ny * c->Mz, c->Mz};
}

static LineSet x_lines(const ihg_ctx *c)
{
    // lines along x: one per (y, kz); points Ny*Mz apart
    const long plane = (long)c->Ny * c->Mz;
    return LineSet{plane, plane, 0, plane};
}

template <int L, class C>
static int launch_c2c(ihg_ctx *c, const C *in, C *out, LineSet ls, int dir,
                      int batch, long cbatch, int fam, const C *tw, int nr,
                      long rr, int cand)
{
    // algorithmic bytes: every complex value read once and written once
    const double bytes = 2.0 * sizeof(C) * c->M * batch * nr;
    return with_cand(cand, typename StridedCands<L>::type{}, [&](auto cd) {
        using CD = decltype(cd);
        using P = typename CD::P;
        constexpr int B = CD::WG / (L / P::E);
        dim3 grid((unsigned)((ls.nlines + B - 1) / B), batch, nr);
        return timed(c, fam, bytes, [&]() {
            if (dir < 0)
                hipLaunchKernelGGL((k_c2c_strided<L, -1, C, P, CD::WG, CD::SWZ>), grid,
                                   dim3(CD::WG), 0, c->stream, in, out, ls,
                                   cbatch, tw, rr);
            else
                hipLaunchKernelGGL((k_c2c_strided<L, +1, C, P, CD::WG, CD::SWZ>), grid,
                                   dim3(CD::WG), 0, c->stream, in, out, ls,
                                   cbatch, tw, rr);
            return hip_check(hipGetLastError(), "k_c2c_strided");
        });
    });
}

int y_pass(ihg_ctx *c, double2 *data, int dir, int batch, long cbatch,
           int nr, long rr)
{
    return dispatch_len(c->Ny, [&](auto Lc) {
        return launch_c2c<decltype(Lc)::value>(c, data, data, y_lines(c), dir,
                                               batch, cbatch, F_Y_PASS, c->tw,
                                               nr, rr, c->cand_y);
    });
}

int x_pass(ihg_ctx *c, double2 *data, int dir, int batch, long cbatch,
           int nr, long rr)
{
    return dispatch_len(c->Nx, [&](auto Lc) {
        return launch_c2c<decltype(Lc)::value>(c, data, data, x_lines(c), dir,
                                               batch, cbatch, F_X_PASS, c->tw,
                                               nr, rr, c->cand_x);
    });
}

int x_pass_f32(ihg_ctx *c, float2 *data, int dir, int batch, long cbatch,
               int nr, long rr)
{
    return dispatch_len(c->Nx, [&](auto Lc) {
        return launch_c2c<decltype(Lc)::value>(c, data, data, x_lines(c), dir,
                                               batch, cbatch, F_X_PASS, c->twf,
                                               nr, rr, c->cand_x);
    });
}

static int x_pass_oop(ihg_ctx *c, const double2 *in, double2 *out, int dir,
                      int batch, long cbatch)
{
    return dispatch_len(c->Nx, [&](auto Lc) {
        return launch_c2c<decltype(Lc)::value>(c, in, out, x_lines(c), dir,
                                               batch, cbatch, F_X_PASS, c->tw,
                                               1, 0L, c->cand_x);
    });
}

int z_r2c_pass(ihg_ctx *c, const double *in, double2 *out, int batch,
               long rbatch, long cbatch, int nr, long rin, long rout)
{
    const long nrows = (long)c->Nx * c->Ny;
    return dispatch_len(c->Nz, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        using TL = Tile<L>;
        dim3 grid((unsigned)((nrows / 2 + TL::B - 1) / TL::B), batch, nr);
        const double bytes = (8.0 * c->N + 16.0 * c->M) * batch * nr;
        return timed(c, F_Z_R2C, bytes, [&]() {
            hipLaunchKernelGGL((k_z_r2c<L>), grid, dim3(FFT_WG), 0, c->stream,
                               in, out, nrows, rbatch, cbatch, c->tw, rin, rout);
            return hip_check(hipGetLastError(), "k_z_r2c");
        });
    });
}

int z_c2r_pass(ihg_ctx *c, const double2 *in, double *out, int batch,
               long cbatch, long rbatch, int nr, long rin, long rout)
{
    const long nrows = (long)c->Nx * c->Ny;
    return dispatch_len(c->Nz, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        using TL = Tile<L>;
        dim3 grid((unsigned)((nrows / 2 + TL::B - 1) / TL::B), batch, nr);
        const double bytes = (8.0 * c->N + 16.0 * c->M) * batch * nr;
        return timed(c, F_Z_C2R, bytes, [&]() {
            hipLaunchKernelGGL((k_z_c2r<L>), grid, dim3(FFT_WG), 0, c->stream,
                               in, out, nrows, cbatch, rbatch, c->tw, rin, rout);
            return hip_check(hipGetLastError(), "k_z_c2r");
        });
    });
}

int fft_r2c(ihg_ctx *c, const double *in, double2 *out, int batch,
            long rbatch, long cbatch)
{
    IHG_TRY(z_r2c_pass(c, in, out, batch, rbatch, cbatch));
    IHG_TRY(y_pass(c, out, -1, batch, cbatch));
    return x_pass(c, out, -1, batch, cbatch);
}

int fft_c2r(ihg_ctx *c, const double2 *in, double *out, double2 *work,
            int batch, long cbatch, long rbatch)
{
    IHG_TRY(x_pass_oop(c, in, work, +1, batch, cbatch));
    IHG_TRY(y_pass(c, work, +1, batch, cbatch));
    return z_c2r_pass(c, work, out, batch, cbatch, rbatch);
}

// Autotuner of the strided passes: every (plan, work-group) candidate of the
// y and of the x line length is timed on the context's own work arrays, in
// the shape the RHS uses most (four arrays, all realizations of an ensemble),
// and the fastest is kept.  Results of different candidates differ only in
// rounding.
template <int L>
static int n_cands()
{
    return StridedCands<L>::type::n;
}

int default_candidate_x(int len)
{
    return dispatch_len(len, [&](auto Lc) {
        return StridedCands<decltype(Lc)::value>::x_default;
    });
}

void candidate_counts(const ihg_ctx *c, int n[2])
{
    n[0] = dispatch_len(c->Ny, [&](auto Lc) { return n_cands<decltype(Lc)::value>(); });
    n[1] = dispatch_len(c->Nx, [&](auto Lc) { return n_cands<decltype(Lc)::value>(); });
}

int autotune(ihg_ctx *c, int reps, int out[2], float ms_out[16])
{
    const int nr = c->nreal;
    const long rr = nr > 1 ? c->ens.w : 0;
    hipEvent_t a, b;
    IHG_TRY(hip_check(hipEventCreate(&a), "hipEventCreate"));
    IHG_TRY(hip_check(hipEventCreate(&b), "hipEventCreate"));
    const bool prof = c->prof_on;
    c->prof_on = false;
    int rc = 0, k = 0;
    for (int axis = 0; axis < 2 && !rc; ++axis) {
        const int len = axis ? c->Nx : c->Ny;
        const int ncand = dispatch_len(len, [&](auto Lc) {
            return n_cands<decltype(Lc)::value>();
        });
        const int def = axis ? default_candidate_x(len) : 0;
        int &slot = axis ? c->cand_x : c->cand_y;
        // three rounds over all candidates, each candidate's fastest round
        // kept (drifts of the clock or of the other work hit every candidate)
        float tmin[16];
        for (int id = 0; id < 16; ++id)
            tmin[id] = 1e30f;
        for (int round = 0; round < 3 && !rc; ++round) {
            for (int id = 0; id < ncand && id < 16 && !rc; ++id) {
                slot = id;
                auto pass = [&]() {
                    return axis ? x_pass(c, c->W, -1, 4, c->M, nr, rr)
                                : y_pass(c, c->W, -1, 4, c->M, nr, rr);
                };
                rc = pass(); // warm-up
                if (!rc) rc = hip_check(hipEventRecord(a, c->stream), "hipEventRecord");
                for (int r = 0; r < reps && !rc; ++r) rc = pass();
                if (!rc) rc = hip_check(hipEventRecord(b, c->stream), "hipEventRecord");
                if (!rc) rc = hip_check(hipEventSynchronize(b), "hipEventSynchronize");
                float ms = 0.0f;
                if (!rc) rc = hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
                if (!rc && ms < tmin[id]) tmin[id] = ms;
            }
        }
        // the default stays unless a candidate is at least 3 % faster
        int pick = def;
        for (int id = 0; id < ncand && id < 16; ++id)
            if (tmin[id] < 0.97f * tmin[def] && tmin[id] < tmin[pick]) pick = id;
        for (int id = 0; id < ncand && id < 16; ++id, ++k)
            if (ms_out && k < 16) ms_out[k] = tmin[id] / reps;
        slot = pick;
        if (out) out[axis] = pick;
    }
    c->prof_on = prof;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

} // namespace ihg
