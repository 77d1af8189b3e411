// fft_kernels.h -- FFT pass kernels for the 3D r2c / c2r transforms.
//
// Layout (FFTW r2c_3d, the state layout of the reference): a real field is
// double[Nx][Ny][Nz] with z fastest; its half spectrum is
// complex[Nx][Ny][Nz/2+1].  A 3D r2c is a z pass (real rows -> half rows),
// then a y pass and an x pass on the half spectrum; c2r is the mirror.
//
// z passes transform two real rows at once as one complex row (real part one
// row, imaginary part the other) and split / merge the two Hermitian halves
// in LDS.  y and x passes are strided: a work-group owns B neighbouring lines,
// so every global access is a run of B consecutive complex values.
#pragma once

#include "fft_core.h"
#include "mem.h"

namespace ihg {

constexpr int FFT_WG = 256;

// Line plans of the z passes.  Their transposes are wavefront-local, so an
// extra stage costs LDS traffic but no work-group barrier, and fewer values
// per thread buy occupancy: 256 points take 8 values per thread in three
// stages (at 16 values the z c2r + rho pass held 182 registers and 70 KB of
// LDS per work-group, 2 waves per SIMD).  (512 points keep 16 values: a
// wavefront must hold both lines of a row pair.)
template <int L> struct ZPlanOf { using P = Plan<L>; };
template <> struct ZPlanOf<256> { using P = PlanT<8, 8, 8, 4>; };

// Tile of a z pass: B lines of L points per work-group, T threads per line
template <int L>
struct Tile {
    using P = typename ZPlanOf<L>::P;
    static constexpr int E = P::E;
    static constexpr int T = L / E;
    static constexpr int B = FFT_WG / T;
    static constexpr int PITCH = Lds<L>::PITCH;
    // lines per wavefront (a line's T threads are consecutive lanes), so
    // each wavefront owns LPW lines and finishes them alone
    static constexpr int LPW = 64 / T;
};

// Geometry of a set of strided lines: line l starts at
// (l / inner) * outer_stride + (l % inner) and its points are `stride` apart.
struct LineSet {
    long nlines, inner, outer_stride, stride;
    __device__ __forceinline__ long base(long l) const
    {
        return (l / inner) * outer_stride + (l % inner);
    }
};

// Work-group coordinates of one pass: a standalone launch takes them from
// the hardware (hw_blk), a grouped launch -- passes of several grid sizes in
// one launch -- from its segment table.  gx is the number of tiles along x.
struct Blk {
    unsigned x, y, z, gx;
};

__device__ __forceinline__ Blk hw_blk()
{
    return Blk{blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x};
}

// Tile index of a work-group along x.  The dispatcher hands work-groups to
// the eight XCDs in turn (work-group w to XCD w mod 8, w the linear
// work-group id; w and w + 8 land on one XCD whatever the other grid
// dimensions are), and every XCD has its own L2.  Strided tiles that are
// neighbours share a 128-byte line at their edges (the half-spectrum rows of
// Nz/2 + 1 complex values start off line boundaries), so with SWZ each XCD
// gets one contiguous run of tiles: XCD x takes tiles
// x*q + min(x, r) + [0, q + (x < r)), q = gx / 8, r = gx mod 8, in order.
// The map is a bijection of [0, gx).
template <bool SWZ>
__device__ __forceinline__ unsigned tile_x(const Blk &bk)
{
    if constexpr (!SWZ) {
        return bk.x;
    } else {
        const unsigned gx = bk.gx, bx = bk.x;
        const unsigned q = gx / 8, r = gx % 8, x = bx % 8;
        return x * q + (x < r ? x : r) + bx / 8;
    }
}

// ---------------------------------------------------------------------------
// strided complex-to-complex pass (y or x), in place or out of place,
// batched over blockIdx.y (arrays `batch` complex values apart).
// Lanes run fastest over the B lines of the tile, so for a fixed point n a
// wavefront touches B consecutive complex values.  The first Stockham stage
// loads straight from HBM into registers and the last stage stores straight
// from registers; only the transposes between stages go through LDS.
// ---------------------------------------------------------------------------
// The line plan P and the work-group size WG are template parameters so
// that an autotuner can pick them per line length (B = WG / (L / P::E)
// lines per work-group).
template <int L, class P = Plan<L>, int WG = FFT_WG>
struct StridedShape {
    static constexpr int B = WG / (L / P::E), PITCH = Lds<L>::PITCH;
    static constexpr int LDS = B * PITCH; // complex values of LDS
};

template <int L, int DIR, class C = double2, class P = Plan<L>,
          int WG = FFT_WG, bool SWZ = true>
__device__ __forceinline__ void c2c_strided_body(
    const Blk &bk, C *lds, const C *in, C *out, LineSet ls, long batch,
    const C *__restrict__ tw, long rr)
{
    constexpr int B = StridedShape<L, P, WG>::B, PITCH = Lds<L>::PITCH;
    const int tid = threadIdx.x;
    in += bk.y * batch + bk.z * rr;
    out += bk.y * batch + bk.z * rr;
    const int b = tid % B, j = tid / B;
    const long line = (long)tile_x<SWZ>(bk) * B + b;
    const bool act = line < ls.nlines;
    const long base = ls.base(act ? line : 0);
    const C *src = in + base;
    C *dst = out + base;
    const long st = ls.stride;
    auto rd = [&](int, int n) {
        const C v = ldn(src + n * st); // in bounds for every lane
        return act ? v : CT<C>::make(0, 0);
    };
    auto wr = [&](int, int n, C v) {
        if (act) stn(dst + n * st, v);
    };
    fft_line<L, DIR, true, true, P>(lds + b * PITCH, j, tw, rd, wr);
}

template <int L, int DIR, class C = double2, class P = Plan<L>,
          int WG = FFT_WG, bool SWZ = true>
__global__ __launch_bounds__(WG) void k_c2c_strided(
    const C *in, C *out, LineSet ls, long batch, const C *__restrict__ tw,
    long rr)
{
    __shared__ C lds[StridedShape<L, P, WG>::LDS];
    c2c_strided_body<L, DIR, C, P, WG, SWZ>(hw_blk(), lds, in, out, ls, batch,
                                            tw, rr);
}

// Candidate (plan, work-group size) pairs of the strided pass per line
// length; entry 0 is the default.  The autotuner times them on the
// context's own arrays and keeps the fastest.
template <class P_, int WG_, bool SWZ_ = true> struct StridedCand {
    using P = P_;
    static constexpr int WG = WG_;
    static constexpr bool SWZ = SWZ_;
};
template <class... Cs> struct CandList {
    static constexpr int n = sizeof...(Cs);
};
template <int L> struct StridedCands {
    using type = CandList<StridedCand<Plan<L>, 256>, StridedCand<Plan<L>, 256, false>>;
    static constexpr int x_default = 0;
};
template <> struct StridedCands<16> {
    using type = CandList<StridedCand<Plan<16>, 256>, StridedCand<Plan<16>, 256, false>,
                          StridedCand<PlanT<4, 4, 4>, 256>, StridedCand<Plan<16>, 128>>;
    static constexpr int x_default = 0;
};
template <> struct StridedCands<32> {
    using type = CandList<StridedCand<Plan<32>, 256>, StridedCand<Plan<32>, 256, false>,
                          StridedCand<PlanT<16, 16, 2>, 256>,
                          StridedCand<PlanT<4, 4, 4, 2>, 256>, StridedCand<Plan<32>, 128>>;
    static constexpr int x_default = 0;
};
template <> struct StridedCands<64> {
    using type = CandList<StridedCand<Plan<64>, 256>, StridedCand<Plan<64>, 256, false>,
                          StridedCand<PlanT<16, 16, 4>, 256>,
                          StridedCand<PlanT<4, 4, 4, 4>, 256>, StridedCand<Plan<64>, 128>>;
    static constexpr int x_default = 0;
};
template <> struct StridedCands<128> {
    using type = CandList<StridedCand<Plan<128>, 256>, StridedCand<Plan<128>, 256, false>,
                          StridedCand<PlanT<16, 16, 8>, 256>,
                          StridedCand<PlanT<4, 4, 4, 4, 2>, 256>, StridedCand<Plan<128>, 128>>;
    static constexpr int x_default = 0;
};
// 256-point x lines (points a whole y-z plane apart) run fastest with 32
// lines per work-group (512 B contiguous per point); y lines with 16.
template <> struct StridedCands<256> {
    using type = CandList<StridedCand<Plan<256>, 256>, StridedCand<Plan<256>, 256, false>,
                          StridedCand<Plan<256>, 512>, StridedCand<PlanT<8, 8, 8, 4>, 512>,
                          StridedCand<PlanT<4, 4, 4, 4, 4>, 512, false>>;
    static constexpr int x_default = 2;
};

// f(StridedCand<...>{}) for candidate `id` of a list (the default for an id
// out of range)
template <class F, class C0, class... Cs>
auto with_cand(int id, CandList<C0, Cs...>, F &&f)
{
    if constexpr (sizeof...(Cs) == 0) {
        return f(C0{});
    } else {
        if (id <= 0) return f(C0{});
        return with_cand(id - 1, CandList<Cs...>{}, f);
    }
}

// ---------------------------------------------------------------------------
// x-axis forward transform followed by NQ backward transforms of the same
// spectrum, each multiplied by its own factor fac(q, line, kx_index, value):
// the spectral derivatives of the right-hand side (NQ = 4, out of place,
// output q at out + q * qstride) and the filter round trip (NQ = 1, in
// place, batched over blockIdx.y with arrays `batch` apart); realizations
// of an ensemble over blockIdx.z, `rr` apart.
// The spectrum never leaves registers: the forward transform's last stage
// keeps its outputs, and each backward transform's first stage takes them
// through chain_slot; HBM is read once and written NQ times.  With FWD =
// false the input is already the full spectrum and only the backward
// transforms run.
// ---------------------------------------------------------------------------
template <int L, int NQ, bool FWD, class Fac, class P = Plan<L>,
          int WG = FFT_WG, bool SWZ = true>
__device__ __forceinline__ void x_fanout_body(
    const Blk &bk, double2 *lds, const double2 *in, double2 *out,
    long qstride, LineSet ls, long batch, const double2 *__restrict__ tw,
    Fac fac, long rr)
{
    constexpr int E = P::E, B = StridedShape<L, P, WG>::B, PITCH = Lds<L>::PITCH;
    const int tid = threadIdx.x;
    in += bk.y * batch + bk.z * rr;
    out += bk.y * batch + bk.z * rr;
    const int b = tid % B, j = tid / B;
    const long line = (long)tile_x<SWZ>(bk) * B + b;
    const bool act = line < ls.nlines;
    const long base = ls.base(act ? line : 0);
    const double2 *src = in + base;
    const long st = ls.stride;
    double2 *my = lds + b * PITCH;
    double2 spec[E];
    auto rd = [&](int, int n) {
        const double2 v = ldn(src + n * st); // in bounds for every lane
        return act ? v : make_double2(0.0, 0.0);
    };
    if constexpr (FWD) {
        auto keep = [&](int slot, int, double2 v) { spec[slot] = v; };
        fft_line<L, -1, true, true, P>(my, j, tw, rd, keep);
    } else {
#pragma unroll
        for (int slot = 0; slot < E; ++slot)
            spec[slot] = rd(slot, stage1_point<L, P>(j, slot));
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        if (FWD || q > 0)
            lds_barrier(); // LDS reads of the previous transform are done
        double2 *dst = out + q * qstride + base;
        auto rq = [&](int slot, int n) {
            return fac(bk, q, line, n, spec[FWD ? chain_slot<L, P>(slot) : slot]);
        };
        auto wq = [&](int, int n, double2 v) {
            if (act) stn(dst + n * st, v);
        };
        fft_line<L, +1, true, true, P>(my, j, tw, rq, wq);
    }
}

template <int L, int NQ, bool FWD, class Fac, class P = Plan<L>,
          int WG = FFT_WG, bool SWZ = true>
__global__ __launch_bounds__(WG) void k_x_fanout(
    const double2 *in, double2 *out, long qstride, LineSet ls, long batch,
    const double2 *__restrict__ tw, Fac fac, long rr)
{
    __shared__ double2 lds[StridedShape<L, P, WG>::LDS];
    x_fanout_body<L, NQ, FWD, Fac, P, WG, SWZ>(hw_blk(), lds, in, out, qstride,
                                               ls, batch, tw, fac, rr);
}

// Line plans of the derivative fan-out.  (At 256 points, 8 values per
// thread in three stages -- PlanT<8, 8, 8, 4>, 166 VGPRs and 3 waves per
// SIMD instead of 1 -- measured 209-219 us against 190-194 us for the
// two-stage 16-value plan, with the XCD-aware tile order on both, and
// still 198-210 against 172-180 us with the stage twiddles read from LDS
// (158 VGPRs); the 16-value plan loads its twiddles once for all five
// transforms, so LDS twiddles gain it nothing either (181-184 us).  The
// default plans stay.)
template <int L> struct FanPlan { using P = Plan<L>; static constexpr int WG = FFT_WG; };

// ---------------------------------------------------------------------------
// z passes
// ---------------------------------------------------------------------------

// Hermitian split after a forward transform of Z = A + iB (A, B real rows):
// A[k] = (Z[k] + conj Z[L-k]) / 2, B[k] = -i (Z[k] - conj Z[L-k]) / 2.
template <int L, class C>
__device__ __forceinline__ C herm_split(const C *line, int k, int s)
{
    using R = typename CT<C>::R;
    const C z1 = line[Lds<L>::idx(k)];
    const C z2 = cconj(line[Lds<L>::idx((L - k) & (L - 1))]);
    const R h = (R)0.5;
    if (s == 0)
        return CT<C>::make(h * (z1.x + z2.x), h * (z1.y + z2.y));
    return CT<C>::make(h * (z1.y - z2.y), -h * (z1.x - z2.x));
}

// Merge of two half rows A, B (Hermitian extension, imaginary parts of the
// k = 0 and k = L/2 bins dropped) into Z[n] = A_ext[n] + i B_ext[n].
// Written without branches: the mirrored index and the sign / zeroing of the
// imaginary parts are selects, so no value the caller computes is left to
// divergent control flow.
template <int L>
__device__ __forceinline__ double2 herm_merge(const double2 *__restrict__ A,
                                              const double2 *__restrict__ Bv,
                                              int n)
{
    const bool upper = n > L / 2;
    const int m = upper ? L - n : n;
    const double2 a = A[m], bb = Bv[m];
    // imaginary parts of the k = 0 and k = L/2 bins are dropped; the upper
    // half uses the conjugates
    const double edge = (m == 0 || m == L / 2) ? 0.0 : 1.0;
    const double sg = upper ? -1.0 : 1.0;
    const double ai = sg * edge * a.y, bi = sg * edge * bb.y;
    return make_double2(a.x - bi, ai + bb.x);
}

// Real rows -> half rows.  Line l = rows (2l, 2l+1) of `in` (nrows rows of
// L reals, batched over y with `rbatch` / `cbatch` strides, realizations
// over z `rin` / `rout` apart).
template <int L>
__device__ __forceinline__ void z_r2c_body(
    const Blk &bk, double2 *lds, const double *in, double2 *out, long nrows,
    long rbatch, long cbatch, const double2 *__restrict__ tw, long rin,
    long rout)
{
    using TL = Tile<L>;
    constexpr int MZ = L / 2 + 1;
    const int tid = threadIdx.x;
    in += bk.y * rbatch + bk.z * rin;
    out += bk.y * cbatch + bk.z * rout;
    const long nlines = nrows / 2;
    const int bc = tid / TL::T;
    const long line = (long)bk.x * TL::B + bc;
    const bool act = line < nlines;
    const double *ra = in + (act ? 2 * line : 0) * L;
    auto rd = [&](int, int n) {
        const double u = ldn(ra + n), v = ldn(ra + L + n); // in bounds
        return act ? make_double2(u, v) : make_double2(0.0, 0.0);
    };
    fft_line_wave<L, -1, typename TL::P>(lds + bc * TL::PITCH, tid % TL::T, tw, rd);
    // each wavefront splits and stores the 2 LPW rows of its own lines
    const int w = tid / 64, lane = tid % 64;
    const long row0 = 2 * ((long)bk.x * TL::B + w * TL::LPW);
    const double2 *wl = lds + w * TL::LPW * TL::PITCH;
    for (int idx = lane; idx < 2 * TL::LPW * MZ; idx += 64) {
        const int r = idx / MZ, k = idx - r * MZ;
        if (row0 + r < nrows)
            stn(out + row0 * MZ + idx,
                herm_split<L>(wl + (r >> 1) * TL::PITCH, k, r & 1));
    }
}

template <int L>
__global__ __launch_bounds__(FFT_WG) void k_z_r2c(
    const double *in, double2 *out, long nrows, long rbatch, long cbatch,
    const double2 *__restrict__ tw, long rin, long rout)
{
    __shared__ double2 lds[Tile<L>::B * Tile<L>::PITCH];
    z_r2c_body<L>(hw_blk(), lds, in, out, nrows, rbatch, cbatch, tw, rin, rout);
}

// Half rows -> real rows (the c2r z pass).
template <int L>
__device__ __forceinline__ void z_c2r_body(
    const Blk &bk, double2 *lds, const double2 *in, double *out, long nrows,
    long cbatch, long rbatch, const double2 *__restrict__ tw, long rin,
    long rout)
{
    using TL = Tile<L>;
    constexpr int MZ = L / 2 + 1;
    const int tid = threadIdx.x;
    in += bk.y * cbatch + bk.z * rin;
    out += bk.y * rbatch + bk.z * rout;
    const long nlines = nrows / 2;
    const int bc = tid / TL::T;
    const long line = (long)bk.x * TL::B + bc;
    const bool act = line < nlines;
    const double2 *A = in + (act ? 2 * line : 0) * MZ;
    auto rd = [&](int, int n) {
        const double2 z = herm_merge<L>(A, A + MZ, n); // in bounds for every lane
        return act ? z : make_double2(0.0, 0.0);
    };
    fft_line_wave<L, +1, typename TL::P>(lds + bc * TL::PITCH, tid % TL::T, tw, rd);
    // each wavefront stores the 2 LPW rows of its own lines
    const int w = tid / 64, lane = tid % 64;
    const long row0 = 2 * ((long)bk.x * TL::B + w * TL::LPW);
    const double2 *wl = lds + w * TL::LPW * TL::PITCH;
    for (int idx = lane; idx < 2 * TL::LPW * L; idx += 64) {
        const int r = idx / L, n = idx - r * L;
        if (row0 + r < nrows) {
            const double2 z = wl[(r >> 1) * TL::PITCH + Lds<L>::idx(n)];
            stn(out + row0 * L + idx, (r & 1) ? z.y : z.x);
        }
    }
}

template <int L>
__global__ __launch_bounds__(FFT_WG) void k_z_c2r(
    const double2 *in, double *out, long nrows, long cbatch, long rbatch,
    const double2 *__restrict__ tw, long rin, long rout)
{
    __shared__ double2 lds[Tile<L>::B * Tile<L>::PITCH];
    z_c2r_body<L>(hw_blk(), lds, in, out, nrows, cbatch, rbatch, tw, rin, rout);
}

} // namespace ihg
