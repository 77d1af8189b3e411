// fft_core.h -- FP64 complex FFT building blocks for gfx950.
//
// A line of L complex points is transformed by a work-group slice of T = L/E
// threads.  Each stage is a radix-R Stockham pass: a thread loads E values
// (E/R butterflies), applies twiddles, does the R-point DFTs in registers and
// writes the results back to LDS in Stockham order, so that after the last
// stage the line sits in LDS in natural order.  The first stage can read its
// inputs through a caller-supplied functor (straight from HBM for the
// contiguous z axis), every later stage reads LDS.
//
// Sign convention: DIR = -1 is the forward transform (e^{-i}), DIR = +1 the
// backward one; both unnormalised, matching FFTW's r2c/c2r.
#pragma once

#include <hip/hip_runtime.h>

namespace ihg {

// complex value types: double2 (FP64 path) and float2 (the FP32 GW source)
template <class C> struct CT;
template <> struct CT<double2> {
    using R = double;
    __device__ static __forceinline__ double2 make(double a, double b)
    {
        return make_double2(a, b);
    }
};
template <> struct CT<float2> {
    using R = float;
    __device__ static __forceinline__ float2 make(float a, float b)
    {
        return make_float2(a, b);
    }
};

template <class C>
__device__ __forceinline__ C cadd(C a, C b)
{
    return CT<C>::make(a.x + b.x, a.y + b.y);
}
template <class C>
__device__ __forceinline__ C csub(C a, C b)
{
    return CT<C>::make(a.x - b.x, a.y - b.y);
}
template <class C>
__device__ __forceinline__ C cmul(C a, C b)
{
    return CT<C>::make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class C>
__device__ __forceinline__ C cconj(C a)
{
    return CT<C>::make(a.x, -a.y);
}
// a * (DIR * i)
template <int DIR, class C>
__device__ __forceinline__ C mul_i(C a)
{
    if constexpr (DIR < 0)
        return CT<C>::make(a.y, -a.x);
    else
        return CT<C>::make(-a.y, a.x);
}

// a * e^{DIR * 2 pi i m / 16} for a compile-time m
template <int DIR, int M, class C>
__device__ __forceinline__ C rot16(C a)
{
    using R = typename CT<C>::R;
    constexpr int m = M & 15;
    constexpr double C1 = 0.92387953251128675613; // cos(pi/8)
    constexpr double S1 = 0.38268343236508977173; // sin(pi/8)
    constexpr double H = 0.70710678118654752440;  // cos(pi/4)
    if constexpr (m == 0) {
        return a;
    } else if constexpr (m == 4) {
        return mul_i<DIR>(a);
    } else if constexpr (m == 8) {
        return CT<C>::make(-a.x, -a.y);
    } else if constexpr (m == 12) {
        return mul_i<-DIR>(a);
    } else {
        // cos / sin of 2 pi m / 16 for m in the four quadrants
        constexpr double cs[16] = {1, C1, H, S1, 0, -S1, -H, -C1,
                                   -1, -C1, -H, -S1, 0, S1, H, C1};
        constexpr double sn[16] = {0, S1, H, C1, 1, C1, H, S1,
                                   0, -S1, -H, -C1, -1, -C1, -H, -S1};
        constexpr R c = (R)cs[m];
        constexpr R s = (R)(DIR * sn[m]);
        return CT<C>::make(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

template <int DIR, class C>
__device__ __forceinline__ void dft2(C *v)
{
    C t = v[0];
    v[0] = cadd(t, v[1]);
    v[1] = csub(t, v[1]);
}

template <int DIR, class C>
__device__ __forceinline__ void dft4(C &x0, C &x1, C &x2, C &x3)
{
    C a0 = cadd(x0, x2), a1 = csub(x0, x2);
    C b0 = cadd(x1, x3), b1 = mul_i<DIR>(csub(x1, x3));
    x0 = cadd(a0, b0);
    x2 = csub(a0, b0);
    x1 = cadd(a1, b1);
    x3 = csub(a1, b1);
}

template <int DIR, class C>
__device__ __forceinline__ void dft4(C *v)
{
    dft4<DIR>(v[0], v[1], v[2], v[3]);
}

template <int DIR, class C>
__device__ __forceinline__ void dft8(C *v)
{
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    o1 = rot16<DIR, 2>(o1);
    o2 = rot16<DIR, 4>(o2);
    o3 = rot16<DIR, 6>(o3);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

// 16 = 4 x 4: n = 4 p + q, k = k1 + 4 k2
template <int DIR, class C>
__device__ __forceinline__ void dft16(C *v)
{
    C y[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        y[q][0] = v[q];
        y[q][1] = v[4 + q];
        y[q][2] = v[8 + q];
        y[q][3] = v[12 + q];
        dft4<DIR>(y[q][0], y[q][1], y[q][2], y[q][3]);
    }
    y[1][1] = rot16<DIR, 1>(y[1][1]);
    y[1][2] = rot16<DIR, 2>(y[1][2]);
    y[1][3] = rot16<DIR, 3>(y[1][3]);
    y[2][1] = rot16<DIR, 2>(y[2][1]);
    y[2][2] = rot16<DIR, 4>(y[2][2]);
    y[2][3] = rot16<DIR, 6>(y[2][3]);
    y[3][1] = rot16<DIR, 3>(y[3][1]);
    y[3][2] = rot16<DIR, 6>(y[3][2]);
    y[3][3] = rot16<DIR, 9>(y[3][3]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        dft4<DIR>(y[0][k1], y[1][k1], y[2][k1], y[3][k1]);
        v[k1] = y[0][k1];
        v[k1 + 4] = y[1][k1];
        v[k1 + 8] = y[2][k1];
        v[k1 + 12] = y[3][k1];
    }
}

template <int R, int DIR, class C>
__device__ __forceinline__ void dft(C *v)
{
    if constexpr (R == 2)
        dft2<DIR>(v);
    else if constexpr (R == 4)
        dft4<DIR>(v);
    else if constexpr (R == 8)
        dft8<DIR>(v);
    else
        dft16<DIR>(v);
}

// Factorisation of a line length: E values per thread, radices R0, R1, R2
// (R1 / R2 = 1 when absent).  Every radix divides E.
template <int E_, int R0_, int R1_ = 1, int R2_ = 1, int R3_ = 1>
struct PlanT {
    static constexpr int E = E_, R0 = R0_, R1 = R1_, R2 = R2_, R3 = R3_;
};
template <int L>
struct Plan;
template <> struct Plan<4>   { static constexpr int E = 4,  R0 = 4,  R1 = 1,  R2 = 1, R3 = 1; };
template <> struct Plan<8>   { static constexpr int E = 8,  R0 = 8,  R1 = 1,  R2 = 1, R3 = 1; };
template <> struct Plan<16>  { static constexpr int E = 16, R0 = 16, R1 = 1,  R2 = 1, R3 = 1; };
template <> struct Plan<32>  { static constexpr int E = 8,  R0 = 8,  R1 = 4,  R2 = 1, R3 = 1; };
template <> struct Plan<64>  { static constexpr int E = 8,  R0 = 8,  R1 = 8,  R2 = 1, R3 = 1; };
// (128 points: 8 values per thread in three stages -- half the LDS and
// registers of the two-stage 16-value plan, which measured 4 % slower for a
// 4 x 128^3 ensemble)
template <> struct Plan<128> { static constexpr int E = 8,  R0 = 8,  R1 = 8,  R2 = 2, R3 = 1; };
template <> struct Plan<256> { static constexpr int E = 16, R0 = 16, R1 = 16, R2 = 1, R3 = 1; };
template <> struct Plan<512> { static constexpr int E = 16, R0 = 16, R1 = 16, R2 = 2, R3 = 1; };

// radix of the last stage of a plan
template <class P>
__host__ __device__ constexpr int last_radix()
{
    return P::R3 > 1 ? P::R3 : (P::R2 > 1 ? P::R2 : (P::R1 > 1 ? P::R1 : P::R0));
}

// LDS geometry of a tile of B lines: one pad slot every 16 points and an odd
// line pitch, so that both row-wise and column-wise sweeps spread over banks.
template <int L>
struct Lds {
    static constexpr int PITCH = L + L / 16 + 1;
    __device__ static __forceinline__ int idx(int n) { return n + (n >> 4); }
};

// Synchronisation of the threads of a line between two Stockham stages.
// WAVE = false: the line's threads may sit in several wavefronts (strided
// passes put lanes on B neighbouring lines), so the work-group barrier.
// WAVE = true: the line's threads are lanes of one wavefront (z passes,
// where a line's L / E threads are consecutive and L / E divides 64), and
// the wavefront's LDS accesses are performed in order: only the compiler
// must keep them in order, and the other wavefronts of the work-group run
// on without waiting (the transposes are wavefront-local exchanges).
// Work-group barrier of an LDS exchange.  __syncthreads() fences every
// address space at work-group scope, which waits for all of the wave's
// outstanding global loads and stores too (s_waitcnt vmcnt(0)): loads issued
// ahead of a transform, or the stores of the previous one, would be drained
// at every stage.  The exchanges need only the LDS operations done and
// ordered, so the fences are on the LDS address space alone (lgkmcnt(0) and
// s_barrier); a global load still in flight is waited for where its value is
// used.  No work-group barrier here orders global memory between threads.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <bool WAVE>
__device__ __forceinline__ void line_sync()
{
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        lds_barrier();
    }
}

// The stage twiddles tw[L, 2L) copied into the LDS table ltw by the calling
// wavefront, and the pointer through which stages index them as tw[L + i].
// Twiddles read from LDS are counted apart from the vector-memory loads,
// which return in order: a twiddle read from global memory waits for every
// load issued before it, so loads issued ahead of a transform would be
// drained at its second stage.  Every wavefront of a work-group writes the
// same values, so no work-group barrier is needed: a wavefront reads its
// own writes, or identical ones.  ltw must lie at least L values into the
// work-group's LDS.
template <int L, class C>
__device__ __forceinline__ const C *twiddles_to_lds_wave(C *ltw,
                                                         const C *__restrict__ tw)
{
    const int lane = (int)(threadIdx.x % 64);
#pragma unroll
    for (int k = 0; k < (L + 63) / 64; ++k)
        if (lane + 64 * k < L)
            ltw[lane + 64 * k] = tw[L + lane + 64 * k];
    line_sync<true>();
    return ltw - L;
}

// One Stockham stage of radix R on a line in LDS.  Ns = product of the
// radices of the earlier stages.  When FIRST, inputs come from rd(slot, n)
// (point n; slot = m * R + r is the register the value lands in) instead of
// LDS; when LAST_TO_WR, outputs go to wr(slot, n, value) instead of LDS.
// The stage reads `line` and writes `out`; when they are different buffers
// (ping-pong) no barrier is needed between the reads and the writes.
template <int L, int R, int Ns, int DIR, bool FIRST, bool LAST_TO_WR,
          class P = Plan<L>, class C, bool WAVE = false, class Rd, class Wr>
__device__ __forceinline__ void stockham_stage(C *line, C *out, int j,
                                               const C *__restrict__ tw,
                                               Rd &rd, Wr &wr)
{
    constexpr int E = P::E;
    constexpr int T = L / E;
    constexpr int NB = E / R; // butterflies per thread
    constexpr int STRIDE_IN = L / R;
    C v[NB][R];
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int jb = j + m * T;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int n = jb + r * STRIDE_IN;
            if constexpr (FIRST)
                v[m][r] = rd(m * R + r, n);
            else
                v[m][r] = line[Lds<L>::idx(n)];
        }
    }
    if constexpr (!FIRST && !LAST_TO_WR)
        if (out == line)
            line_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int jb = j + m * T;
        if constexpr (Ns > 1) {
            const int k = jb % Ns;
            constexpr int TS = L / (Ns * R);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C w = tw[L + k * r * TS];
                if constexpr (DIR > 0)
                    w = cconj(w);
                v[m][r] = cmul(v[m][r], w);
            }
        }
        dft<R, DIR>(v[m]);
        const int k = (Ns > 1) ? jb % Ns : 0;
        const int base = (jb / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (LAST_TO_WR)
                wr(m * R + r, base + r * Ns, v[m][r]);
            else
                out[Lds<L>::idx(base + r * Ns)] = v[m][r];
        }
    }
    if constexpr (!LAST_TO_WR)
        line_sync<WAVE>();
}

// Point index of register slot `slot` (= m * R0 + r) in the first stage of
// a line transform: n = j + m T + r L / R0.
template <int L, class P = Plan<L>>
__device__ __forceinline__ int stage1_point(int j, int slot)
{
    constexpr int T = L / P::E;
    return j + (slot / P::R0) * T + (slot % P::R0) * (L / P::R0);
}

// Point index held by register slot `slot` (= m * RL + r) after the last
// stage (radix RL) of a line transform that ends in registers:
// n = j + m T + r L / RL.
template <int L, class P = Plan<L>>
__device__ __forceinline__ int last_point(int j, int slot)
{
    constexpr int RL = last_radix<P>();
    constexpr int T = L / P::E;
    return j + (slot / RL) * T + (slot % RL) * (L / RL);
}

// Register slots when one transform feeds the next on the same thread.  The
// last stage of a transform leaves thread j holding points j + T u of the
// line (u < E), and the first stage of the next transform needs the same
// points: slot s = m R + r of the last stage (radix RL) holds
// u = m + r E / RL, first-stage slot s1 = m1 R0 + r1 needs
// u = m1 + r1 E / R0.  Returns the last-stage slot of first-stage slot s1.
template <int L, class P = Plan<L>>
__device__ __forceinline__ constexpr int chain_slot(int s1)
{
    constexpr int RL = last_radix<P>();
    constexpr int ML = P::E / RL;
    const int u = s1 / P::R0 + (s1 % P::R0) * (P::E / P::R0);
    return (u % ML) * RL + u / ML;
}

struct NoRead {
    __device__ double2 operator()(int, int) const { return make_double2(0.0, 0.0); }
};
struct NoWrite {
    template <class C>
    __device__ void operator()(int, int, C) const {}
};

// Full line transform.  On entry either the line is in LDS (FIRST_FROM_RD =
// false; callers must have synchronised) or `rd` supplies it.  On exit the
// transformed line is in LDS in natural order and the block is synchronised,
// unless LAST_TO_WR: then `wr` receives every output point and the LDS line
// may still be read by other threads of the block (no trailing barrier).
// WAVE: the line's threads are lanes of one wavefront (line_sync); on exit
// only this wavefront is synchronised.
template <int L, int DIR, bool FIRST_FROM_RD, bool LAST_TO_WR, class P,
          class C, bool WAVE, class Rd, class Wr>
__device__ __forceinline__ void fft_line_impl(C *line, C *line2, int j,
                                              const C *__restrict__ tw, Rd rd,
                                              Wr wr)
{
    static_assert(P::E % P::R0 == 0 && P::E % P::R1 == 0 &&
                      P::E % P::R2 == 0 && P::E % P::R3 == 0,
                  "every radix divides E");
    static_assert(P::R0 * P::R1 * P::R2 * P::R3 == L, "radices multiply to L");
    static_assert(!WAVE || 64 % (L / P::E) == 0,
                  "a line's threads within one wavefront");
    constexpr int NST = 1 + (P::R1 > 1) + (P::R2 > 1) + (P::R3 > 1);
    // buffers of stages 2.. (line2 == line: in place)
    C *const b0 = line, *const b1 = line2;
    stockham_stage<L, P::R0, 1, DIR, FIRST_FROM_RD, LAST_TO_WR && NST == 1, P,
                   C, WAVE>(b0, FIRST_FROM_RD ? b0 : b1, j, tw, rd, wr);
    C *const s2 = FIRST_FROM_RD ? b0 : b1, *const d2 = FIRST_FROM_RD ? b1 : b0;
    if constexpr (P::R1 > 1)
        stockham_stage<L, P::R1, P::R0, DIR, false, LAST_TO_WR && NST == 2, P,
                       C, WAVE>(s2, d2, j, tw, rd, wr);
    if constexpr (P::R2 > 1)
        stockham_stage<L, P::R2, P::R0 * P::R1, DIR, false,
                       LAST_TO_WR && NST == 3, P, C, WAVE>(d2, s2, j, tw, rd, wr);
    if constexpr (P::R3 > 1)
        stockham_stage<L, P::R3, P::R0 * P::R1 * P::R2, DIR, false, LAST_TO_WR,
                       P, C, WAVE>(s2, d2, j, tw, rd, wr);
}

// number of Stockham stages of a plan
template <class P>
constexpr int plan_stages()
{
    return 1 + (P::R1 > 1) + (P::R2 > 1) + (P::R3 > 1);
}

// radix of stage S of a plan, and the product of the radices before it
template <class P, int S>
__host__ __device__ constexpr int stage_radix()
{
    return S == 0 ? P::R0 : S == 1 ? P::R1 : S == 2 ? P::R2 : P::R3;
}
template <class P, int S>
__host__ __device__ constexpr int stage_span()
{
    if constexpr (S == 0)
        return 1;
    else
        return stage_span<P, S - 1>() * stage_radix<P, S - 1>();
}

// Stages S, S + 1, ... of a line transform on the LDS line `line` (in
// place), whose threads are lanes of one wavefront (wavefront-local
// exchanges); the last stage's outputs go to wr(slot, n, value).  On entry
// the line holds the output of stage S - 1 and the wavefront is
// synchronised with its writers.
template <int L, int DIR, class P, int S, class C, class Wr>
__device__ __forceinline__ void fft_stages_wave(C *line, int j,
                                                const C *__restrict__ tw,
                                                Wr &wr)
{
    constexpr int NST = plan_stages<P>();
    if constexpr (S < NST) {
        NoRead rd;
        stockham_stage<L, stage_radix<P, S>(), stage_span<P, S>(), DIR, false,
                       S == NST - 1, P, C, true>(line, line, j, tw, rd, wr);
        fft_stages_wave<L, DIR, P, S + 1, C>(line, j, tw, wr);
    }
}

template <int L, int DIR, bool FIRST_FROM_RD = false, bool LAST_TO_WR = false,
          class P = Plan<L>, class C = double2, class Rd = NoRead,
          class Wr = NoWrite>
__device__ __forceinline__ void fft_line(C *line, int j,
                                         const C *__restrict__ tw,
                                         Rd rd = Rd(), Wr wr = Wr())
{
    fft_line_impl<L, DIR, FIRST_FROM_RD, LAST_TO_WR, P, C, false>(line, line, j,
                                                                  tw, rd, wr);
}

// A line whose threads are lanes of one wavefront, first stage from `rd`,
// the result in LDS (the z passes): wavefront-local transposes.
template <int L, int DIR, class P = Plan<L>, class C = double2, class Rd>
__device__ __forceinline__ void fft_line_wave(C *line, int j,
                                              const C *__restrict__ tw, Rd rd)
{
    fft_line_impl<L, DIR, true, false, P, C, true>(line, line, j, tw, rd,
                                                   NoWrite());
}

} // namespace ihg
