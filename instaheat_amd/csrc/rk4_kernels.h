// rk4_kernels.h -- the stage arithmetic of run_rk4 (src/rk4.c:78-102) as
// device bodies over a work-group index range, shared by the per-context
// kernels and the grouped launches of several ensembles, and the same
// arithmetic as an epilogue of the kernels that produce a slope.
#pragma once

#include <hip/hip_runtime.h>

#include "mem.h"
#include "reduce.h"

namespace ihg {

// one element of a stage: y + dt k / 2 (half) or y + dt k
__device__ __forceinline__ double rk4_stage_val(double y, double k, double dt,
                                                int half)
{
    return half ? y + dt * k / 2.0 : y + dt * k;
}

// one element of the update: y + dt (k1 + 2 k2 + 2 k3 + k4) / 6
__device__ __forceinline__ double rk4_final_val(double y, double k1, double k2,
                                                double k3, double k4, double dt)
{
    return y + dt * (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0;
}

// RK4 stage: out = y + dt * k / 2 (half) or y + dt * k, over the pairs of
// [0, nlive) and the element `tail` (the scale factor; -1: none); work-group
// bx of gx
template <bool NT>
__device__ __forceinline__ void rk4_stage_body(
    unsigned bx, unsigned gx, double *__restrict__ out,
    const double *__restrict__ y, const double *__restrict__ k, double dt,
    int half, long nlive, long tail)
{
    const long npairs = nlive / 2;
    const long stride = (long)gx * RED_WG;
    const double2 *y2 = reinterpret_cast<const double2 *>(y);
    const double2 *k2 = reinterpret_cast<const double2 *>(k);
    double2 *o2 = reinterpret_cast<double2 *>(out);
    for (long i = (long)bx * RED_WG + threadIdx.x; i < npairs; i += stride) {
        const double2 yy = ld2<NT>(y2 + i), kk = ld2<NT>(k2 + i);
        st2<NT>(o2 + i, make_double2(rk4_stage_val(yy.x, kk.x, dt, half),
                                     rk4_stage_val(yy.y, kk.y, dt, half)));
    }
    if (tail >= 0 && bx == 0 && threadIdx.x == 0)
        out[tail] = rk4_stage_val(y[tail], k[tail], dt, half);
}

// RK4 update: y += dt * (k1 + 2 k2 + 2 k3 + k4) / 6
template <bool NT>
__device__ __forceinline__ void rk4_final_body(
    unsigned bx, unsigned gx, double *__restrict__ y,
    const double *__restrict__ k1, const double *__restrict__ k2,
    const double *__restrict__ k3, const double *__restrict__ k4, double dt,
    long nlive, long tail)
{
    const long npairs = nlive / 2;
    const long stride = (long)gx * RED_WG;
    double2 *y2 = reinterpret_cast<double2 *>(y);
    for (long i = (long)bx * RED_WG + threadIdx.x; i < npairs; i += stride) {
        const double2 a = ld2<NT>(reinterpret_cast<const double2 *>(k1) + i);
        const double2 b = ld2<NT>(reinterpret_cast<const double2 *>(k2) + i);
        const double2 c = ld2<NT>(reinterpret_cast<const double2 *>(k3) + i);
        const double2 d = ld2<NT>(reinterpret_cast<const double2 *>(k4) + i);
        const double2 yy = ld2<NT>(y2 + i);
        st2<NT>(y2 + i, make_double2(rk4_final_val(yy.x, a.x, b.x, c.x, d.x, dt),
                                     rk4_final_val(yy.y, a.y, b.y, c.y, d.y, dt)));
    }
    if (tail >= 0 && bx == 0 && threadIdx.x == 0)
        y[tail] = rk4_final_val(y[tail], k1[tail], k2[tail], k3[tail], k4[tail], dt);
}

// ---------------------------------------------------------------------------
// The RK4 step folded into the kernels that write a slope k (the update of
// phi, psi and their rates, the TT pass of the h blocks, the means of a):
// besides k they write, for the same elements,
//   RK4_STAGE: out = y + dt k / 2 (half) or y + dt k -- the next RHS input;
//   RK4_FINAL: y = y + dt (k1 + 2 k2 + 2 k3 + k) / 6 in place, and not k
//              (the step's last slope is needed nowhere else).
// The values are those of rk4_stage_body / rk4_final_body bit for bit: the
// slope enters the same expressions, made opaque first so that the compiler
// cannot fuse its last operation into them (as it cannot fuse a load).
// Pointers are those of the realization the kernel works on.
// ---------------------------------------------------------------------------
enum { RK4_NONE = 0, RK4_STAGE = 1, RK4_FINAL = 2 };

struct Rk4Ep {
    int mode, half;
    double dt;
    const double *y;
    double *out;
    const double *k1, *k2, *k3;

    __device__ __forceinline__ Rk4Ep shifted(long off) const
    {
        Rk4Ep e = *this;
        e.y += off;
        e.out += off;
        if (mode == RK4_FINAL) {
            e.k1 += off;
            e.k2 += off;
            e.k3 += off;
        }
        return e;
    }
};

__device__ __forceinline__ double opaque(double v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// the pair of elements 2p, 2p + 1 whose slope is k; kdst: where k goes
template <bool NT>
__device__ __forceinline__ void rk4_ep_pair(const Rk4Ep &ep, long p, double2 k,
                                            double2 *kdst)
{
    k = make_double2(opaque(k.x), opaque(k.y));
    if (ep.mode == RK4_FINAL) {
        double2 *y2 = reinterpret_cast<double2 *>(ep.out); // y, in place
        const double2 yy = ld2<NT>(y2 + p);
        const double2 a = ld2<NT>(reinterpret_cast<const double2 *>(ep.k1) + p);
        const double2 b = ld2<NT>(reinterpret_cast<const double2 *>(ep.k2) + p);
        const double2 c = ld2<NT>(reinterpret_cast<const double2 *>(ep.k3) + p);
        st2<NT>(y2 + p, make_double2(rk4_final_val(yy.x, a.x, b.x, c.x, k.x, ep.dt),
                                     rk4_final_val(yy.y, a.y, b.y, c.y, k.y, ep.dt)));
    } else {
        st2<NT>(kdst, k);
        const double2 yy = ld2<NT>(reinterpret_cast<const double2 *>(ep.y) + p);
        st2<NT>(reinterpret_cast<double2 *>(ep.out) + p,
                make_double2(rk4_stage_val(yy.x, k.x, ep.dt, ep.half),
                             rk4_stage_val(yy.y, k.y, ep.dt, ep.half)));
    }
}

// one element i whose slope is k
__device__ __forceinline__ void rk4_ep_one(const Rk4Ep &ep, long i, double k,
                                           double *kdst)
{
    k = opaque(k);
    if (ep.mode == RK4_FINAL) {
        ep.out[i] = rk4_final_val(ep.out[i], ep.k1[i], ep.k2[i], ep.k3[i], k, ep.dt);
    } else {
        *kdst = k;
        ep.out[i] = rk4_stage_val(ep.y[i], k, ep.dt, ep.half);
    }
}

} // namespace ihg
