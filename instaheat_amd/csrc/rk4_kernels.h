// rk4_kernels.h -- the stage arithmetic of run_rk4 (src/rk4.c:78-102) as
// device bodies over a work-group index range, shared by the per-context
// kernels and the grouped launches of several ensembles.
#pragma once

#include <hip/hip_runtime.h>

#include "mem.h"
#include "reduce.h"

namespace ihg {

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
        if (half)
            st2<NT>(o2 + i, make_double2(yy.x + dt * kk.x / 2.0,
                                         yy.y + dt * kk.y / 2.0));
        else
            st2<NT>(o2 + i, make_double2(yy.x + dt * kk.x, yy.y + dt * kk.y));
    }
    if (tail >= 0 && bx == 0 && threadIdx.x == 0) {
        const long i = tail;
        out[i] = half ? y[i] + dt * k[i] / 2.0 : y[i] + dt * k[i];
    }
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
        double2 yy = ld2<NT>(y2 + i);
        yy.x += dt * (a.x + 2.0 * b.x + 2.0 * c.x + d.x) / 6.0;
        yy.y += dt * (a.y + 2.0 * b.y + 2.0 * c.y + d.y) / 6.0;
        st2<NT>(y2 + i, yy);
    }
    if (tail >= 0 && bx == 0 && threadIdx.x == 0) {
        const long i = tail;
        y[i] += dt * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]) / 6.0;
    }
}

} // namespace ihg
