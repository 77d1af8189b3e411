// reduce.h -- deterministic block reductions (wave64 shuffles, then LDS).
//
// Every reduction in the library is two-level: each block writes its partial
// to HBM in a fixed order and one block combines the partials in a fixed
// order, so results do not depend on scheduling.
#pragma once

#include <hip/hip_runtime.h>

namespace ihg {

constexpr int RED_WG = 256;

struct OpSum {
    __device__ static double id() { return 0.0; }
    __device__ static double op(double a, double b) { return a + b; }
};
struct OpMax {
    __device__ static double id() { return -__builtin_inf(); }
    __device__ static double op(double a, double b) { return a > b ? a : b; }
};
struct OpMin {
    __device__ static double id() { return __builtin_inf(); }
    __device__ static double op(double a, double b) { return a < b ? a : b; }
};

template <class Op>
__device__ __forceinline__ double wave_reduce(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = Op::op(v, __shfl_down(v, o, 64));
    return v;
}

// Reduce v over a block of `nthreads` (a multiple of 64, at most 1024).
// The result is valid in thread 0.  `sh` holds nthreads / 64 doubles.
// Must be called by all threads; synchronises.
template <class Op>
__device__ __forceinline__ double block_reduce(double v, double *sh,
                                               int nthreads)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_reduce<Op>(v);
    __syncthreads();
    if (lane == 0)
        sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int i = 1; i < nthreads / 64; ++i)
            r = Op::op(r, sh[i]);
        v = r;
    }
    return v;
}

} // namespace ihg
