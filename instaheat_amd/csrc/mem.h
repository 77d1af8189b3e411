// mem.h -- 16-byte loads and stores with an optional streaming hint.
//
// Arrays much larger than the caches (the 1 GB state vectors at 256^3) are
// read and written once per kernel: nontemporal accesses keep them from
// evicting each other in L2 and the Infinity Cache, which raises the rate
// of the multi-stream stage kernels by 5-10 % on gfx950.  Small arrays (an
// ensemble of 32^3 states fits in the 256 MB Infinity Cache) keep the
// default policy, so that the next kernel can still hit in the cache.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace ihg {

typedef double vd2 __attribute__((ext_vector_type(2)));

// arrays at least this large take the streaming hint
constexpr double NT_BYTES = 256.0 * 1024.0 * 1024.0;

template <bool NT>
__device__ __forceinline__ double2 ld2(const double2 *p)
{
    if constexpr (NT) {
        const vd2 v = __builtin_nontemporal_load(reinterpret_cast<const vd2 *>(p));
        return make_double2(v.x, v.y);
    } else {
        return *p;
    }
}

template <bool NT>
__device__ __forceinline__ void st2(double2 *p, double2 v)
{
    if constexpr (NT) {
        const vd2 w = {v.x, v.y};
        __builtin_nontemporal_store(w, reinterpret_cast<vd2 *>(p));
    } else {
        *p = v;
    }
}

typedef float vf2 __attribute__((ext_vector_type(2)));

// Loads and stores of the FFT passes: streaming when FFT_NT (their arrays,
// 135 MB per field at 256^3, are read and written once per pass)
constexpr bool FFT_NT = false;

template <class T>
__device__ __forceinline__ T ldn(const T *p)
{
    if constexpr (!FFT_NT) {
        return *p;
    } else if constexpr (std::is_same<T, double2>::value) {
        return ld2<true>(p);
    } else if constexpr (std::is_same<T, float2>::value) {
        const vf2 v = __builtin_nontemporal_load(reinterpret_cast<const vf2 *>(p));
        return make_float2(v.x, v.y);
    } else {
        return __builtin_nontemporal_load(p);
    }
}

template <class T>
__device__ __forceinline__ void stn(T *p, T v)
{
    if constexpr (!FFT_NT) {
        *p = v;
    } else if constexpr (std::is_same<T, double2>::value) {
        st2<true>(p, v);
    } else if constexpr (std::is_same<T, float2>::value) {
        const vf2 w = {v.x, v.y};
        __builtin_nontemporal_store(w, reinterpret_cast<vf2 *>(p));
    } else {
        __builtin_nontemporal_store(v, p);
    }
}

} // namespace ihg
