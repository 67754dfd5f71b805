// Shared helpers of the MI355X oa-fdtd kernels: storage conversions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oafd {

constexpr int MAX_POLES = 16;

// Field storage type S, arithmetic type R. 16-bit storage converts through float.
template <typename S, typename R, typename I = long long>
__device__ __forceinline__ R ldS(const void* base, I i) {
    if constexpr (__is_same(S, __bf16) || __is_same(S, _Float16))
        return (R)(float)(static_cast<const S*>(base)[i]);
    else
        return (R)(static_cast<const S*>(base)[i]);
}
template <typename S, typename R, typename I = long long>
__device__ __forceinline__ void stS(void* base, I i, R v) {
    if constexpr (__is_same(S, __bf16) || __is_same(S, _Float16))
        static_cast<S*>(base)[i] = (S)(float)v;
    else
        static_cast<S*>(base)[i] = (S)v;
}

}  // namespace oafd
