// Device helper of the split-K reduce kernels: the sum of S FP32 partial
// matrices at one position, in slice order 0, 1, ..., S-1 (deterministic).
//
// All S loads are issued before the first add, so a thread waits for one memory
// round trip instead of S dependent ones. The partials were written by another
// kernel, usually from another XCD, so each load misses this XCD's L2; with
// sequential loads the reduce of a 16-way split took 5-6 us even for a 64 x 512
// output. The adds keep the order of a plain loop, so results are bit-identical
// to it.
#pragma once

#include <hip/hip_runtime.h>

#include "gemm_tile.hpp"
#include "plan.hpp"

namespace mmsk {

template <int S>
__device__ __forceinline__ f32x4 sum_splits4_fixed(const float* __restrict__ P, long long stride) {
    f32x4 v[S];
#pragma unroll
    for (int k = 0; k < S; ++k) v[k] = *reinterpret_cast<const f32x4u*>(P + k * stride);
    f32x4 s = v[0];
#pragma unroll
    for (int k = 1; k < S; ++k) s += v[k];
    return s;
}

// Dispatch of a wave-uniform split count onto the unrolled bodies S .. kMaxSplit.
template <int S>
__device__ __forceinline__ f32x4 sum_splits4_from(const float* __restrict__ P, long long stride, int splits) {
    if constexpr (S >= kMaxSplit) {
        return sum_splits4_fixed<kMaxSplit>(P, stride);
    } else {
        if (splits == S) return sum_splits4_fixed<S>(P, stride);
        return sum_splits4_from<S + 1>(P, stride, splits);
    }
}

// Four consecutive values (P[0..3] of every slice, dword-aligned) summed over
// `splits` slices, 1 <= splits <= kMaxSplit. The split count is uniform across
// the wave.
__device__ __forceinline__ f32x4 sum_splits4(const float* __restrict__ P, long long stride, int splits) {
    return sum_splits4_from<1>(P, stride, splits);
}

}  // namespace mmsk
