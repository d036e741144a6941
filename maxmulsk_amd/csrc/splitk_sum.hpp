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

static_assert(kMaxSplit == 16, "sum_splits4 dispatches split counts 1..16");

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

// Four consecutive values (P[0..3] of every slice, dword-aligned) summed over
// `splits` slices, 1 <= splits <= 16. The switch is uniform across the wave.
__device__ __forceinline__ f32x4 sum_splits4(const float* __restrict__ P, long long stride, int splits) {
    switch (splits) {
        case 1: return sum_splits4_fixed<1>(P, stride);
        case 2: return sum_splits4_fixed<2>(P, stride);
        case 3: return sum_splits4_fixed<3>(P, stride);
        case 4: return sum_splits4_fixed<4>(P, stride);
        case 5: return sum_splits4_fixed<5>(P, stride);
        case 6: return sum_splits4_fixed<6>(P, stride);
        case 7: return sum_splits4_fixed<7>(P, stride);
        case 8: return sum_splits4_fixed<8>(P, stride);
        case 9: return sum_splits4_fixed<9>(P, stride);
        case 10: return sum_splits4_fixed<10>(P, stride);
        case 11: return sum_splits4_fixed<11>(P, stride);
        case 12: return sum_splits4_fixed<12>(P, stride);
        case 13: return sum_splits4_fixed<13>(P, stride);
        case 14: return sum_splits4_fixed<14>(P, stride);
        case 15: return sum_splits4_fixed<15>(P, stride);
        default: return sum_splits4_fixed<16>(P, stride);
    }
}

}  // namespace mmsk
