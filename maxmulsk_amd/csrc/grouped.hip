// Grouped persistent GEMM: one launch runs every tile of every problem of a
// batch. Workgroups claim work items (problem, k-slice, tile) from an atomic
// counter until the flattened list is exhausted, so uneven problems balance
// themselves; the host orders the list most-expensive-first. One kernel per
// mode serves every tile geometry (switch on the problem's geometry), with LDS
// sized for the largest.
//
// Batching counterpart in the reference: run_batch over struct Gemm
// (sme/v4/sme-1x4-v4c-onestream.cpp:683-694), which runs the problems one after
// another inside one streaming region.
#include <hip/hip_runtime.h>

#include "grouped_body.hpp"
#include "kernels.hpp"
#include "splitk_sum.hpp"

namespace mmsk {

// Split modes (the plain kernel is grouped_plain_kernel, grouped_plain.hip).
template <int MODE>
__global__ __launch_bounds__(grouped_threads(MODE)) void grouped_kernel(GroupLaunch g) {
    static_assert(MODE != 0, "plain mode runs grouped_plain_kernel");
    MMSK_GROUPED_BODY(MODE)
}

__global__ __launch_bounds__(256) void grouped_reduce_kernel(const GroupProblem* __restrict__ problems,
                                                             const int* __restrict__ block_problem) {
    const GroupProblem& p = problems[block_problem[blockIdx.x]];
    const long long mn = p.M * p.N;
    const long long base = (long long)(blockIdx.x - p.reduce_begin) * 1024 + threadIdx.x * 4;
    if (base + 3 < mn) {
        *reinterpret_cast<f32x4u*>(p.C + base) = sum_splits4(p.P + base, mn, p.splits);
    } else {
        for (long long j = base; j < base + 4 && j < mn; ++j) {
            float s = p.P[j];
            for (int k = 1; k < p.splits; ++k) s += p.P[k * mn + j];
            p.C[j] = s;
        }
    }
}

hipError_t launch_grouped(int mode, const GroupLaunch& g, int grid, hipStream_t stream) {
    if (grid <= 0) return hipErrorInvalidValue;
    switch (mode) {
        case 0: return launch_grouped_plain(g, grid, stream);
        case 1: hipLaunchKernelGGL(grouped_kernel<1>, dim3(grid), dim3(grouped_threads(1)), 0, stream, g); break;
        case 2: hipLaunchKernelGGL(grouped_kernel<2>, dim3(grid), dim3(grouped_threads(2)), 0, stream, g); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int grouped_blocks_per_cu(int mode) {
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (mode) {
        case 0: return grouped_plain_blocks_per_cu();
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_kernel<1>, grouped_threads(1), 0); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_kernel<2>, grouped_threads(2), 0); break;
        default: break;
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}

hipError_t launch_grouped_reduce(const GroupProblem* problems, const int* block_problem,
                                 int total_blocks, hipStream_t stream) {
    if (total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(grouped_reduce_kernel, dim3(total_blocks), dim3(256), 0, stream, problems,
                       block_problem);
    return hipGetLastError();
}

}  // namespace mmsk
