// Grouped persistent kernel, plain mode (see grouped.hip for the launch
// structure and grouped_body.hpp for the shared work loop).
#include <hip/hip_runtime.h>

#include "grouped_body.hpp"
#include "kernels.hpp"

namespace mmsk {

// Register budget of two waves per SIMD, with the 16-deep k-blocks of
// grouped_bk: two 4-wave workgroups share a CU.
__global__ __launch_bounds__(grouped_threads(0)) __attribute__((amdgpu_waves_per_eu(2))) void grouped_plain_kernel(
    GroupLaunch g) {
    MMSK_GROUPED_BODY(0)
}

hipError_t launch_grouped_plain(const GroupLaunch& g, int grid, hipStream_t stream) {
    if (grid <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(grouped_plain_kernel, dim3(grid), dim3(grouped_threads(0)), 0, stream, g);
    return hipGetLastError();
}

int grouped_plain_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_plain_kernel, grouped_threads(0), 0) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}

}  // namespace mmsk
