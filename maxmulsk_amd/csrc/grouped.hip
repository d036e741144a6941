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

#include "gemm_tile.hpp"
#include "kernels.hpp"
#include "splitk_sum.hpp"

namespace mmsk {

// One register set of staging per geometry: the kernel's register allocation
// is the maximum over its geometries, and two sets make the plain 128 x 256
// case spill. Split modes: eight waves (2 x 4) per tile, 512 threads.
template <int MODE, int CFG>
using CfgTile = Tile<MODE, kTileCfgs[CFG].bm, kTileCfgs[CFG].bn, kTileCfgs[CFG].bk[MODE],
                     MODE == 0 ? kTileCfgs[CFG].wgm : 2, MODE == 0 ? kTileCfgs[CFG].wgn : 4, 1>;

// The persistent kernel serves geometries 0..kGroupedMaxCfg (up to 128 x 256):
// with the 256 x 256 geometry in the same kernel the register allocation of the
// combined kernel spills.
template <int MODE>
constexpr int max_smem() {
    int m = 0;
    if constexpr (grouped_min_cfg(MODE) == 0) m = CfgTile<MODE, 0>::SMEM;
    m = m > CfgTile<MODE, 1>::SMEM ? m : CfgTile<MODE, 1>::SMEM;
    m = m > CfgTile<MODE, 2>::SMEM ? m : CfgTile<MODE, 2>::SMEM;
    m = m > CfgTile<MODE, 3>::SMEM ? m : CfgTile<MODE, 3>::SMEM;
    m = m > CfgTile<MODE, 4>::SMEM ? m : CfgTile<MODE, 4>::SMEM;
    return m;
}
static_assert(kGroupedMaxCfg == 4, "switch below covers geometries 0..4");

template <int MODE, int CFG>
__device__ __forceinline__ void run_item(const GroupProblem& p, int local, char* smem) {
    constexpr TileCfg c = kTileCfgs[CFG];
    const int ntiles = p.tiles_m * p.tiles_n;
    const int split = local / ntiles;
    int tm, tn;
    grouped_tile(local - split * ntiles, p.tiles_m, p.tiles_n, tm, tn);
    const long long kbeg = split * p.k_len;
    const long long kend = min(p.K, kbeg + p.k_len);
    float* dst = p.splits > 1 ? p.P + split * p.M * p.N : p.C;
    CfgTile<MODE, CFG> t;
    t.run(p.A, p.B, dst, p.M, p.N, p.K, p.N, p.N, (long long)tm * c.bm, (long long)tn * c.bn, kbeg,
          kend, smem);
}

template <int MODE>
__global__ __launch_bounds__(grouped_threads(MODE)) void grouped_kernel(GroupLaunch g) {
    __shared__ __attribute__((aligned(16))) char smem[max_smem<MODE>()];
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(g.counter, 1u);
        __syncthreads();
        const int item = s_item;
        __syncthreads();  // every wave has read s_item before it is overwritten
        if (item >= g.total_items) break;
        const GroupProblem& p = g.problems[g.item_problem[item]];
        const int local = item - p.item_begin;
        switch (p.cfg) {
            case 0:
                if constexpr (grouped_min_cfg(MODE) == 0) run_item<MODE, 0>(p, local, smem);
                break;
            case 1: run_item<MODE, 1>(p, local, smem); break;
            case 2: run_item<MODE, 2>(p, local, smem); break;
            case 3: run_item<MODE, 3>(p, local, smem); break;
            default: run_item<MODE, 4>(p, local, smem); break;
        }
    }
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
        case 0: hipLaunchKernelGGL(grouped_kernel<0>, dim3(grid), dim3(grouped_threads(0)), 0, stream, g); break;
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
        case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_kernel<0>, grouped_threads(0), 0); break;
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
