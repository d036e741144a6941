// Device code shared by the grouped persistent kernels (grouped.hip: split
// modes, reduce and launchers; grouped_plain.hip: plain mode). The plain
// kernel lives in its own translation unit so that its register budget and
// k-block depth leave the split-mode kernels' code generation as it is.
#pragma once

#include <hip/hip_runtime.h>

#include "gemm_tile.hpp"
#include "kernels.hpp"

namespace mmsk {

// One register set of staging per geometry: the kernel's register allocation
// is the maximum over its geometries, and two sets make the plain 128 x 256
// case spill. Split modes: eight waves (2 x 4) per tile, 512 threads.
//
// Plain mode: k-blocks of 16 (LDS 48 KiB instead of 96) and a register budget
// of two waves per SIMD, so two 4-wave workgroups share a CU where one did.
// The budget spills 276 bytes per lane, and the 256-problem batch still went
// from 30.1 to 26.3 ms (101 -> 116 TFLOP/s). The split modes keep 32-deep
// blocks and their own allocation: under the same budget split3 was 4% slower.
// The k-block depth changes no product and no summation order.
constexpr int grouped_bk(int mode, int cfg) { return mode == 0 ? 16 : kTileCfgs[cfg].bk[mode]; }

template <int MODE, int CFG>
using CfgTile = Tile<MODE, kTileCfgs[CFG].bm, kTileCfgs[CFG].bn, grouped_bk(MODE, CFG),
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

// Workgroups claim items from *counter until the list is exhausted, so every
// wave reaches the exit.
#define MMSK_GROUPED_BODY(MODE)                                                              \
    __shared__ __attribute__((aligned(16))) char smem[max_smem<MODE>()];                     \
    __shared__ int s_item;                                                                   \
    for (;;) {                                                                               \
        if (threadIdx.x == 0) s_item = (int)atomicAdd(g.counter, 1u);                        \
        __syncthreads();                                                                     \
        const int item = s_item;                                                             \
        __syncthreads(); /* every wave has read s_item before it is overwritten */           \
        if (item >= g.total_items) break;                                                    \
        const GroupProblem& p = g.problems[g.item_problem[item]];                            \
        const int local = item - p.item_begin;                                               \
        switch (p.cfg) {                                                                     \
            case 0:                                                                          \
                if constexpr (grouped_min_cfg(MODE) == 0) run_item<MODE, 0>(p, local, smem); \
                break;                                                                       \
            case 1: run_item<MODE, 1>(p, local, smem); break;                                \
            case 2: run_item<MODE, 2>(p, local, smem); break;                                \
            case 3: run_item<MODE, 3>(p, local, smem); break;                                \
            default: run_item<MODE, 4>(p, local, smem); break;                               \
        }                                                                                    \
    }

}  // namespace mmsk
