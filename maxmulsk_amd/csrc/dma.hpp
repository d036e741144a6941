// Global -> LDS DMA helpers shared by the kernels that stage operands with
// global_load_lds_dwordx4 into a ring of LDS buffers.
#pragma once

#include <hip/hip_runtime.h>

namespace mmsk {

// s_waitcnt with only the vector-memory counter constrained (gfx9 encoding:
// vmcnt[3:0], expcnt[6:4], lgkmcnt[11:8], vmcnt[5:4] in bits 15:14).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

template <int D>
__device__ __forceinline__ void wait_stages_behind(int stages) {
    // Outstanding DMA allowed: `stages` later stages of D instructions each.
    switch (stages) {
        case 0: wait_vmcnt<0>(); break;
        case 1: wait_vmcnt<D>(); break;
        case 2: wait_vmcnt<2 * D>(); break;
        default: wait_vmcnt<3 * D>(); break;
    }
}

// global_load_lds_dwordx4: lane l's 16 bytes at gsrc land at lds + 16 * l.
__device__ __forceinline__ void dma16(const char* gsrc, char* lds) {
    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)gsrc,
                                     (void __attribute__((address_space(3)))*)lds, 16, 0, 0);
}

// Workgroup barrier without the release/acquire fence of __syncthreads: the
// fence would make the compiler drain every outstanding DMA (vmcnt(0)) at each
// barrier and collapse the ring to one block of lookahead. Each wave waits for
// its own DMA of the stage it is about to read (wait_stages_behind) before the
// barrier, and its LDS reads of the previous stage have completed (their
// MFMAs consumed them), so after the barrier the stage is complete for every
// wave and the freed buffer may be refilled. The empty asm statements keep the
// compiler from moving LDS accesses across the barrier.
__device__ __forceinline__ void stage_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt left at their maxima
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

}  // namespace mmsk
