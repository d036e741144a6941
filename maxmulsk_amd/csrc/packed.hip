// Packed GEMM path for compute-bound shapes (the 256 x 256 geometry).
//
// 1. pack: A (rows = M) and B (rows = N, read column-wise) are rewritten once
//    into MFMA fragments, split into their BF16 terms (or copied as FP32 in
//    plain mode), zero-padded to whole tiles and k-blocks. Fragment (term,
//    rblk, kg) is 1 KiB: lane l holds row rblk*32 + l%32 and the 16-byte chunk
//    of k values kg*KGK + (l/32)*KGK/2 .. +KGK/2 (KGK = 8 FP32 or 16 BF16).
//    Counterpart of pack_A / pack_B in the reference
//    (sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:58-175), with the BF16 split
//    folded in.
// 2. gemm: per k-block every wave moves whole fragments global -> LDS with
//    global_load_lds_dwordx4 (one 1 KiB fragment per instruction, no VGPR
//    staging) into a STAGES-deep ring, so loads run STAGES-1 blocks ahead of the
//    MFMAs; fragments are read back with one ds_read_b128 per lane at
//    consecutive addresses (conflict-free without a swizzle).
#include <hip/hip_runtime.h>

#include "dma.hpp"
#include "gemm_tile.hpp"
#include "kernels.hpp"

namespace mmsk {

// ---- pack ------------------------------------------------------------------------

// One workgroup per 32-row x 256-k super-tile: the FP32 tile is read from
// global memory along the operand's contiguous direction (k for A, rows for B)
// into LDS, then written out as the 256/KGK fragments (x terms) it covers,
// which are contiguous in the packed layout. Rows >= R and k >= K are zero.
// Element (row, k) is at X[row * ld_row + k * ld_k]; k_major: ld_row == 1.
constexpr int kPackKT = 256;        // k values per super-tile
constexpr int kPackPitch = kPackKT + 4;  // LDS row pitch (floats)

// One launch may pack two operands of the same K (A and B of one product):
// workgroups [0, a.rblocks) serve a, the rest serve b.
template <int MODE>
__global__ __launch_bounds__(256) void pack_kernel(PackJob a, PackJob b, long long K, int kgroups) {
    constexpr int LU = mode_lu(MODE);  // k values per 16-byte chunk
    constexpr int KGK = 2 * LU;        // k values per fragment
    constexpr int TERMS = mode_terms(MODE);
    __shared__ __attribute__((aligned(16))) float tile[32 * kPackPitch];
    const bool second = (int)blockIdx.x >= a.rblocks;
    const float* __restrict__ X = second ? b.X : a.X;
    const long long R = second ? b.R : a.R;
    const long long ld_row = second ? b.ld_row : a.ld_row;
    const long long ld_k = second ? b.ld_k : a.ld_k;
    const int k_major = ld_row == 1 && ld_k != 1;
    char* __restrict__ dst = second ? b.dst : a.dst;
    const int rblocks = second ? b.rblocks : a.rblocks;
    const int rb = second ? (int)blockIdx.x - a.rblocks : (int)blockIdx.x;
    const long long r0 = (long long)rb * 32;
    const long long k0 = (long long)blockIdx.y * kPackKT;
    const int tid = threadIdx.x;
    // A super-tile wholly inside the operand and contiguous along its read
    // direction issues all eight loads before its first LDS write, so they
    // are in flight together; the guarded loads below cannot be, because the
    // compiler waits for each one inside its branch.
    const bool whole = r0 + 32 <= R && k0 + kPackKT <= K && (k_major ? ld_row == 1 : ld_k == 1);
    if (whole) {
        f32x4 v[8];
        if (!k_major) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int idx = tid + 256 * i;
                v[i] = *reinterpret_cast<const f32x4u*>(X + (r0 + (idx >> 6)) * ld_row + k0 + 4 * (idx & 63));
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int idx = tid + 256 * i;
                *reinterpret_cast<f32x4*>(&tile[(idx >> 6) * kPackPitch + 4 * (idx & 63)]) = v[i];
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int idx = tid + 256 * i;
                v[i] = *reinterpret_cast<const f32x4u*>(X + (k0 + (idx >> 3)) * ld_k + r0 + 4 * (idx & 7));
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int idx = tid + 256 * i;
#pragma unroll
                for (int e = 0; e < 4; ++e) tile[(4 * (idx & 7) + e) * kPackPitch + (idx >> 3)] = v[i][e];
            }
        }
    } else if (!k_major) {
        // rows of A: 256 contiguous k per row, 64 float4 per row
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int idx = tid + 256 * i;
            const int row = idx >> 6, c4 = idx & 63;
            const long long gr = r0 + row, gk = k0 + 4 * c4;
            f32x4 v = {0.f, 0.f, 0.f, 0.f};
            if (gr < R) {
                const float* src = X + gr * ld_row + gk * ld_k;
                if (gk + 3 < K && ld_k == 1) {
                    v = *reinterpret_cast<const f32x4u*>(src);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (gk + e < K) v[e] = src[e * ld_k];
                }
            }
            *reinterpret_cast<f32x4*>(&tile[row * kPackPitch + 4 * c4]) = v;
        }
    } else {
        // columns of B: for each k, 32 contiguous rows (8 float4)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int idx = tid + 256 * i;
            const int kk = idx >> 3, r4 = idx & 7;
            const long long gk = k0 + kk, gr = r0 + 4 * r4;
            f32x4 v = {0.f, 0.f, 0.f, 0.f};
            if (gk < K) {
                const float* src = X + gk * ld_k + gr * ld_row;
                if (gr + 3 < R && ld_row == 1) {
                    v = *reinterpret_cast<const f32x4u*>(src);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (gr + e < R) v[e] = src[e * ld_row];
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) tile[(4 * r4 + e) * kPackPitch + kk] = v[e];
        }
    }
    __syncthreads();
    // Fragments kg0 .. kg0 + nf - 1 of row block rb, 64 lane-chunks each.
    const int kg0 = (int)(k0 / KGK);
    int nf = kPackKT / KGK;
    if (kg0 + nf > kgroups) nf = kgroups - kg0;
    const long long plane = (long long)rblocks * kgroups * 1024;  // bytes per term
    char* base = dst + ((long long)rb * kgroups + kg0) * 1024;
    for (int o = tid; o < nf * 64; o += 256) {
        const int f = o >> 6, lane = o & 63;
        const float* src = &tile[(lane & 31) * kPackPitch + f * KGK + (lane >> 5) * LU];
        float v[LU];
#pragma unroll
        for (int q = 0; q < LU / 4; ++q) {
            const f32x4 w = *reinterpret_cast<const f32x4*>(src + 4 * q);
            v[4 * q + 0] = w[0];
            v[4 * q + 1] = w[1];
            v[4 * q + 2] = w[2];
            v[4 * q + 3] = w[3];
        }
        char* out = base + (long long)o * 16;
        if constexpr (MODE == 0) {
            const f32x4 w = {v[0], v[1], v[2], v[3]};
            *reinterpret_cast<f32x4*>(out) = w;
        } else {
            uint32_t w[TERMS][4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t t[TERMS];
                split_pair<TERMS>(v[2 * q], v[2 * q + 1], t);
#pragma unroll
                for (int s = 0; s < TERMS; ++s) w[s][q] = t[s];
            }
#pragma unroll
            for (int s = 0; s < TERMS; ++s) {
                const u32x4 u = {w[s][0], w[s][1], w[s][2], w[s][3]};
                *reinterpret_cast<u32x4*>(out + s * plane) = u;
            }
        }
    }
}

// ---- packed GEMM tile ---------------------------------------------------------------

template <int MODE, int BM, int BN, int BK, int WGM, int WGN, int STAGES>
struct PackedTile {
    static constexpr int TERMS = mode_terms(MODE);
    static constexpr int KGK = 2 * mode_lu(MODE);        // k per fragment
    static constexpr int KGB = BK / KGK;                 // fragments deep per k-block
    static constexpr int RA = BM / 32, RB = BN / 32;
    static constexpr int NFA = TERMS * RA * KGB, NFB = TERMS * RB * KGB;
    static constexpr int NW = WGM * WGN, NT = 64 * NW;
    static constexpr int NF = NFA + NFB;
    static constexpr int D = NF / NW;                    // DMA instructions per wave per stage
    static constexpr int STAGE = NF * 1024;
    static constexpr int SMEM = STAGES * STAGE;
    static constexpr int WTM = BM / WGM, WTN = BN / WGN, TM = WTM / 32, TN = WTN / 32;
    static_assert(NF % NW == 0, "fragments split evenly over the waves");
    static_assert(SMEM <= 160 * 1024, "LDS budget of one CU");
    static_assert(STAGES >= 2 && STAGES <= 4, "pipeline depth");
    static_assert(BK % KGK == 0, "whole fragments per k-block");


    f32x16 acc[TM][TN];

    // Issue the DMA of k-block kb into stage buffer `st`.
    __device__ __forceinline__ void issue(const char* __restrict__ PA, const char* __restrict__ PB,
                                          long long planeA, long long planeB, int kgtA, int kgtB, int rb0A,
                                          int rb0B, long long kb, char* st) const {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const int f = wave + NW * i;
            const char* src;
            if (f < NFA) {
                const int t = f / (RA * KGB), rb = (f / KGB) % RA, kg = f % KGB;
                src = PA + t * planeA + ((long long)(rb0A + rb) * kgtA + kb * KGB + kg) * 1024;
            } else {
                const int g = f - NFA;
                const int t = g / (RB * KGB), rb = (g / KGB) % RB, kg = g % KGB;
                src = PB + t * planeB + ((long long)(rb0B + rb) * kgtB + kb * KGB + kg) * 1024;
            }
            dma16(src + lane * 16, st + f * 1024);
        }
    }

    // Per k-group: plain mode reads B fragments once and A fragments one 32-row
    // block at a time (the other wave on the SIMD covers the LDS latency); the
    // split modes read every fragment of the k-group first.
    __device__ __forceinline__ void compute(const char* st) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int wm = wave / WGN, wn = wave % WGN;
        const char* sa = st + lane * 16;
        const char* sb = st + NFA * 1024 + lane * 16;
#pragma unroll
        for (int kg = 0; kg < KGB; ++kg) {
            if constexpr (MODE == 0) {
                f32x4 b[TN];
#pragma unroll
                for (int tj = 0; tj < TN; ++tj)
                    b[tj] = *reinterpret_cast<const f32x4*>(sb + ((wn * TN + tj) * KGB + kg) * 1024);
#pragma unroll
                for (int ti = 0; ti < TM; ++ti) {
                    const f32x4 a = *reinterpret_cast<const f32x4*>(sa + ((wm * TM + ti) * KGB + kg) * 1024);
#pragma unroll
                    for (int s = 0; s < 4; ++s)
#pragma unroll
                        for (int tj = 0; tj < TN; ++tj)
                            acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(b[tj][s], a[s], acc[ti][tj], 0, 0, 0);
                }
            } else {
                bf16x8 b[TERMS][TN];
#pragma unroll
                for (int s = 0; s < TERMS; ++s)
#pragma unroll
                    for (int tj = 0; tj < TN; ++tj)
                        b[s][tj] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(
                                                                  sb + ((s * RB + wn * TN + tj) * KGB + kg) * 1024));
                // All A fragments of the k-group, then for each B column block the products in
                // order (smallest first per accumulator) over the TM row blocks: consecutive
                // MFMAs share their B operand, which measured 1.5-2% faster at 4096^3
                // (split3 and split6) than sharing the A operand two at a time.
                bf16x8 a[TERMS][TM];
#pragma unroll
                for (int ti = 0; ti < TM; ++ti)
#pragma unroll
                    for (int s = 0; s < TERMS; ++s)
                        a[s][ti] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(
                                                              sa + ((s * RA + wm * TM + ti) * KGB + kg) * 1024));
                constexpr int NP = MODE == 1 ? 3 : 6;
#pragma unroll
                for (int tj = 0; tj < TN; ++tj)
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        const int ta = product_term(MODE, p, 0);
                        const int tb = product_term(MODE, p, 1);
#pragma unroll
                        for (int ti = 0; ti < TM; ++ti)
                            acc[ti][tj] =
                                __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[tb][tj], a[ta][ti], acc[ti][tj], 0, 0, 0);
                    }
            }
        }
    }

    __device__ __forceinline__ void run(const char* __restrict__ PA, const char* __restrict__ PB, long long planeA,
                                        long long planeB, int kgtA, int kgtB, int tm, int tn, long long kb0,
                                        long long nkb, float* __restrict__ C, long long M, long long N, long long ldc,
                                        char* smem) {
#pragma unroll
        for (int ti = 0; ti < TM; ++ti)
#pragma unroll
            for (int tj = 0; tj < TN; ++tj)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[ti][tj][e] = 0.f;
        const int rb0A = tm * RA, rb0B = tn * RB;
        // Ring of STAGES buffers; block j lives in buffer j % STAGES and is
        // DMA'd STAGES-1 iterations before it is computed.
#pragma unroll
        for (int s = 0; s < STAGES - 1; ++s)
            if (s < nkb) issue(PA, PB, planeA, planeB, kgtA, kgtB, rb0A, rb0B, kb0 + s, smem + s * STAGE);
        // Iteration kb: wait for this wave's DMA of block kb (blocks issued
        // after it: up to kb+STAGES-2), barrier (block kb complete for every
        // wave; every wave is done reading block kb-1), refill buffer
        // (kb-1) % STAGES with block kb+STAGES-1, then read and compute block kb.
        for (long long kb = 0; kb < nkb; ++kb) {
            wait_stages_behind<D>((int)min(nkb - 1 - kb, (long long)STAGES - 2));
            stage_barrier();
            if (kb + STAGES - 1 < nkb)
                issue(PA, PB, planeA, planeB, kgtA, kgtB, rb0A, rb0B, kb0 + kb + STAGES - 1,
                      smem + ((kb + STAGES - 1) % STAGES) * STAGE);
            compute(smem + (kb % STAGES) * STAGE);
        }
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int wm = wave / WGN, wn = wave % WGN;
        const int r = lane & 31, h = lane >> 5;
        const long long m0 = (long long)tm * BM, n0 = (long long)tn * BN;
        // Epilogue through LDS, 32 rows of the wave tile at a time: the
        // accumulators (lane = row, four columns per register quad) are
        // written to a padded row-major block, then read back four whole
        // rows per instruction, so every global store covers complete
        // 128-byte lines of C instead of 32-byte pieces of 32 rows.
        constexpr int PITCH = WTN + 4;  // floats; 16-byte row shift keeps both patterns conflict-free
        static_assert(NW * 32 * PITCH * 4 <= SMEM, "epilogue blocks fit in the ring's LDS");
        float* blk = reinterpret_cast<float*>(smem) + wave * 32 * PITCH;
        constexpr int LPR = WTN / 4;  // lanes per row when reading
        constexpr int RPI = 64 / LPR; // rows per read instruction
        const int rr = lane / LPR, cc = (lane % LPR) * 4;
        __syncthreads();  // every wave has finished reading the ring
#pragma unroll
        for (int ti = 0; ti < TM; ++ti) {
#pragma unroll
            for (int tj = 0; tj < TN; ++tj)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const f32x4 v = {acc[ti][tj][4 * q + 0], acc[ti][tj][4 * q + 1], acc[ti][tj][4 * q + 2],
                                     acc[ti][tj][4 * q + 3]};
                    *reinterpret_cast<f32x4*>(blk + r * PITCH + tj * 32 + 8 * q + 4 * h) = v;
                }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 32 / RPI; ++i) {
                const int row = i * RPI + rr;
                const long long m = m0 + wm * WTM + ti * 32 + row;
                const long long n = n0 + wn * WTN + cc;
                const f32x4 v = *reinterpret_cast<const f32x4*>(blk + row * PITCH + cc);
                if (m < M) {
                    float* crow = C + m * ldc;
                    if (n + 3 < N) {
                        *reinterpret_cast<f32x4u*>(crow + n) = v;
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (n + e < N) crow[n + e] = v[e];
                    }
                }
            }
            __syncthreads();
        }
    }
};

template <int MODE>
struct PackedCfg;
// 256 x 256 tiles, eight waves of 128 x 64 (two waves per SIMD, so one
// wave's LDS reads and DMA issue overlap the other's MFMAs); k-block depth
// (packed_bk) and ring depth per mode.
template <> struct PackedCfg<0> { using T = PackedTile<0, 256, 256, 32, 2, 4, 2>; };
template <> struct PackedCfg<1> { using T = PackedTile<1, 256, 256, 32, 2, 4, 2>; };
template <> struct PackedCfg<2> { using T = PackedTile<2, 256, 256, 16, 2, 4, 3>; };
static_assert(PackedCfg<0>::T::KGB * PackedCfg<0>::T::KGK == packed_bk(0), "plain k-block depth");
static_assert(PackedCfg<1>::T::KGB * PackedCfg<1>::T::KGK == packed_bk(1), "split3 k-block depth");
static_assert(PackedCfg<2>::T::KGB * PackedCfg<2>::T::KGK == packed_bk(2), "split6 k-block depth");

template <int MODE>
__global__ __launch_bounds__(PackedCfg<MODE>::T::NT) void gemm_packed_kernel(PackedParams p) {
    using T = typename PackedCfg<MODE>::T;
    __shared__ __attribute__((aligned(1024))) char smem[T::SMEM];
    const int ntiles = p.tiles_m * p.tiles_n;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int split = bid / ntiles;
    int tm, tn;
    grouped_tile(bid - split * ntiles, p.tiles_m, p.tiles_n, tm, tn);
    const long long kb0 = (long long)split * p.kblocks_per_split;
    long long nkb = p.kblocks - kb0;
    if (nkb > p.kblocks_per_split) nkb = p.kblocks_per_split;
    T t;
    t.run(p.PA, p.PB, p.planeA, p.planeB, p.kgtA, p.kgtB, tm, tn, kb0, nkb, p.C + split * p.c_split_stride, p.M, p.N,
          p.ldc, smem);
}

hipError_t launch_pack2(int mode, const PackJob& a, const PackJob& b, long long K, int kgroups,
                        hipStream_t stream) {
    const long long kpad = (long long)kgroups * packed_kgk(mode);
    const long long blocks = (long long)a.rblocks + b.rblocks;
    if (blocks <= 0 || blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)blocks, (unsigned)((kpad + kPackKT - 1) / kPackKT));
    switch (mode) {
        case 0: hipLaunchKernelGGL(pack_kernel<0>, grid, dim3(256), 0, stream, a, b, K, kgroups); break;
        case 1: hipLaunchKernelGGL(pack_kernel<1>, grid, dim3(256), 0, stream, a, b, K, kgroups); break;
        case 2: hipLaunchKernelGGL(pack_kernel<2>, grid, dim3(256), 0, stream, a, b, K, kgroups); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pack(int mode, const float* X, long long R, long long K, long long ld_row, long long ld_k,
                       void* dst, int rblocks, int kgroups, hipStream_t stream) {
    const PackJob a{X, R, ld_row, ld_k, static_cast<char*>(dst), rblocks};
    const PackJob none{nullptr, 0, 1, 1, nullptr, 0};
    return launch_pack2(mode, a, none, K, kgroups, stream);
}

hipError_t launch_gemm_packed(int mode, const PackedParams& p, hipStream_t stream) {
    const long long grid = (long long)p.tiles_m * p.tiles_n * p.splits;
    if (grid <= 0 || grid > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (mode) {
        case 0: hipLaunchKernelGGL(gemm_packed_kernel<0>, dim3((unsigned)grid), dim3(PackedCfg<0>::T::NT), 0, stream, p); break;
        case 1: hipLaunchKernelGGL(gemm_packed_kernel<1>, dim3((unsigned)grid), dim3(PackedCfg<1>::T::NT), 0, stream, p); break;
        case 2: hipLaunchKernelGGL(gemm_packed_kernel<2>, dim3((unsigned)grid), dim3(PackedCfg<2>::T::NT), 0, stream, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mmsk
