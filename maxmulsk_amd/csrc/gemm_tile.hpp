// Device-side GEMM tile: one workgroup (4 or 8 waves) computes one BM x BN output
// tile over one K range. Shared by the one-tile-per-workgroup kernel and the
// grouped persistent kernel.
//
// Pipeline per BK-deep k-block (two LDS stages, one barrier per block):
//   MFMA over the LDS stage of block kb
//   registers of block kb+1 -> (split to BF16 terms) -> the other LDS stage
//   global -> registers for block kb+2 (in flight across the barrier)
//   barrier
// Edge tiles and the last partial k-block take the masked, element-wise load
// path (zero fill outside M, N and the K range); full tiles take dwordx4 loads.
//
// Counterparts in the reference: pack_B_streaming
// (sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:127-175) and the A panel transpose
// NB_PACK_A_KP_CHUNK (:58-117) become the register -> LDS stores below;
// micro_kernel_1x4 (:196-327) becomes the MFMA loop; svzero_za (:554) the
// zero-initialised accumulators; store_za / the masked edge copy (:342-405,
// :561-588) the epilogue.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lds_layout.hpp"

namespace mmsk {

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// Global-memory vectors that only promise dword alignment: rows of A, B and C
// start at any float when K or N is not a multiple of 4. Multi-dword global
// loads and stores need only dword alignment on gfx950.
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

// Two FP32 values -> two BF16 values (round to nearest even), packed with the
// first value in the low half (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t pack_bf16x2(float x0, float x1) {
    f32x2 v = {x0, x1};
    bf16x2 h = __builtin_convertvector(v, bf16x2);
    return __builtin_bit_cast(uint32_t, h);
}

// Exact FP32 values of a packed BF16 pair.
__device__ __forceinline__ void unpack_bf16x2(uint32_t w, float& f0, float& f1) {
    f0 = __uint_as_float(w << 16);
    f1 = __uint_as_float(w & 0xffff0000u);
}

// Splits two FP32 values into TERMS BF16 terms each: hi = RN(x), then each next
// term rounds the exact FP32 residual of the previous ones. x - hi (and the
// second residual) is exact in FP32 because hi agrees with x in its leading bits.
template <int TERMS>
__device__ __forceinline__ void split_pair(float x0, float x1, uint32_t (&out)[TERMS]) {
    const uint32_t h = pack_bf16x2(x0, x1);
    out[0] = h;
    float h0, h1;
    unpack_bf16x2(h, h0, h1);
    const float r0 = x0 - h0, r1 = x1 - h1;
    const uint32_t m = pack_bf16x2(r0, r1);
    out[1] = m;
    if constexpr (TERMS == 3) {
        float m0, m1;
        unpack_bf16x2(m, m0, m1);
        out[2] = pack_bf16x2(r0 - m0, r1 - m1);
    }
}

// Workgroup id -> logical tile id so that the workgroups each XCD receives
// (hardware dispatches round-robin: id % 8) hold a contiguous logical range,
// which the grouped tile order below turns into a compact block of C whose A
// and B panels stay in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int grid) {
    const int per = grid >> 3, extra = grid & 7;
    const int xcd = bid & 7, local = bid >> 3;
    return (xcd < extra ? xcd * (per + 1) : extra * (per + 1) + (xcd - extra) * per) + local;
}

// Logical tile id -> (tm, tn), walking GROUP_M tile rows at a time.
__device__ __forceinline__ void grouped_tile(int t, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int GROUP_M = 8;
    const int per_group = GROUP_M * tiles_n;
    const int group = t / per_group;
    const int first_m = group * GROUP_M;
    const int gsz = min(tiles_m - first_m, GROUP_M);
    const int in = t - group * per_group;
    tm = first_m + in % gsz;
    tn = in / gsz;
}

// Term index of the A (side 0) or B (side 1) factor of cross product p.
// split3: (lo,hi) (hi,lo) (hi,hi); split6: (lo,hi) (hi,lo) (mid,mid) (mid,hi)
// (hi,mid) (hi,hi), i.e. smallest magnitude first.
__host__ __device__ constexpr int product_term(int mode, int p, int side) {
    return mode == 1 ? (side == 0 ? (p == 0 ? 1 : 0) : (p == 1 ? 1 : 0))
                     : (side == 0 ? (p == 0 ? 2 : (p == 2 || p == 3) ? 1 : 0)
                                  : (p == 1 ? 2 : (p == 2 || p == 4) ? 1 : 0));
}

template <int MODE, int BM, int BN, int BK, int WGM, int WGN, int PF_ = prefetch_sets(MODE, BM, BN)>
struct Tile {
    static constexpr int LU = mode_lu(MODE);
    static constexpr int TERMS = mode_terms(MODE);
    static constexpr int CPR = BK / LU;             // 16-byte chunks per row of a k-block
    static constexpr int KG = CPR / 2;              // 2-chunk k-groups (one MFMA fragment deep)
    static constexpr int A_UNITS = BM * CPR;        // A unit: one chunk of one row
    // B unit: LU k-rows x BW columns.
    static constexpr int BW = b_unit_cols(MODE, BN, BK, WGM * WGN);
    static constexpr int B_UNITS = CPR * (BN / BW);
    static constexpr int NT = 64 * WGM * WGN;       // threads per workgroup
    static constexpr int UA = (A_UNITS + NT - 1) / NT;
    static constexpr int UB = (B_UNITS + NT - 1) / NT;
    static constexpr int A_BYTES = operand_bytes(MODE, BM, BK);
    static constexpr int B_BYTES = operand_bytes(MODE, BN, BK);
    static constexpr int STAGE = A_BYTES + B_BYTES;
    static constexpr int SMEM = 2 * STAGE;
    static constexpr int WTM = BM / WGM, WTN = BN / WGN;
    static constexpr int TM = WTM / 32, TN = WTN / 32;

    static_assert(WGM * WGN == 4 || WGM * WGN == 8, "four or eight waves per workgroup");
    static_assert(BK % (2 * LU) == 0, "k-block must hold whole fragments");
    static_assert(TM >= 1 && TN >= 1, "wave tile must hold a 32x32 block");
    static_assert(SMEM <= 160 * 1024, "LDS budget of one CU");

    // Register sets for k-blocks in flight between global memory and LDS.
    static constexpr int PF = PF_;
    static_assert(PF >= 1 && PF <= 4, "one to four register sets");
    f32x4 ra[PF][UA][LU / 4];  // A unit i: LU consecutive k of one row
    using bvec = float __attribute__((ext_vector_type(BW)));
    using bvecu = float __attribute__((ext_vector_type(BW), aligned(4)));
    bvec rb[PF][UB][LU];       // B unit i: row e holds BW consecutive columns
    f32x16 acc[TM][TN];

    // Global -> registers for the k-block at k0. fast: the whole block is inside
    // M, N and [k0, kend), so every unit is a plain vector load. Otherwise each
    // unit is a vector load when it is inside, zero when it is wholly outside,
    // and element-wise only when it straddles an edge.
    template <int S>
    __device__ __forceinline__ void load(const float* __restrict__ A, const float* __restrict__ B,
                                         long long M, long long N, long long lda, long long ldb,
                                         long long m0, long long n0, long long k0, long long kend,
                                         bool fast) {
        load_as<S>(threadIdx.x, A, B, M, N, lda, ldb, m0, n0, k0, kend, fast);
    }

    // load() for staging thread `tid` (0..NT-1) of the tile, which need not be
    // threadIdx.x (the decoupled feed stages with waves of its own).
    template <int S>
    __device__ __forceinline__ void load_as(int tid, const float* __restrict__ A, const float* __restrict__ B,
                                            long long M, long long N, long long lda, long long ldb,
                                            long long m0, long long n0, long long k0, long long kend,
                                            bool fast) {
#pragma unroll
        for (int i = 0; i < UA; ++i) {
            const int u = tid + NT * i;
            if (A_UNITS % NT == 0 || u < A_UNITS) {
                const int row = u / CPR, c = u % CPR;
                const long long gm = m0 + row, gk = k0 + c * LU;
                const float* src = A + gm * lda + gk;
                if (fast || (gm < M && gk + LU <= kend)) {
#pragma unroll
                    for (int q = 0; q < LU / 4; ++q) ra[S][i][q] = *reinterpret_cast<const f32x4u*>(src + 4 * q);
                } else {
#pragma unroll
                    for (int e = 0; e < LU; ++e) {
                        float x = 0.f;
                        if (gm < M && gk + e < kend) x = src[e];
                        ra[S][i][e / 4][e % 4] = x;
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < UB; ++i) {
            const int u = tid + NT * i;
            if (B_UNITS % NT == 0 || u < B_UNITS) {
                const int nb = u % (BN / BW), c = u / (BN / BW);
                const long long gn = n0 + BW * nb, gk = k0 + c * LU;
                const float* src = B + gk * ldb + gn;
                if (fast || (gn + BW <= N && gk + LU <= kend)) {
#pragma unroll
                    for (int e = 0; e < LU; ++e) rb[S][i][e] = *reinterpret_cast<const bvecu*>(src + e * ldb);
                } else {
#pragma unroll
                    for (int e = 0; e < LU; ++e) {
#pragma unroll
                        for (int j = 0; j < BW; ++j) {
                            float x = 0.f;
                            if (gk + e < kend && gn + j < N) x = src[e * ldb + j];
                            rb[S][i][e][j] = x;
                        }
                    }
                }
            }
        }
    }

    // load_as() for a k-block wholly inside the K range on a tile that may
    // cross M or N (N at least BW). Every unit is one unconditional vector
    // load: A units past M read the clamped last row and are zeroed by select;
    // a B unit that crosses N reads the last BW columns and shifts them into
    // place by selects, and one past N reads them and is zeroed. Without
    // branches around the loads, a caller that keeps several blocks of loads in
    // flight keeps them in flight (the compiler's wait counting loses their
    // order across conditional loads).
    template <int S>
    __device__ __forceinline__ void load_edge_as(int tid, const float* __restrict__ A, const float* __restrict__ B,
                                                 long long M, long long N, long long lda, long long ldb,
                                                 long long m0, long long n0, long long k0) {
#pragma unroll
        for (int i = 0; i < UA; ++i) {
            const int u = tid + NT * i;
            if (A_UNITS % NT == 0 || u < A_UNITS) {
                const int row = u / CPR, c = u % CPR;
                const long long gm = m0 + row, gk = k0 + c * LU;
                const bool in = gm < M;
                const float* src = A + (in ? gm : M - 1) * lda + gk;
#pragma unroll
                for (int q = 0; q < LU / 4; ++q) {
                    const f32x4 v = *reinterpret_cast<const f32x4u*>(src + 4 * q);
                    ra[S][i][q] = in ? v : f32x4{0.f, 0.f, 0.f, 0.f};
                }
            }
        }
#pragma unroll
        for (int i = 0; i < UB; ++i) {
            const int u = tid + NT * i;
            if (B_UNITS % NT == 0 || u < B_UNITS) {
                const int nb = u % (BN / BW), c = u / (BN / BW);
                const long long gn = n0 + BW * nb, gk = k0 + c * LU;
                const long long base = gn + BW <= N ? gn : N - BW;  // in range: N >= BW
                const int sh = gn < N ? (int)(gn - base) : 0;       // 0 inside, 1..BW-1 across N
                const float* src = B + gk * ldb + base;
#pragma unroll
                for (int e = 0; e < LU; ++e) {
                    const bvec v = *reinterpret_cast<const bvecu*>(src + e * ldb);
                    bvec o;
#pragma unroll
                    for (int j = 0; j < BW; ++j) {
                        float x = v[j];
#pragma unroll
                        for (int t = 1; t < BW - j; ++t) x = sh == t ? v[j + t] : x;
                        o[j] = gn + j < N ? x : 0.f;
                    }
                    rb[S][i][e] = o;
                }
            }
        }
    }

    template <int S>
    __device__ __forceinline__ void store(char* stage) const { store_as<S>(threadIdx.x, stage); }

    template <int S>
    __device__ __forceinline__ void store_as(int tid, char* stage) const {
#pragma unroll
        for (int i = 0; i < UA; ++i) {
            const int u = tid + NT * i;
            if (A_UNITS % NT == 0 || u < A_UNITS) {
                const int row = u / CPR, c = u % CPR;
                if constexpr (MODE == 0) {
                    *reinterpret_cast<f32x4*>(stage + chunk_offset(MODE, 0, BM, BK, row, c, 0)) = ra[S][i][0];
                } else {
                    uint32_t w[TERMS][4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        uint32_t t[TERMS];
                        split_pair<TERMS>(ra[S][i][q / 2][2 * (q % 2)], ra[S][i][q / 2][2 * (q % 2) + 1], t);
#pragma unroll
                        for (int s = 0; s < TERMS; ++s) w[s][q] = t[s];
                    }
#pragma unroll
                    for (int s = 0; s < TERMS; ++s) {
                        const u32x4 v = {w[s][0], w[s][1], w[s][2], w[s][3]};
                        *reinterpret_cast<u32x4*>(stage + chunk_offset(MODE, 0, BM, BK, row, c, s)) = v;
                    }
                }
            }
        }
        char* bstage = stage + A_BYTES;
#pragma unroll
        for (int i = 0; i < UB; ++i) {
            const int u = tid + NT * i;
            if (B_UNITS % NT == 0 || u < B_UNITS) {
                const int nb = u % (BN / BW), c = u / (BN / BW);
#pragma unroll
                for (int j = 0; j < BW; ++j) {
                    const int n = BW * nb + j;
                    if constexpr (MODE == 0) {
                        const f32x4 v = {rb[S][i][0][j], rb[S][i][1][j], rb[S][i][2][j], rb[S][i][3][j]};
                        *reinterpret_cast<f32x4*>(bstage + chunk_offset(MODE, 1, BN, BK, n, c, 0)) = v;
                    } else {
                        uint32_t w[TERMS][4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            uint32_t t[TERMS];
                            split_pair<TERMS>(rb[S][i][2 * q][j], rb[S][i][2 * q + 1][j], t);
#pragma unroll
                            for (int s = 0; s < TERMS; ++s) w[s][q] = t[s];
                        }
#pragma unroll
                        for (int s = 0; s < TERMS; ++s) {
                            const u32x4 v = {w[s][0], w[s][1], w[s][2], w[s][3]};
                            *reinterpret_cast<u32x4*>(bstage + chunk_offset(MODE, 1, BN, BK, n, c, s)) = v;
                        }
                    }
                }
            }
        }
    }

    __device__ __forceinline__ void compute(const char* stage) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int wm = wave / WGN, wn = wave % WGN;
        const int r = lane & 31, h = lane >> 5;
        const char* bstage = stage + A_BYTES;
#pragma unroll
        for (int kg = 0; kg < KG; ++kg) {
            const int c = 2 * kg + h;
            if constexpr (MODE == 0) {
                f32x4 a[TM], b[TN];
#pragma unroll
                for (int ti = 0; ti < TM; ++ti)
                    a[ti] = *reinterpret_cast<const f32x4*>(
                        stage + chunk_offset(MODE, 0, BM, BK, wm * WTM + ti * 32 + r, c, 0));
#pragma unroll
                for (int tj = 0; tj < TN; ++tj)
                    b[tj] = *reinterpret_cast<const f32x4*>(
                        bstage + chunk_offset(MODE, 1, BN, BK, wn * WTN + tj * 32 + r, c, 0));
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int ti = 0; ti < TM; ++ti)
#pragma unroll
                        for (int tj = 0; tj < TN; ++tj)
                            acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                                b[tj][s], a[ti][s], acc[ti][tj], 0, 0, 0);
            } else {
                // B fragments once per k-group; A fragments two 32-row blocks
                // at a time (register budget of two waves per SIMD), products
                // outer, so MFMAs into one accumulator are 2*TN apart.
                bf16x8 b[TERMS][TN];
#pragma unroll
                for (int s = 0; s < TERMS; ++s)
#pragma unroll
                    for (int tj = 0; tj < TN; ++tj)
                        b[s][tj] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(
                            bstage + chunk_offset(MODE, 1, BN, BK, wn * WTN + tj * 32 + r, c, s)));
                constexpr int TP = TM >= 2 ? 2 : 1;
#pragma unroll
                for (int t0 = 0; t0 < TM; t0 += TP) {
                    bf16x8 a[TERMS][TP];
#pragma unroll
                    for (int i = 0; i < TP; ++i)
#pragma unroll
                        for (int s = 0; s < TERMS; ++s)
                            a[s][i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(
                                stage + chunk_offset(MODE, 0, BM, BK, wm * WTM + (t0 + i) * 32 + r, c, s)));
                    // Cross products, smallest first, all into the one FP32
                    // accumulator: split3 = lo*hi + hi*lo + hi*hi; split6 adds the
                    // mid terms (Henry, Tang & Heinecke, ARITH 2019).
                    constexpr int NP = MODE == 1 ? 3 : 6;
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        const int ta = product_term(MODE, p, 0);
                        const int tb = product_term(MODE, p, 1);
#pragma unroll
                        for (int i = 0; i < TP; ++i)
#pragma unroll
                            for (int tj = 0; tj < TN; ++tj)
                                acc[t0 + i][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                    b[tb][tj], a[ta][i], acc[t0 + i][tj], 0, 0, 0);
                    }
                }
            }
        }
    }

    // Accumulator (lane l, register 4q + e) of block (ti, tj) holds
    // C[m0 + wm*WTM + ti*32 + l%32][n0 + wn*WTN + tj*32 + 8q + 4(l/32) + e]:
    // four consecutive columns per q, stored as one dwordx4 inside N.
    __device__ __forceinline__ void epilogue(float* __restrict__ C, long long M, long long N,
                                             long long ldc, long long m0, long long n0) const {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int wm = wave / WGN, wn = wave % WGN;
        const int r = lane & 31, h = lane >> 5;
#pragma unroll
        for (int ti = 0; ti < TM; ++ti) {
            const long long m = m0 + wm * WTM + ti * 32 + r;
            if (m >= M) continue;
            float* crow = C + m * ldc;
#pragma unroll
            for (int tj = 0; tj < TN; ++tj) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const long long n = n0 + wn * WTN + tj * 32 + 8 * q + 4 * h;
                    const f32x4 v = {acc[ti][tj][4 * q + 0], acc[ti][tj][4 * q + 1],
                                     acc[ti][tj][4 * q + 2], acc[ti][tj][4 * q + 3]};
                    if (n + 3 < N) {
                        *reinterpret_cast<f32x4u*>(crow + n) = v;
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (n + e < N) crow[n + e] = v[e];
                    }
                }
            }
        }
    }

    // The epilogue through LDS (smem free again, every wave past the main
    // loop's last barrier), 32 rows of the wave tile at a time: the
    // accumulators are written to a padded row-major block per wave, then
    // read back whole rows per instruction, so every global store covers
    // complete 128-byte lines of C instead of 32-byte pieces of 32 rows. Same
    // values as epilogue(). Used by the 256 x 256 tile, whose C writes it cuts
    // from 1.46x to 1.00x the size of C (4096^3 split3: 93.2 -> 64.0 MiB) at
    // the same speed.
    __device__ __forceinline__ void epilogue_lds(float* __restrict__ C, long long M, long long N, long long ldc,
                                                 long long m0, long long n0, char* smem) const {
        constexpr int NW = WGM * WGN;
        constexpr int PITCH = epi_pitch(WTN);  // floats (layout in lds_layout.hpp)
        static_assert(NW * 32 * PITCH * 4 <= SMEM, "epilogue blocks fit in the tile's LDS");
        constexpr int RPI = 256 / WTN;         // rows per read instruction (WTN / 4 lanes per row)
        static_assert(WTN == 64, "the staging layout is conflict-free at 64-wide wave tiles");
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int wm = wave / WGN, wn = wave % WGN;
        float* blk = reinterpret_cast<float*>(smem) + wave * 32 * PITCH;
        const int cc = epi_read_col(WTN, lane);
#pragma unroll
        for (int ti = 0; ti < TM; ++ti) {
#pragma unroll
            for (int tj = 0; tj < TN; ++tj)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const f32x4 v = {acc[ti][tj][4 * q + 0], acc[ti][tj][4 * q + 1], acc[ti][tj][4 * q + 2],
                                     acc[ti][tj][4 * q + 3]};
                    *reinterpret_cast<f32x4*>(blk + epi_write_offset(WTN, lane, tj, q)) = v;
                }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 32 / RPI; ++i) {
                const int row = epi_read_row(WTN, lane, i);
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

    // Main loop over the k-blocks of [kbeg, kend). ALL_FAST: every block is a
    // full block (vector loads, no bounds checks), so the steady-state
    // body is one basic block and the scheduler can interleave the MFMAs of
    // stage kb with the split/store of stage kb+1.
    //
    // One register set (PF = 1): registers for block kb+2 are loaded once the
    // store of kb+1 has consumed them, one full iteration (plus the barrier)
    // before they are needed. Two sets (PF = 2): the sets alternate and the
    // loop is unrolled by two, so blocks kb+2 and kb+3 are both in flight
    // while block kb is multiplied.
    //
    // (Measured slower here and not used: the feed's clamped edge loads,
    // load_edge_as, which add selects on tiles inside M and N and re-read the
    // clamped row on tiles mostly past M (the 256-problem grouped batch went
    // from 12.7 to 14.2 ms); and choosing per k-block between a straight run
    // of vector loads and the general path (split6 128 x 128 on 2000^3 went
    // from 109 to 127 us).)
    template <bool ALL_FAST>
    __device__ __forceinline__ void mainloop(const float* __restrict__ A, const float* __restrict__ B,
                                             long long M, long long N, long long lda, long long ldb,
                                             long long m0, long long n0, long long kbeg, long long kend,
                                             bool full_mn, char* smem) {
        static_assert(PF <= 2, "the tile's own loop runs one or two register sets");
        const long long nkb = (kend - kbeg + BK - 1) / BK;
        auto fast = [&](long long k0) { return ALL_FAST || (full_mn && k0 + BK <= kend); };
        if constexpr (PF == 1) {
            auto fetch = [&](long long k0) { load<0>(A, B, M, N, lda, ldb, m0, n0, k0, kend, fast(k0)); };
            fetch(kbeg);
            store<0>(smem);
            if (nkb > 1) fetch(kbeg + BK);
            __syncthreads();
            for (long long kb = 0; kb + 1 < nkb; ++kb) {
                compute(smem + (kb & 1) * STAGE);
                store<0>(smem + ((kb + 1) & 1) * STAGE);
                if (kb + 2 < nkb) fetch(kbeg + (kb + 2) * BK);
                __syncthreads();
            }
            compute(smem + ((nkb - 1) & 1) * STAGE);
        } else {
            auto fetch0 = [&](long long k0) { load<0>(A, B, M, N, lda, ldb, m0, n0, k0, kend, fast(k0)); };
            auto fetch1 = [&](long long k0) { load<1>(A, B, M, N, lda, ldb, m0, n0, k0, kend, fast(k0)); };
            // Loop invariant at even kb: stage 0 holds block kb, register set 1
            // block kb+1, register set 0 block kb+2 (those that exist).
            fetch0(kbeg);
            if (nkb > 1) fetch1(kbeg + BK);
            store<0>(smem);
            if (nkb > 2) fetch0(kbeg + 2 * BK);
            __syncthreads();
            long long kb = 0;
            for (; kb + 2 < nkb; kb += 2) {
                compute(smem);
                store<1>(smem + STAGE);
                if (kb + 3 < nkb) fetch1(kbeg + (kb + 3) * BK);
                __syncthreads();
                compute(smem + STAGE);
                store<0>(smem);
                if (kb + 4 < nkb) fetch0(kbeg + (kb + 4) * BK);
                __syncthreads();
            }
            compute(smem);
            if (kb + 1 < nkb) {
                store<1>(smem + STAGE);
                __syncthreads();
                compute(smem + STAGE);
            }
        }
        __syncthreads();
    }

    // Whole tile: C[m0:m0+BM, n0:n0+BN] = A[m0:, kbeg:kend] * B[kbeg:kend, n0:].
    // smem holds SMEM bytes; on return every wave has passed the last barrier,
    // so the caller may reuse smem for the next tile without another barrier.
    __device__ __forceinline__ void run(const float* __restrict__ A, const float* __restrict__ B,
                                        float* __restrict__ C, long long M, long long N,
                                        long long lda, long long ldb, long long ldc,
                                        long long m0, long long n0, long long kbeg, long long kend,
                                        char* smem) {
#pragma unroll
        for (int ti = 0; ti < TM; ++ti)
#pragma unroll
            for (int tj = 0; tj < TN; ++tj)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[ti][tj][e] = 0.f;

        const bool full_mn = (m0 + BM <= M) && (n0 + BN <= N);
        if (full_mn && (kend - kbeg) % BK == 0)
            mainloop<true>(A, B, M, N, lda, ldb, m0, n0, kbeg, kend, full_mn, smem);
        else
            mainloop<false>(A, B, M, N, lda, ldb, m0, n0, kbeg, kend, full_mn, smem);
        if constexpr (BM == 256 && BN == 256)
            epilogue_lds(C, M, N, ldc, m0, n0, smem);
        else
            epilogue(C, M, N, ldc, m0, n0);
    }
};

}  // namespace mmsk
