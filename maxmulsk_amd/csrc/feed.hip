// Decoupled producer/consumer feed for the 64 x 64 tile (geometry 0).
//
// Counterpart of the reference's Trio (sme/trio/trio_sme.cpp:274-334,
// sme/trio/trio_packers.cpp:31-107): one compute thread that only runs the
// micro-kernel and writes C, fed by separate threads that move and rearrange
// the operands into small rings, talking through per-slot ready/consumed
// sequence counters; same blocking and product order as the base kernel, so
// the result is bit-identical to it.
//
// On gfx950 the threads are the waves of one 512-thread workgroup:
//   waves 0..3 (consumers): the 2 x 2 wave layout of the direct 64 x 64 tile;
//     they only read LDS fragments, run the MFMAs and write the tile (or its
//     split-K partial);
//   waves 4..7 (producers): the tile's global -> register staging with their
//     own thread index, the split into BF16 terms and the fragment-major,
//     swizzled LDS stores, into a ring of NS slots of one BK-deep k-block
//     each, with up to PFR k-blocks of loads in flight per thread.
// Slot j % NS holds k-block j. Per slot two counters in LDS count waves:
//   ready[s]  += 1 by each producer wave once its stores of a block are done;
//   freed[s]  += 1 by each consumer wave once its reads of a block are done.
// A producer fills k-block j after freed[j % NS] reaches 4 * (j / NS); a
// consumer reads it after ready[j % NS] reaches 4 * (j / NS + 1). Counts only
// grow, so no reset is needed and no slot is read before it is written or
// overwritten before it is read. Every wait is bounded: past the bound the
// wave raises a workgroup abort flag, every wave leaves its loop, and the
// consumers write NaN to their tile, so a broken handshake shows up as wrong
// output and never as a hang.
#include <hip/hip_runtime.h>

#include <atomic>
#include <type_traits>
#include <cstdlib>
#include <cstring>

#include "gemm_tile.hpp"
#include "kernels.hpp"

namespace mmsk {

namespace {

// About 2^21 polls of one LDS round trip each (about 100 cycles): tens of
// milliseconds, far beyond any real wait (a k-block is microseconds). The
// polls do not sleep: with s_sleep between polls every variant measured
// slower. MAXMULSK_TEST_FEED_SPIN_LIMIT (read once) replaces the bound, so a
// test can drive the abort path from the host.
constexpr unsigned kFeedSpinLimit = 1u << 21;

__device__ __forceinline__ unsigned lds_peek(unsigned* p) {
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
}

// Waits until *cnt >= target. False when the workgroup aborted or this wave
// ran out of polls (it then raises the abort flag).
__device__ __forceinline__ bool wait_count(unsigned* cnt, unsigned target, unsigned* abort_flag, unsigned limit) {
    for (unsigned spins = 0;; ++spins) {
        if (lds_peek(cnt) >= target) break;
        if (lds_peek(abort_flag) != 0u) return false;
        if (spins >= limit) {
            __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return false;
        }
    }
    // LDS accesses of the slot stay after the observed count.
    asm volatile("" ::: "memory");
    return true;
}

// This wave's LDS accesses so far are complete (LDS executes one wave's
// accesses in order), then its lane 0 adds one to *cnt.
__device__ __forceinline__ void signal(unsigned* cnt) {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt left at their maxima
    asm volatile("" ::: "memory");
    if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// A BM x BN tile with a CWM x CWN wave layout of consumers (the direct tile's
// MFMA layout) and a PWM x PWN layout of producers (the direct tile's staging
// map for that many threads); both sides address the same fragment-major slot
// layout, which depends only on the mode, BM, BN and BK.
template <int MODE, int BM, int BN, int BK, int NS, int PFR, int CWM, int CWN, int PWM, int PWN>
struct Feed {
    using Cons = Tile<MODE, BM, BN, BK, CWM, CWN, 1>;
    using Prod = Tile<MODE, BM, BN, BK, PWM, PWN, PFR>;
    static constexpr int CT = Cons::NT, PT = Prod::NT;  // consumer, producer threads
    static constexpr unsigned CWAVES = CT / 64, PWAVES = PT / 64;
    static constexpr int STAGE = Cons::STAGE;
    static constexpr int RING = NS * STAGE;
    static_assert(Prod::STAGE == STAGE, "one slot layout");
    static_assert(CT + PT <= 1024, "one workgroup");
    static_assert(RING <= 160 * 1024 - 256, "LDS budget of one CU");

    struct Shared {
        unsigned ready[NS];
        unsigned freed[NS];
        unsigned abort_flag;
    };

    // Producer step for k-block j from register set S (when j < nkb), then
    // the loads of k-block j + PFR into the same set. The loads are issued on
    // every step, past the last block as repeats of it (the block index is
    // clamped), so each step's loads are unconditional: with conditional
    // loads the compiler's wait counting lost the order of the sets in flight
    // and waited for all of them before every store.
    template <int S, bool ALL_FAST>
    __device__ __forceinline__ static bool produce(Prod& t, int ptid, long long j, long long nkb, char* ring,
                                                   Shared& sh, const GemmParams& p, long long m0, long long n0,
                                                   long long kbeg, long long kend, bool full_mn, unsigned limit) {
        bool ok = true;
        if (j < nkb) {
            const int slot = (int)(j % NS);
            const unsigned round = (unsigned)(j / NS);
            ok = round == 0 || wait_count(&sh.freed[slot], CWAVES * round, &sh.abort_flag, limit);
            if (ok) {
                t.template store_as<S>(ptid, ring + slot * STAGE);
                signal(&sh.ready[slot]);
            }
        }
        const long long jn = min(j + PFR, nkb - 1);
        const long long k0 = kbeg + jn * BK;
        t.template load_as<S>(ptid, p.A, p.B, p.M, p.N, p.lda, p.ldb, m0, n0, k0, kend,
                              ALL_FAST || (full_mn && k0 + BK <= kend));
        return ok;
    }

    template <bool ALL_FAST>
    __device__ __forceinline__ static void producer(char* ring, Shared& sh, const GemmParams& p, long long m0,
                                                    long long n0, long long kbeg, long long kend, bool full_mn,
                                                    long long nkb, unsigned limit) {
        if (nkb <= 0) return;
        Prod t;
        const int ptid = (int)threadIdx.x - CT;
        auto fetch = [&](auto set, long long j) {
            const long long k0 = kbeg + min(j, nkb - 1) * BK;
            t.template load_as<decltype(set)::value>(ptid, p.A, p.B, p.M, p.N, p.lda, p.ldb, m0, n0, k0, kend,
                                                     ALL_FAST || (full_mn && k0 + BK <= kend));
        };
        // Loads of the first PFR k-blocks, one register set each.
        fetch(std::integral_constant<int, 0>{}, 0);
        if constexpr (PFR > 1) fetch(std::integral_constant<int, 1>{}, 1);
        if constexpr (PFR > 2) fetch(std::integral_constant<int, 2>{}, 2);
        if constexpr (PFR > 3) fetch(std::integral_constant<int, 3>{}, 3);
        for (long long j = 0; j < nkb; j += PFR) {
            // Every step runs (its loads are unconditional); an aborted wait
            // skips only that step's stores, and the loop ends after the group.
            bool ok = produce<0, ALL_FAST>(t, ptid, j, nkb, ring, sh, p, m0, n0, kbeg, kend, full_mn, limit);
            if constexpr (PFR > 1) ok &= produce<1, ALL_FAST>(t, ptid, j + 1, nkb, ring, sh, p, m0, n0, kbeg, kend, full_mn, limit);
            if constexpr (PFR > 2) ok &= produce<2, ALL_FAST>(t, ptid, j + 2, nkb, ring, sh, p, m0, n0, kbeg, kend, full_mn, limit);
            if constexpr (PFR > 3) ok &= produce<3, ALL_FAST>(t, ptid, j + 3, nkb, ring, sh, p, m0, n0, kbeg, kend, full_mn, limit);
            if (!ok) return;
        }
    }

    __device__ __forceinline__ static void consumer(char* ring, Shared& sh, float* dst, const GemmParams& p,
                                                    long long m0, long long n0, long long nkb, unsigned limit) {
        Cons t;
#pragma unroll
        for (int i = 0; i < Cons::TM; ++i)
#pragma unroll
            for (int j = 0; j < Cons::TN; ++j)
#pragma unroll
                for (int e = 0; e < 16; ++e) t.acc[i][j][e] = 0.f;
        bool ok = true;
        for (long long j = 0; j < nkb; ++j) {
            const int slot = (int)(j % NS);
            const unsigned round = (unsigned)(j / NS);
            if (!wait_count(&sh.ready[slot], PWAVES * (round + 1), &sh.abort_flag, limit)) {
                ok = false;
                break;
            }
            t.compute(ring + slot * STAGE);
            signal(&sh.freed[slot]);
        }
        if (!ok) {
#pragma unroll
            for (int i = 0; i < Cons::TM; ++i)
#pragma unroll
                for (int j = 0; j < Cons::TN; ++j)
#pragma unroll
                    for (int e = 0; e < 16; ++e) t.acc[i][j][e] = __builtin_nanf("");
        }
        t.epilogue(dst, p.M, p.N, p.ldc, m0, n0);
    }
};

}  // namespace

template <int MODE, int BM, int BN, int BK, int NS, int PFR, int CWM, int CWN, int PWM, int PWN>
__global__ __launch_bounds__(64 * (CWM * CWN + PWM * PWN)) void gemm_feed_kernel(GemmParams p, unsigned spin_limit) {
    using F = Feed<MODE, BM, BN, BK, NS, PFR, CWM, CWN, PWM, PWN>;
    __shared__ __attribute__((aligned(16))) char ring[F::RING];
    __shared__ typename F::Shared sh;
    if (threadIdx.x < NS) {
        sh.ready[threadIdx.x] = 0u;
        sh.freed[threadIdx.x] = 0u;
    }
    if (threadIdx.x == 0) sh.abort_flag = 0u;
    __syncthreads();

    const int ntiles = p.tiles_m * p.tiles_n;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int split = bid / ntiles;
    int tm, tn;
    grouped_tile(bid - split * ntiles, p.tiles_m, p.tiles_n, tm, tn);
    const long long kbeg = split * p.k_len;
    const long long kend = min(p.K, kbeg + p.k_len);
    const long long m0 = (long long)tm * BM, n0 = (long long)tn * BN;
    const long long nkb = (kend - kbeg + BK - 1) / BK;
    const bool full_mn = (m0 + BM <= p.M) && (n0 + BN <= p.N);
    if (threadIdx.x < F::CT) {
        F::consumer(ring, sh, p.C + split * p.c_split_stride, p, m0, n0, nkb, spin_limit);
    } else if (full_mn && (kend - kbeg) % BK == 0) {
        F::template producer<true>(ring, sh, p, m0, n0, kbeg, kend, full_mn, nkb, spin_limit);
    } else {
        F::template producer<false>(ring, sh, p, m0, n0, kbeg, kend, full_mn, nkb, spin_limit);
    }
}

namespace {

std::atomic<int> g_feed{-1};

unsigned spin_limit() {
    static const unsigned v = [] {
        const char* e = std::getenv("MAXMULSK_TEST_FEED_SPIN_LIMIT");
        return e && *e ? (unsigned)std::strtoul(e, nullptr, 10) : kFeedSpinLimit;
    }();
    return v;
}

int feed_from_env() {
    const char* e = std::getenv("MAXMULSK_FEED");
    if (e && std::strcmp(e, "decoupled") == 0) return 1;
    if (e && std::strcmp(e, "direct") == 0) return 0;
    return 2;
}

template <int MODE, int BK, int NS, int PFR, int BM = 64, int BN = 64, int CWM = 2, int CWN = 2, int PWM = 2,
          int PWN = 2>
hipError_t launch_feed_cfg(const GemmParams& p, hipStream_t stream) {
    const long long grid = (long long)p.tiles_m * p.tiles_n * p.splits;
    if (grid <= 0 || grid > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((gemm_feed_kernel<MODE, BM, BN, BK, NS, PFR, CWM, CWN, PWM, PWN>), dim3((unsigned)grid),
                       dim3(64 * (CWM * CWN + PWM * PWN)), 0, stream, p, spin_limit());
    return hipGetLastError();
}

}  // namespace

int feed_mode() {
    int v = g_feed.load(std::memory_order_relaxed);
    if (v < 0) {
        int expect = -1;
        g_feed.compare_exchange_strong(expect, feed_from_env());
        v = g_feed.load(std::memory_order_relaxed);
    }
    return v;
}

void set_feed_mode(int v) { g_feed.store(v, std::memory_order_relaxed); }

// Variants (k-block depth BK, ring slots NS, k-blocks of loads in flight PFR)
// and where the automatic setting takes them. Measured against the direct
// tile on one MI355X: per-call times of chosen shapes, and every 64-tile
// candidate of 1200 shapes of the selector's table (M, N, K log-uniform in
// 1..8192, all three modes) timed by the autotuner under both feeds.
//   3 split6 BK 64, NS 2, PFR 2: one workgroup per CU and K slices of at
//     least 512: 1.06-1.40x on every such candidate with slices of 1024 or
//     more (median 1.10 at 512-1023), e.g. 768 x 768 x 4096 (64,1) 76.3 ->
//     63.0 us. Shorter slices lose (median 0.50 below 64, 0.93 at 128-255).
//   1 split3 BK 128, NS 2, PFR 2, and 2 split3 BK 128, NS 2, PFR 1: split3
//     gains only on grids of 64..256 workgroups with slices of at least 1024
//     (1024^3 (64,1) 14.6 -> 13.0 us with 2, 768 x 768 x 4096 (64,1) 48.4 ->
//     38.9 us with 1, 83 x 4819 x 6314 (64,1) 1.12x), and is even or 2-7%
//     slower on smaller grids and shorter slices. Variant 1 (two blocks of
//     loads in flight) for slices of 4096 or more, 2 below.
//   7 split3 128 x 128, BK 64, NS 2, PFR 1, eight consumer and eight producer
//     waves, and 8 split6 128 x 128, BK 32, NS 2, PFR 2, eight consumer and
//     four producer waves: grids of at most one workgroup per CU whose tiles
//     and k-blocks are all full (M, N multiples of 128; K and the slice
//     multiples of 64), slices of at least 1024: 2048^3 (128,1) split3 61.7 ->
//     59.7 us, 512 x 4096 x 4096 94.5 -> 83.7 us, 1152 x 1152 x 8192 184.2 ->
//     161.6 us; split6 1.03-1.10x. Ragged tiles or a partial last k-block put
//     the element-wise edge loads on the producers alone: 3-18% slower
//     (1920 x 1920 x 2000, 1000 x 1100 x 3000), so those stay direct, as do
//     grids of more than one workgroup per CU (4096 x 4096 x 1024: 0.94x).
//     Plain keeps the direct 128 x 128 tile.
//   4 plain BK 64, NS 2, PFR 2; 5 split3 BK 64, NS 2, PFR 2; 6 split6 BK 32,
//     NS 4, PFR 2: only when the feed is forced. Plain was 3-40% slower on
//     small grids with every variant, and on grids of more than one
//     workgroup per CU all modes were 1-25% slower (the direct tile hides
//     latency with several workgroups per CU instead).
// Slices shorter than the k-block depth are zero-padded, as in the deep
// direct variants: results are those of the direct tile bit for bit, except
// that an exact -0 may come out as +0.
int feed_variant(int mode, int cfg, long long M, long long N, long long K, long long grid, long long k_len,
                 int setting) {
    if (setting == 0 || mode < 0 || mode > 2) return 0;
    const bool small_grid = grid > 0 && grid <= kDeepMaxGrid;
    if (cfg == 3) {
        // 128 x 128 (split modes): only where every tile and k-block is full;
        // plain keeps the direct tile.
        if (mode == 0) return 0;
        if (setting == 1) return mode == 1 ? 7 : 8;
        const bool whole = M % 128 == 0 && N % 128 == 0 && K % 64 == 0 && k_len % 64 == 0;
        return (small_grid && whole && k_len >= 1024) ? (mode == 1 ? 7 : 8) : 0;
    }
    if (cfg != 0) return 0;
    if (setting == 2) {  // measured rule
        if (!small_grid) return 0;
        if (mode == 1) return (grid >= 64 && k_len >= 1024) ? (k_len >= 4096 ? 1 : 2) : 0;
        return (mode == 2 && k_len >= 512) ? 3 : 0;
    }
    if (mode == 0) return 4;
    if (mode == 1) return small_grid ? (k_len >= 4096 ? 1 : 2) : 5;
    return small_grid ? 3 : 6;
}

hipError_t launch_gemm_feed(int variant, const GemmParams& p, hipStream_t stream) {
    switch (variant) {
        case 1: return launch_feed_cfg<1, 128, 2, 2>(p, stream);
        case 2: return launch_feed_cfg<1, 128, 2, 1>(p, stream);
        case 3: return launch_feed_cfg<2, 64, 2, 2>(p, stream);
        case 4: return launch_feed_cfg<0, 64, 2, 2>(p, stream);
        case 5: return launch_feed_cfg<1, 64, 2, 2>(p, stream);
        case 6: return launch_feed_cfg<2, 32, 4, 2>(p, stream);
        case 7: return launch_feed_cfg<1, 64, 2, 1, 128, 128, 2, 4, 2, 4>(p, stream);
        case 8: return launch_feed_cfg<2, 32, 2, 2, 128, 128, 2, 4, 2, 2>(p, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mmsk
