// Decoupled producer/consumer feed: the direct tile's work split between
// waves that only multiply and waves that only stage operands, for the 64 x 64
// tile (geometry 0), the split modes' 128 x 128 tile (geometry 3), and the
// split modes' grouped persistent launch (geometries 1..4).
//
// Counterpart of the reference's Trio (sme/trio/trio_sme.cpp:274-334,
// sme/trio/trio_packers.cpp:31-107): one compute thread that only runs the
// micro-kernel and writes C, fed by separate threads that move and rearrange
// the operands into small rings, talking through per-slot ready/consumed
// sequence counters; same blocking and product order as the base kernel, so
// the result is bit-identical to it.
//
// On gfx950 the threads are the waves of one workgroup, e.g. for 64 x 64:
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
// With P producer and C consumer waves, a producer fills k-block j after
// freed[j % NS] reaches C * (j / NS); a consumer reads it after
// ready[j % NS] reaches P * (j / NS + 1). Counts only grow, so no reset is
// needed and no slot is read before it is written or overwritten before it is
// read. Every wait is bounded: past the bound the wave raises a workgroup
// abort flag, every wave leaves its loop, and the consumers write NaN to their
// tile, so a broken handshake shows up as wrong output and never as a hang.
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

// Ring header: per-slot counters and the workgroup abort flag.
template <int NS>
struct FeedShared {
    unsigned ready[NS];
    unsigned freed[NS];
    unsigned abort_flag;
};

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

    static constexpr int kBM = BM, kBN = BN, kBK = BK, kNS = NS;
    using Shared = FeedShared<NS>;

    // How a producer step loads a k-block: every unit a plain vector load
    // (k-block inside M, N and K), one clamped vector load per unit (k-block
    // inside K, tile across M or N), or the tile's general path (vector loads
    // inside, element-wise across an edge).
    enum { kLoadFull = 0, kLoadEdge = 1, kLoadGeneral = 2 };

    template <int S, int POL>
    __device__ __forceinline__ static void load_block(Prod& t, int ptid, const GemmParams& p, long long m0,
                                                      long long n0, long long k0, long long kend, bool full_mn) {
        if constexpr (POL == kLoadFull)
            t.template load_as<S>(ptid, p.A, p.B, p.M, p.N, p.lda, p.ldb, m0, n0, k0, kend, true);
        else if constexpr (POL == kLoadEdge)
            t.template load_edge_as<S>(ptid, p.A, p.B, p.M, p.N, p.lda, p.ldb, m0, n0, k0);
        else
            t.template load_as<S>(ptid, p.A, p.B, p.M, p.N, p.lda, p.ldb, m0, n0, k0, kend,
                                  full_mn && k0 + BK <= kend);
    }

    // Producer step for k-block j from register set S (when j < nloop), then
    // the loads of k-block j + PFR into the same set. The loads are issued on
    // every step, past the loop's last block as repeats of it (the block index
    // is clamped), so each step's loads are unconditional: with conditional
    // loads the compiler's wait counting lost the order of the sets in flight
    // and waited for all of them before every store.
    template <int S, int POL>
    __device__ __forceinline__ static bool produce(Prod& t, int ptid, long long j, long long nloop, char* ring,
                                                   Shared& sh, const GemmParams& p, long long m0, long long n0,
                                                   long long kbeg, long long kend, bool full_mn, long long jb,
                                                   unsigned limit) {
        bool ok = true;
        if (j < nloop) {
            const int slot = (int)((jb + j) % NS);
            const unsigned round = (unsigned)((jb + j) / NS);
            ok = round == 0 || wait_count(&sh.freed[slot], CWAVES * round, &sh.abort_flag, limit);
            if (ok) {
                t.template store_as<S>(ptid, ring + slot * STAGE);
                signal(&sh.ready[slot]);
            }
        }
        const long long jn = min(j + PFR, nloop - 1);
        load_block<S, POL>(t, ptid, p, m0, n0, kbeg + jn * BK, kend, full_mn);
        return ok;
    }

    // k-blocks [0, nloop) with PFR blocks of loads in flight, every block
    // loaded under policy POL. False when a wait ran out.
    template <int POL>
    __device__ __forceinline__ static bool produce_blocks(Prod& t, int ptid, char* ring, Shared& sh,
                                                          const GemmParams& p, long long m0, long long n0,
                                                          long long kbeg, long long kend, bool full_mn,
                                                          long long nloop, long long jb, unsigned limit) {
        auto fetch = [&](auto set, long long j) {
            load_block<decltype(set)::value, POL>(t, ptid, p, m0, n0, kbeg + min(j, nloop - 1) * BK, kend, full_mn);
        };
        // Loads of the first PFR k-blocks, one register set each.
        fetch(std::integral_constant<int, 0>{}, 0);
        if constexpr (PFR > 1) fetch(std::integral_constant<int, 1>{}, 1);
        if constexpr (PFR > 2) fetch(std::integral_constant<int, 2>{}, 2);
        if constexpr (PFR > 3) fetch(std::integral_constant<int, 3>{}, 3);
        for (long long j = 0; j < nloop; j += PFR) {
            // Every step runs (its loads are unconditional); an aborted wait
            // skips only that step's stores, and the loop ends after the group.
            bool ok = produce<0, POL>(t, ptid, j, nloop, ring, sh, p, m0, n0, kbeg, kend, full_mn, jb, limit);
            if constexpr (PFR > 1) ok &= produce<1, POL>(t, ptid, j + 1, nloop, ring, sh, p, m0, n0, kbeg, kend, full_mn, jb, limit);
            if constexpr (PFR > 2) ok &= produce<2, POL>(t, ptid, j + 2, nloop, ring, sh, p, m0, n0, kbeg, kend, full_mn, jb, limit);
            if constexpr (PFR > 3) ok &= produce<3, POL>(t, ptid, j + 3, nloop, ring, sh, p, m0, n0, kbeg, kend, full_mn, jb, limit);
            if (!ok) return false;
        }
        return true;
    }

    // Full tiles with whole k-blocks: every block on the plain vector path.
    // Otherwise the k-blocks inside K run the plain vector path (tile inside
    // M and N) or the clamped one (tile across M or N, N at least the B unit
    // width), and a last partial k-block is loaded and stored on its own
    // afterwards; with N below the unit width every block takes the general
    // path.
    template <bool ALL_FAST>
    __device__ __forceinline__ static void producer(char* ring, Shared& sh, const GemmParams& p, long long m0,
                                                    long long n0, long long kbeg, long long kend, bool full_mn,
                                                    long long nkb, long long jb, unsigned limit) {
        if (nkb <= 0) return;
        Prod t;
        const int ptid = (int)threadIdx.x - CT;
        if constexpr (ALL_FAST) {
            (void)produce_blocks<kLoadFull>(t, ptid, ring, sh, p, m0, n0, kbeg, kend, full_mn, nkb, jb, limit);
        } else {
            const long long nwhole = (kend - kbeg) / BK;
            if (p.N >= Prod::BW && nwhole >= 1) {
                const bool whole_ok =
                    full_mn ? produce_blocks<kLoadFull>(t, ptid, ring, sh, p, m0, n0, kbeg, kend, full_mn, nwhole, jb, limit)
                            : produce_blocks<kLoadEdge>(t, ptid, ring, sh, p, m0, n0, kbeg, kend, full_mn, nwhole, jb, limit);
                if (!whole_ok) return;
                if (nwhole < nkb) {  // the partial last k-block
                    const long long j = nwhole;
                    load_block<0, kLoadGeneral>(t, ptid, p, m0, n0, kbeg + j * BK, kend, full_mn);
                    const int slot = (int)((jb + j) % NS);
                    const unsigned round = (unsigned)((jb + j) / NS);
                    if (round > 0 && !wait_count(&sh.freed[slot], CWAVES * round, &sh.abort_flag, limit)) return;
                    t.template store_as<0>(ptid, ring + slot * STAGE);
                    signal(&sh.ready[slot]);
                }
            } else {
                (void)produce_blocks<kLoadGeneral>(t, ptid, ring, sh, p, m0, n0, kbeg, kend, full_mn, nkb, jb, limit);
            }
        }
    }

    __device__ __forceinline__ static void consumer(char* ring, Shared& sh, float* dst, const GemmParams& p,
                                                    long long m0, long long n0, long long nkb, long long jb,
                                                    unsigned limit) {
        Cons t;
#pragma unroll
        for (int i = 0; i < Cons::TM; ++i)
#pragma unroll
            for (int j = 0; j < Cons::TN; ++j)
#pragma unroll
                for (int e = 0; e < 16; ++e) t.acc[i][j][e] = 0.f;
        bool ok = true;
        for (long long j = 0; j < nkb; ++j) {
            const int slot = (int)((jb + j) % NS);
            const unsigned round = (unsigned)((jb + j) / NS);
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
        F::consumer(ring, sh, p.C + split * p.c_split_stride, p, m0, n0, nkb, 0, spin_limit);
    } else if (full_mn && (kend - kbeg) % BK == 0) {
        F::template producer<true>(ring, sh, p, m0, n0, kbeg, kend, full_mn, nkb, 0, spin_limit);
    } else {
        F::template producer<false>(ring, sh, p, m0, n0, kbeg, kend, full_mn, nkb, 0, spin_limit);
    }
}

// ---- grouped persistent launch on the feed (split modes) ---------------------------
//
// The persistent workgroups of the grouped launch (grouped.hip) made of
// consumer and producer waves: each workgroup claims work items from the
// launch's counter and runs each one through its ring. The ring's counters
// keep counting across items: jb, the number of k-blocks the workgroup has
// passed through the ring so far, is the same in every wave (every wave runs
// every item's blocks), so the workgroup's block jb + j goes to slot
// (jb + j) % NS in round (jb + j) / NS. The kernel serves the direct grouped
// kernel's geometries 1..4 (64 x 128, 64 x 256, 128 x 128, 128 x 256) with the
// same work items, so every element gets the products of the direct launch
// in the same order. All geometries of a mode share the wave counts (four
// consumer waves in a 2 x 2 layout, so a 128 x 256 tile gives each 64 x 128
// and every fragment read from LDS feeds four MFMAs; four producer waves; 250
// VGPRs per lane in split3, 256 in split6 with a 2-byte spill), the ring slot
// count and one LDS ring sized for the largest. Measured against the direct
// grouped kernel on one MI355X, bit-identical every time:
//   the 256-problem batch (M ~ U{1..512}, N in {512, 2112, 4096, 7168}, K in
//     {512, 1536, 7168, 16384}; nearly every problem on 128 x 256 tiles):
//     split3 12.77 -> 11.69 ms, split6 18.90 -> 17.12 ms;
//   other batches, split3 / split6: K in {256, 512} 1.03x / 1.05x; M <= 128
//     with K in {8192, 16384} 1.20x / 1.15x; M <= 64 1.26x / 1.12x; M in
//     256..512 1.08x / 1.10x; N in {64, 96, 130, 200} 1.37x / 1.26x; eight
//     problems 1.23x / 1.31x; sixteen with K = 16384 1.12x / 1.07x.
// Variants measured on the 256-problem batch and not kept (split3 / split6,
// same session as the kept 11.69 / 17.12 ms unless noted): eight consumer
// waves of 64 x 64 with four producers, 12.40 / 17.50 ms (twice the LDS
// fragment reads per product); eight consumers and eight producers leave 128
// registers per lane and spill (13.3-16.9 / 20.7-21.8 ms, other sessions);
// four consumers with one k-block of loads in flight in split3, 1.05x over
// direct against 1.09x; eight producers with four consumers spill 857 bytes
// in split6 (0.14x); 16-deep k-blocks (21.6 / 20.1 ms, eight consumers);
// everything on 128 x 128 tiles with the standalone feed's 64-deep blocks
// (twice the items; split3 15.0 ms).
template <int MODE>
struct GroupedFeed;
template <>
struct GroupedFeed<1> {
    template <int BM, int BN>
    using G = Feed<1, BM, BN, 32, 2, 2, 2, 2, 2, 2>;
};
template <>
struct GroupedFeed<2> {
    template <int BM, int BN>
    using G = Feed<2, BM, BN, 32, 2, 1, 2, 2, 2, 2>;
};

template <int MODE>
constexpr int grouped_feed_nt() {
    using F = typename GroupedFeed<MODE>::template G<128, 256>;
    return F::CT + F::PT;
}

// One work item of problem q (item `local` of the problem) on geometry F;
// returns its k-block count.
template <class F>
__device__ __forceinline__ long long grouped_feed_item(const GroupProblem& q, int local, char* ring,
                                                       typename F::Shared& sh, long long jb, unsigned limit) {
    const int ntiles = q.tiles_m * q.tiles_n;
    const int split = local / ntiles;
    int tm, tn;
    grouped_tile(local - split * ntiles, q.tiles_m, q.tiles_n, tm, tn);
    const long long kbeg = split * q.k_len;
    const long long kend = min(q.K, kbeg + q.k_len);
    const long long m0 = (long long)tm * F::kBM, n0 = (long long)tn * F::kBN;
    const long long nkb = (kend - kbeg + F::kBK - 1) / F::kBK;
    const bool full_mn = (m0 + F::kBM <= q.M) && (n0 + F::kBN <= q.N);
    GemmParams p;
    p.A = q.A;
    p.B = q.B;
    p.C = nullptr;
    p.M = q.M;
    p.N = q.N;
    p.K = q.K;
    p.lda = q.K;
    p.ldb = q.N;
    p.ldc = q.N;
    p.k_len = q.k_len;
    p.c_split_stride = 0;
    p.splits = q.splits;
    p.tiles_m = q.tiles_m;
    p.tiles_n = q.tiles_n;
    float* dst = q.splits > 1 ? q.P + split * q.M * q.N : q.C;
    if (threadIdx.x < F::CT)
        F::consumer(ring, sh, dst, p, m0, n0, nkb, jb, limit);
    else if (full_mn && (kend - kbeg) % F::kBK == 0)
        F::template producer<true>(ring, sh, p, m0, n0, kbeg, kend, full_mn, nkb, jb, limit);
    else
        F::template producer<false>(ring, sh, p, m0, n0, kbeg, kend, full_mn, nkb, jb, limit);
    return nkb;
}

template <int MODE>
__global__ __launch_bounds__(grouped_feed_nt<MODE>()) void grouped_feed_kernel(GroupLaunch g, unsigned spin_limit) {
    using GF = GroupedFeed<MODE>;
    using G1 = typename GF::template G<64, 128>;
    using G2 = typename GF::template G<64, 256>;
    using G3 = typename GF::template G<128, 128>;
    using G4 = typename GF::template G<128, 256>;
    static_assert(G1::CT == G4::CT && G1::PT == G4::PT && G2::CT == G4::CT && G2::PT == G4::PT &&
                      G3::CT == G4::CT && G3::PT == G4::PT,
                  "one workgroup shape");
    static_assert(G4::RING >= G1::RING && G4::RING >= G2::RING && G4::RING >= G3::RING, "ring of the largest");
    static_assert(kGroupedMaxCfg == 4 && grouped_min_cfg(MODE) == 1, "geometries 1..4");
    __shared__ __attribute__((aligned(16))) char ring[G4::RING];
    __shared__ FeedShared<G4::kNS> sh;
    __shared__ int s_item;
    if (threadIdx.x < G4::kNS) {
        sh.ready[threadIdx.x] = 0u;
        sh.freed[threadIdx.x] = 0u;
    }
    if (threadIdx.x == 0) sh.abort_flag = 0u;
    long long jb = 0;
    // Every wave leaves when the list is exhausted; an aborted handshake only
    // turns the remaining items' waits into immediate failures (NaN tiles).
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(g.counter, 1u);
        __syncthreads();
        const int item = s_item;
        __syncthreads();  // every wave has read s_item before it is overwritten
        if (item >= g.total_items) break;
        const GroupProblem& q = g.problems[g.item_problem[item]];
        const int local = item - q.item_begin;
        switch (q.cfg) {
            case 1: jb += grouped_feed_item<G1>(q, local, ring, sh, jb, spin_limit); break;
            case 2: jb += grouped_feed_item<G2>(q, local, ring, sh, jb, spin_limit); break;
            case 3: jb += grouped_feed_item<G3>(q, local, ring, sh, jb, spin_limit); break;
            default: jb += grouped_feed_item<G4>(q, local, ring, sh, jb, spin_limit); break;
        }
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
// tile on one MI355X: per-call times of chosen shapes, and every 64- and
// 128-tile candidate of 1200 shapes of the selector's table (M, N, K
// log-uniform in 1..8192, all three modes; nearly all ragged) timed by the
// autotuner under both feeds.
//   3 split6 64 x 64, BK 64, NS 2, PFR 2: one workgroup per CU and K slices of
//     at least 256: median 1.21x (p10 1.02x) over such candidates, 1.23x
//     from 512 (768 x 768 x 4096 (64,1) 76.5 -> 63.4 us, 700 x 900 x 3000
//     110.2 -> 55.9 us). Shorter slices lose (median 0.5-0.6x below 128).
//   1 split3 64 x 64, BK 128, NS 2, PFR 2, and 2 split3 64 x 64, BK 128, NS 2,
//     PFR 1: one workgroup per CU and slices of at least 1024: median 1.02x
//     on the table's shapes, 1.05-1.33x on grids of 64..256 workgroups
//     (1000^3 (64,1) 23.6 -> 17.5 us, 768 x 768 x 4096 48.3 -> 40.0 us).
//     Variant 1 (two blocks of loads in flight) for slices of 4096 or more.
//   7 split3 128 x 128, BK 64, NS 2, PFR 1, eight consumer and eight producer
//     waves: one workgroup per CU and slices of at least 1024, when every
//     tile and k-block is full (2048^3 (128,1) 63.8 -> 59.1 us, 512 x 4096 x
//     4096 94.9 -> 84.4 us), or when M and N are at least 256 and slices are
//     at least 2048 (1000 x 1000 x 4000 136.4 -> 92.0 us;
//     median 1.12x on the table's shapes). Smaller or thinner tiles lose up
//     to 10%.
//   8 split6 128 x 128, BK 32, NS 2, PFR 2, eight consumer and four producer
//     waves: one workgroup per CU, every tile and k-block full, slices of at
//     least 1024 (1152 x 1152 x 8192 (128,1) 315.7 -> 288.9 us); on the
//     table's ragged shapes it lost 10-25%.
//   4 plain 64 x 64, BK 64, NS 2, PFR 2; 5 split3 BK 64, NS 2, PFR 2; 6 split6
//     BK 32, NS 4, PFR 2: only when the feed is forced. Plain was 10-40%
//     slower on small grids with every variant, and on grids of more than one
//     workgroup per CU all modes were 1-25% slower (the direct tile hides
//     latency with several workgroups per CU instead).
// Producers load a k-block inside K on a tile that crosses M or N with
// clamped vector loads (load_edge_as), and a partial last k-block on its own
// after the others; the direct tile keeps element-wise edge loads. Slices shorter than the k-block depth
// are zero-padded, as in the deep direct variants: results are those of the
// direct tile bit for bit, except that an exact -0 may come out as +0.
int feed_variant(int mode, int cfg, long long M, long long N, long long K, long long grid, long long k_len,
                 int setting) {
    if (setting == 0 || mode < 0 || mode > 2) return 0;
    const bool small_grid = grid > 0 && grid <= kDeepMaxGrid;
    if (cfg == 3) {
        // 128 x 128 (split modes); plain keeps the direct tile.
        if (mode == 0) return 0;
        if (setting == 1) return mode == 1 ? 7 : 8;
        const bool whole = M % 128 == 0 && N % 128 == 0 && K % 64 == 0 && k_len % 64 == 0;
        if (!small_grid || k_len < 1024) return 0;
        if (mode == 1) return (whole || (M >= 256 && N >= 256 && k_len >= 2048)) ? 7 : 0;
        return whole ? 8 : 0;
    }
    if (cfg != 0) return 0;
    if (setting == 2) {  // measured rule
        if (!small_grid) return 0;
        if (mode == 1) return k_len >= 1024 ? (k_len >= 4096 ? 1 : 2) : 0;
        return (mode == 2 && k_len >= 256) ? 3 : 0;
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

bool grouped_feed(int mode, int setting) { return mode != 0 && setting != 0; }

int grouped_feed_threads(int mode) {
    return mode == 1 ? grouped_feed_nt<1>() : (mode == 2 ? grouped_feed_nt<2>() : 0);
}

hipError_t launch_grouped_feed(int mode, const GroupLaunch& g, int grid, hipStream_t stream) {
    if (grid <= 0) return hipErrorInvalidValue;
    switch (mode) {
        case 1:
            hipLaunchKernelGGL(grouped_feed_kernel<1>, dim3(grid), dim3(grouped_feed_nt<1>()), 0, stream, g,
                               spin_limit());
            break;
        case 2:
            hipLaunchKernelGGL(grouped_feed_kernel<2>, dim3(grid), dim3(grouped_feed_nt<2>()), 0, stream, g,
                               spin_limit());
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int grouped_feed_blocks_per_cu(int mode) {
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    if (mode == 1)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_feed_kernel<1>, grouped_feed_nt<1>(), 0);
    else if (mode == 2)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_feed_kernel<2>, grouped_feed_nt<2>(), 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}

}  // namespace mmsk
