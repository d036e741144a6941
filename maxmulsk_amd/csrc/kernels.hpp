// Host-visible kernel launch interface (no HIP device code here).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace mmsk {

// Tile geometries: wgm x wgn waves over the output tile. Geometries 0..4 are
// compiled for every mode as the direct tile kernel (and serve the grouped
// kernel); geometry 5 (256 x 256, eight waves, two per SIMD) runs on the
// packed path, whose k-block depth is packed_bk(mode).
struct TileCfg {
    int bm, bn;
    int wgm, wgn;
    int bk[3];  // tile depth per mode: plain, split3, split6
};

constexpr int kNumTileCfgs = 6;
constexpr TileCfg kTileCfgs[kNumTileCfgs] = {
    {64, 64, 2, 2, {32, 32, 32}},
    {64, 128, 2, 2, {32, 32, 32}},
    {64, 256, 1, 4, {32, 32, 32}},
    {128, 128, 2, 2, {32, 32, 32}},
    {128, 256, 2, 2, {32, 32, 32}},
    {256, 256, 2, 4, {32, 32, 16}},
};

// Deep variant of geometry 0 (64 x 64): k-blocks of kDeepBK[mode] instead of
// 32, launched when the whole grid fits in one workgroup per CU. Such launches
// (small M and N, long K split into slices) run one k-block loop per CU, whose
// time is a chain of per-block steps (load, split, store, barrier, MFMA);
// deeper blocks halve or quarter the number of steps. With more workgroups
// than CUs the larger LDS stages would cost occupancy instead. Slices keep the
// boundaries of the 32-deep resolution, so the products and their order are
// those of the regular kernel; a slice shorter than one deep block only adds
// MFMAs of zero padding, which can turn an exact -0 result into +0. 0 = no
// deep variant for that mode: split6 measured slower with 64-deep blocks (its
// three BF16 terms make the stages large).
constexpr int kDeepBK[3] = {64, 128, 0};
constexpr int kDeepMaxGrid = 256;
// Grids of up to two workgroups per CU (kDeepMaxGrid < grid <= 2 * kDeepMaxGrid)
// run geometry 0 with k-blocks of kDeep2BK[mode]: 64 KiB of LDS per workgroup,
// so the whole grid is resident at once. Measured: 67 x 513 x 8191 (plan 64,16,
// 288 workgroups) plain 41.8 -> 37.0 us and split3 28.9 -> 26.6 us; 1000^3
// plain 35.3 -> 32.1 us. Same slice boundaries and products as above.
constexpr int kDeep2BK[3] = {64, 64, 0};
// Deep variant of geometry 3 (128 x 128, eight waves in the split modes) for
// grids of at most kDeepMaxGrid workgroups: k-blocks of kDeep128BK[mode] (128 KiB
// of LDS in split3). Measured in split3: 2048^3 (plan 128,1) 67.2 -> 60.4 us,
// 2048 x 1024 x 2048 42.5 -> 39.1 us, 512 x 4096 x 4096 69.5 -> 64.2 us. Plain
// measured within +-1% and keeps 32; split6 would need 192 KiB.
constexpr int kDeep128BK[3] = {0, 64, 0};

// Direct launches of geometries 3 (128 x 128) and 4 (128 x 256) in the split
// modes run eight waves (2 x 4) instead of the table's four: twice the waves
// per SIMD to cover LDS and barrier latency. Measured 2-7% faster on 128 x 128
// (2048^3 split3 71.0 -> 67.0 us) and up to 14% on 128 x 256 (96 x 8192 x 8192
// split6 117.6 -> 100.6 us). Plain mode keeps four (eight measured 4-9%
// slower). The wave layout changes no product and no summation order, so
// results are identical to the four-wave tiles.
constexpr int kSplitWideWaves[2] = {2, 4};

// Wave layout (rows x columns of waves) of the direct launch of geometry cfg
// (0..4) in a mode.
constexpr bool direct_wide(int mode, int cfg) { return mode != 0 && (cfg == 3 || cfg == 4); }
constexpr int direct_wgm(int mode, int cfg) { return direct_wide(mode, cfg) ? kSplitWideWaves[0] : kTileCfgs[cfg].wgm; }
constexpr int direct_wgn(int mode, int cfg) { return direct_wide(mode, cfg) ? kSplitWideWaves[1] : kTileCfgs[cfg].wgn; }

// One launch of the tiled kernel: an output grid of tiles_m x tiles_n tiles,
// each computed over `splits` K slices of k_len elements. With splits > 1 every
// slice writes its own FP32 partial matrix at C + split * c_split_stride and a
// reduce kernel sums them in slice order.
struct GemmParams {
    const float* A;
    const float* B;
    float* C;
    long long M, N, K;
    long long lda, ldb, ldc;
    long long k_len;
    long long c_split_stride;
    int splits;
    int tiles_m, tiles_n;
};

hipError_t launch_gemm(int mode, int cfg, const GemmParams& p, hipStream_t stream);

// Decoupled producer/consumer feed of geometries 0 and 3 (feed.hip): producer
// waves stage and split operands into an LDS ring, consumer waves only
// multiply; same K slices, products and order as the direct tile. feed_mode():
// 0 = direct tile kernels only, 1 = every 64 x 64 launch and every split-mode
// 128 x 128 launch takes the feed, 2 = the feed where it measured faster (the
// default). Split-mode grouped launches take the feed under 1 and 2
// (grouped_feed). MAXMULSK_FEED=direct|decoupled sets 0 or 1 at first use;
// set_feed_mode overrides.
int feed_mode();
void set_feed_mode(int v);
// Feed variant of a launch of geometry cfg (0: 64 x 64, 3: 128 x 128) for an
// M x N x K product on `grid` workgroups with K slices of k_len, under a feed
// setting; 0 = the direct tile.
int feed_variant(int mode, int cfg, long long M, long long N, long long K, long long grid, long long k_len,
                 int setting);
hipError_t launch_gemm_feed(int variant, const GemmParams& p, hipStream_t stream);

// ---- packed path (256 x 256 geometry) -------------------------------------------
// k-block depth of the packed GEMM per mode and k values per packed fragment.
// Plain and split3 take 32-deep k-blocks in a two-stage ring (128 KiB of LDS):
// half the barriers of 16-deep blocks in a four-stage ring, measured 1.5-2%
// faster at 4096^3 and 4096 x 24576 x 1536 with bit-identical results; split6
// keeps 16 (a 32-deep stage of its three terms is 96 KiB).
constexpr int packed_bk(int mode) { return mode == 2 ? 16 : 32; }
static_assert(kTileCfgs[5].bk[0] == packed_bk(0) && kTileCfgs[5].bk[1] == packed_bk(1) &&
                  kTileCfgs[5].bk[2] == packed_bk(2),
              "the tile table lists the packed k-block depths");
constexpr int kPackedCfg = 5;
constexpr int packed_kgk(int mode) { return mode == 0 ? 8 : 16; }

struct PackedParams {
    const char* PA;
    const char* PB;
    long long planeA, planeB;  // bytes per BF16 term (or the FP32 copy)
    int kgtA, kgtB;            // fragments per row block along K
    float* C;
    long long M, N, ldc;
    long long c_split_stride;
    long long kblocks;            // k-blocks of packed_bk(mode) in the packed operands
    long long kblocks_per_split;
    int splits, tiles_m, tiles_n;
};
// Packs X (element (row, k) at X[row * ld_row + k * ld_k], rows < R, k < K)
// into rblocks x kgroups fragments per term, zero-padded.
hipError_t launch_pack(int mode, const float* X, long long R, long long K, long long ld_row, long long ld_k,
                       void* dst, int rblocks, int kgroups, hipStream_t stream);
// Two operands of the same K packed by one launch.
struct PackJob {
    const float* X;
    long long R, ld_row, ld_k;
    char* dst;
    int rblocks;
};
hipError_t launch_pack2(int mode, const PackJob& a, const PackJob& b, long long K, int kgroups,
                        hipStream_t stream);
hipError_t launch_gemm_packed(int mode, const PackedParams& p, hipStream_t stream);

// C[i] = sum_{s < splits} P[s * stride + i] for i < M*N, summed in s order.
hipError_t launch_splitk_reduce(const float* P, float* C, long long mn, int splits,
                                long long stride, hipStream_t stream);

// ---- grouped persistent launch ------------------------------------------------

// One problem of a grouped launch as the device sees it.
struct GroupProblem {
    const float* A;
    const float* B;
    float* C;      // final output
    float* P;      // split-K partials, splits x M x N (nullptr when splits == 1)
    long long M, N, K;
    long long k_len;
    int splits;
    int cfg;       // tile geometry index into kTileCfgs
    int tiles_m, tiles_n;
    int item_begin;   // first work item of this problem in the flattened list
    int reduce_begin; // first reduce block of this problem (split problems only)
};

// Persistent kernel over the flattened (problem, k-slice, tile) work list. Each
// workgroup claims items from *counter (must be zero on entry) until the list
// is exhausted; item_problem maps an item to its problem.
struct GroupLaunch {
    const GroupProblem* problems;
    const int* item_problem;
    int total_items;
    unsigned int* counter;
};
hipError_t launch_grouped(int mode, const GroupLaunch& g, int grid, hipStream_t stream);
// Plain-mode kernel of launch_grouped (its own translation unit) and its occupancy.
hipError_t launch_grouped_plain(const GroupLaunch& g, int grid, hipStream_t stream);
int grouped_plain_blocks_per_cu();
// Largest tile geometry the persistent kernel serves.
constexpr int kGroupedMaxCfg = 4;
// The split-mode persistent kernels run 512-thread workgroups and every tile
// with eight waves (2 x 4): one workgroup fits per CU (LDS), so eight waves
// give two per SIMD where four gave one (measured 10% faster in split3, 6% in
// split6, on the 256-problem batch). Geometry 0 (64 x 64) cannot hold eight
// 32 x 32 wave tiles, so these kernels serve geometries 1..4 and a 64-tile
// plan runs on 64 x 128 tiles (same K slices, same products).
constexpr int grouped_min_cfg(int mode) { return mode == 0 ? 0 : 1; }
constexpr int grouped_threads(int mode) { return mode == 0 ? 256 : 512; }
// Workgroups of the persistent kernel that fit on one CU at once.
int grouped_blocks_per_cu(int mode);
// The split modes' persistent kernel on the decoupled feed (feed.hip): the
// same work items and geometries (1..4) as grouped_kernel, with producer and
// consumer waves per workgroup. grouped_feed tells whether a launch of `mode`
// built under feed setting `setting` takes it: every split-mode launch unless
// the setting is 0 (direct).
bool grouped_feed(int mode, int setting);
int grouped_feed_threads(int mode);
hipError_t launch_grouped_feed(int mode, const GroupLaunch& g, int grid, hipStream_t stream);
int grouped_feed_blocks_per_cu(int mode);
// Split-K reduce of every split problem: reduce block b serves problem
// block_problem[b], elements [(b - reduce_begin) * 1024, +1024) of its M x N.
hipError_t launch_grouped_reduce(const GroupProblem* problems, const int* block_problem,
                                 int total_blocks, hipStream_t stream);
// Zero-fill (graph-safe replacement for hipMemsetAsync).
hipError_t launch_fill_zero_u32(unsigned int* p, size_t n, hipStream_t stream);
hipError_t launch_fill_zero_f32(float* p, size_t n, hipStream_t stream);
// dst[i] = src[i * stride] for i < count.
hipError_t launch_gather_strided(const float* src, size_t stride, size_t count, float* dst,
                                 hipStream_t stream);
// FP64 reference of the same strided sample of C = A*B (A M x K, B K x N,
// row-major): for i < count, flat index i * stride = r * N + c,
// ref[i] = sum_k A[r][k] B[k][c] and absref[i] = sum_k |A[r][k]| |B[k][c]|,
// accumulated in FP64 (the autotuner's correctness gate).
hipError_t launch_sample_dot_f64(const float* A, const float* B, size_t K, size_t N, size_t stride,
                                 size_t count, double* ref, double* absref, hipStream_t stream);

// ---- MLP (learned selector) ---------------------------------------------------
constexpr int kMlpIn = 4;
constexpr int kMlpHidden = 64;
int mlp_num_params(int nout);
hipError_t launch_mlp_forward(const float* params, const float* X, int batch, int nout,
                              float* H1, float* H2, float* Y, hipStream_t stream);
// scratch: device floats, at least 64 + 2 * batch * kMlpHidden.
hipError_t launch_mlp_backward(const float* params, const float* X, int batch, int nout,
                               const float* H1, const float* H2, const float* Y,
                               const float* T, const float* Mask, float* dparams,
                               float* loss, hipStream_t stream, float* scratch);
hipError_t launch_mlp_adam(float* params, const float* g, float* m, float* v, int n, int step,
                           float lr, float b1, float b2, float eps, hipStream_t stream);
hipError_t launch_mlp_infer(const float* params, const float* x, int nout, float* y,
                            hipStream_t stream);

}  // namespace mmsk
