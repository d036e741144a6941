// Launch planning: plan validation, tile geometry, split-K resolution and the
// plan tiers (environment override -> autotuner cache -> learned selector ->
// shape rule). Pure host C++; no HIP calls except the selector tier.
#pragma once

#include <cstddef>

#include "maxmulsk.h"

namespace mmsk {

constexpr int kCandTiles[3] = {64, 128, 256};
constexpr int kCandSplits[6] = {1, 2, 4, 8, 16, 32};
constexpr int kNumSplitCands = 6;
constexpr int kNumCandidates = 3 * kNumSplitCands;  // index = tile_index * 6 + split_index
constexpr int kMaxSplit = 32;
constexpr int kNumCUs = 256;        // MI355X compute units

// Candidate i of the fixed candidate set.
maxmulsk_plan candidate(int i);
// Index of a plan in the candidate set, -1 when it is not a candidate.
int candidate_index(maxmulsk_plan p);

// Tile geometry index (into kTileCfgs) of tile edge `tile` for M rows: the N
// edge is the tile edge; the M edge shrinks to 64 or 128 for narrow M.
int cfg_for(int tile, size_t M);

// Whether a call on the 256 x 256 geometry packs its operands first (the
// packed GEMM) or splits them on the fly in the direct tile kernel. The split
// modes take the direct tile when every tile and k-block is whole (M and N
// multiples of 256, K a multiple of the k-block depth): with the pack pass
// gone it measured 1.04-1.19x faster in split3 and 1.06-1.19x in split6 on
// such shapes (4096^3 split3 367.7 -> 345.8 us, split6 640.3 -> 584.7 us),
// bit-identical. On ragged shapes its element-wise edge loads lose (split3
// 0.76-1.08x, split6 0.94-1.07x), and plain mode is 1-7% slower direct, so
// those keep the pack.
bool packed_path(int mode, size_t M, size_t N, size_t K);

// A plan resolved against a shape: geometry, K slicing, grid and workspace.
struct Resolved {
    int cfg;
    int bm, bn, bk;
    int splits;          // K slices actually used (<= plan.split_k)
    long long k_len;     // K elements per slice, a multiple of bk
    int tiles_m, tiles_n;
    size_t ws_bytes;     // split-K partials + packed operands
    // Packed path (256 x 256 geometry unless packed_path says otherwise):
    // operands rewritten once into MFMA fragments in the workspace, after the
    // partials.
    bool packed;
    long long kblocks;   // packed k-blocks (packed_bk(mode) deep)
    int kgroups;         // fragments per row block along K
    size_t partial_bytes, pack_a_bytes, pack_b_bytes;
};

// Counterpart of classify (sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:410-423)
// and the clamps of run_multiplication (:638-654): rejects a plan whose tile is
// not 64/128/256 or whose split is outside 1..kMaxSplit, then clamps the split to the
// shape (no empty slices). Shapes must be non-empty.
// max_cfg caps the geometry (the grouped kernel serves geometries 0..4);
// min_cfg raises it (the split-mode grouped kernel serves geometries 1..4).
bool resolve(size_t M, size_t K, size_t N, int mode, maxmulsk_plan p, Resolved* out,
             int max_cfg = 5, int min_cfg = 0);

// Shape rule (tier 3): the largest tile that still gives at least half a wave
// of workgroups, then split-K until the grid covers the CUs, keeping slices at
// least 256 deep.
maxmulsk_plan rule_plan(size_t M, size_t K, size_t N, int mode);

// MAXMULSK_PLAN="tile,split" (the analogue of MAXMULSK_BLOCKING,
// sme/support/gemm_tuning.cpp:66-77). Malformed values are ignored.
bool env_plan(maxmulsk_plan* out);
// MAXMULSK_MODE=plain|split3|split6 for the two reference entry points;
// anything else leaves `dflt`.
int env_mode(int dflt);

// Autotuner cache.
bool cache_get(size_t M, size_t K, size_t N, int mode, maxmulsk_plan* out);
void cache_put(size_t M, size_t K, size_t N, int mode, maxmulsk_plan p);
void cache_clear();
size_t cache_size();
bool cache_save(const char* path, const char* device_name);
int cache_load(const char* path, const char* device_name);  // 0 ok, 1 unsupported, 2 io

// Learned selector tier (defined with the MLP code).
bool selector_pick(size_t M, size_t K, size_t N, int mode, maxmulsk_plan* out);

// Full planner: tiers 0..3. source receives the tier used.
maxmulsk_plan choose_plan(size_t M, size_t K, size_t N, int mode, int* source);
// Tiers 0, 1 and 3 (override, autotuner cache, shape rule): pure host code.
maxmulsk_plan choose_plan_no_selector(size_t M, size_t K, size_t N, int mode);

}  // namespace mmsk
