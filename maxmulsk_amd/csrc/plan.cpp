#include "plan.hpp"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "kernels.hpp"

namespace mmsk {

maxmulsk_plan candidate(int i) {
    if (i < 0 || i >= kNumCandidates) return maxmulsk_plan{0, 0};
    return maxmulsk_plan{kCandTiles[i / kNumSplitCands], kCandSplits[i % kNumSplitCands]};
}

int candidate_index(maxmulsk_plan p) {
    for (int i = 0; i < kNumCandidates; ++i) {
        const maxmulsk_plan c = candidate(i);
        if (c.tile == p.tile && c.split_k == p.split_k) return i;
    }
    return -1;
}

int cfg_for(int tile, size_t M) {
    switch (tile) {
        case 64: return 0;
        case 128: return M <= 64 ? 1 : 3;
        case 256: return M <= 64 ? 2 : (M <= 128 ? 4 : 5);
        default: return -1;
    }
}

bool packed_path(int mode, size_t M, size_t N, size_t K) {
    if (mode == 0) return true;
    return !(M % 256 == 0 && N % 256 == 0 && K % (size_t)packed_bk(mode) == 0);
}

bool resolve(size_t M, size_t K, size_t N, int mode, maxmulsk_plan p, Resolved* out, int max_cfg, int min_cfg) {
    if (mode < 0 || mode > 2) return false;
    if (M == 0 || K == 0 || N == 0) return false;
    // Grid and index arithmetic inside the kernels is 32-bit for tile counts.
    if (M > 0x7fffffffULL || N > 0x7fffffffULL || K > 0x7fffffffULL) return false;
    int cfg = cfg_for(p.tile, M);
    if (cfg < 0) return false;
    if (cfg > max_cfg) cfg = max_cfg;
    if (cfg < min_cfg) cfg = min_cfg;
    if (p.split_k < 1 || p.split_k > kMaxSplit) return false;
    const TileCfg& tc = kTileCfgs[cfg];
    Resolved r{};
    r.cfg = cfg;
    r.bm = tc.bm;
    r.bn = tc.bn;
    r.packed = cfg == kPackedCfg && packed_path(mode, M, N, K);
    r.bk = cfg == kPackedCfg ? packed_bk(mode) : tc.bk[mode];
    const long long kblocks = ((long long)K + r.bk - 1) / r.bk;
    long long per = (kblocks + p.split_k - 1) / p.split_k;  // k-blocks per slice
    if (per < 1) per = 1;
    r.k_len = per * r.bk;
    r.splits = (int)((kblocks + per - 1) / per);
    r.tiles_m = (int)((M + r.bm - 1) / r.bm);
    r.tiles_n = (int)((N + r.bn - 1) / r.bn);
    const long long grid = (long long)r.tiles_m * r.tiles_n * r.splits;
    if (grid > 0x7fffffffLL) return false;
    auto up = [](unsigned long long b) { return (size_t)((b + 255) / 256 * 256); };
    r.partial_bytes = r.splits > 1 ? up((unsigned long long)r.splits * M * N * sizeof(float)) : 0;
    r.kblocks = kblocks;
    r.kgroups = 0;
    r.pack_a_bytes = r.pack_b_bytes = 0;
    if (r.packed) {
        const int terms = mode == 0 ? 1 : (mode == 1 ? 2 : 3);
        r.kgroups = (int)(kblocks * (packed_bk(mode) / packed_kgk(mode)));
        const unsigned long long frag = 1024ULL * (unsigned long long)r.kgroups * terms;
        r.pack_a_bytes = up(frag * (unsigned long long)r.tiles_m * (r.bm / 32));
        r.pack_b_bytes = up(frag * (unsigned long long)r.tiles_n * (r.bn / 32));
    }
    r.ws_bytes = r.partial_bytes + r.pack_a_bytes + r.pack_b_bytes;
    *out = r;
    return true;
}

maxmulsk_plan rule_plan(size_t M, size_t K, size_t N, int mode) {
    (void)mode;
    auto tiles = [&](int t) {
        const int cfg = cfg_for(t, M);
        const TileCfg& tc = kTileCfgs[cfg];
        return (long long)((M + tc.bm - 1) / tc.bm) * (long long)((N + tc.bn - 1) / tc.bn);
    };
    int tile = 64;
    for (int t : {256, 128}) {
        if (tiles(t) >= kNumCUs / 2) { tile = t; break; }
    }
    const long long nt = tiles(tile);
    int split = 1;
    while (split < kMaxSplit && nt * split < kNumCUs && (long long)K / (split * 2) >= 256) split *= 2;
    return maxmulsk_plan{tile, split};
}

bool env_plan(maxmulsk_plan* out) {
    const char* s = std::getenv("MAXMULSK_PLAN");
    if (!s || !*s) return false;
    int t = 0, k = 0;
    char tail = 0;
    if (std::sscanf(s, "%d,%d%c", &t, &k, &tail) != 2) return false;
    if (cfg_for(t, 1) < 0 || k < 1 || k > kMaxSplit) return false;
    *out = maxmulsk_plan{t, k};
    return true;
}

int env_mode(int dflt) {
    const char* s = std::getenv("MAXMULSK_MODE");
    if (!s) return dflt;
    if (!std::strcmp(s, "plain")) return 0;
    if (!std::strcmp(s, "split3")) return 1;
    if (!std::strcmp(s, "split6")) return 2;
    return dflt;
}

// ---- autotuner cache ----------------------------------------------------------

namespace {
using Key = std::tuple<unsigned long long, unsigned long long, unsigned long long, int>;
std::mutex g_cache_mu;
std::map<Key, maxmulsk_plan> g_cache;

constexpr char kMagic[8] = {'M', 'M', 'S', 'K', 'T', 'U', 'N', 'E'};
constexpr uint32_t kVersion = 1;
constexpr size_t kNameBytes = 64;

void put_u64(FILE* f, uint64_t v) {
    unsigned char b[8];
    for (int i = 0; i < 8; ++i) b[i] = (unsigned char)(v >> (8 * i));
    std::fwrite(b, 1, 8, f);
}
bool get_u64(FILE* f, uint64_t* v) {
    unsigned char b[8];
    if (std::fread(b, 1, 8, f) != 8) return false;
    uint64_t x = 0;
    for (int i = 0; i < 8; ++i) x |= (uint64_t)b[i] << (8 * i);
    *v = x;
    return true;
}
}  // namespace

bool cache_get(size_t M, size_t K, size_t N, int mode, maxmulsk_plan* out) {
    std::lock_guard<std::mutex> g(g_cache_mu);
    auto it = g_cache.find(Key{M, K, N, mode});
    if (it == g_cache.end()) return false;
    *out = it->second;
    return true;
}

void cache_put(size_t M, size_t K, size_t N, int mode, maxmulsk_plan p) {
    std::lock_guard<std::mutex> g(g_cache_mu);
    g_cache[Key{M, K, N, mode}] = p;
}

void cache_clear() {
    std::lock_guard<std::mutex> g(g_cache_mu);
    g_cache.clear();
}

size_t cache_size() {
    std::lock_guard<std::mutex> g(g_cache_mu);
    return g_cache.size();
}

// File layout (little endian): magic[8], u64 version, name[64] (device name,
// zero padded), u64 count, then count x {u64 M, u64 K, u64 N, u64 mode,
// u64 tile, u64 split}.
bool cache_save(const char* path, const char* device_name) {
    std::lock_guard<std::mutex> g(g_cache_mu);
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    std::fwrite(kMagic, 1, 8, f);
    put_u64(f, kVersion);
    char name[kNameBytes] = {0};
    std::strncpy(name, device_name ? device_name : "", kNameBytes - 1);
    std::fwrite(name, 1, kNameBytes, f);
    put_u64(f, g_cache.size());
    for (const auto& kv : g_cache) {
        put_u64(f, std::get<0>(kv.first));
        put_u64(f, std::get<1>(kv.first));
        put_u64(f, std::get<2>(kv.first));
        put_u64(f, (uint64_t)std::get<3>(kv.first));
        put_u64(f, (uint64_t)kv.second.tile);
        put_u64(f, (uint64_t)kv.second.split_k);
    }
    const bool ok = std::ferror(f) == 0;
    return std::fclose(f) == 0 && ok;
}

int cache_load(const char* path, const char* device_name) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return 2;
    char magic[8];
    char name[kNameBytes];
    uint64_t ver = 0, count = 0;
    int rc = 0;
    std::map<Key, maxmulsk_plan> loaded;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kMagic, 8) != 0 || !get_u64(f, &ver) ||
        ver != kVersion || std::fread(name, 1, kNameBytes, f) != kNameBytes || !get_u64(f, &count)) {
        rc = 1;
    } else {
        name[kNameBytes - 1] = 0;
        char want[kNameBytes] = {0};
        std::strncpy(want, device_name ? device_name : "", kNameBytes - 1);
        if (std::strcmp(name, want) != 0 || count > (1u << 24)) rc = 1;
        for (uint64_t i = 0; rc == 0 && i < count; ++i) {
            uint64_t v[6];
            for (int j = 0; j < 6; ++j)
                if (!get_u64(f, &v[j])) rc = 1;
            if (rc) break;
            const maxmulsk_plan p{(int)v[4], (int)v[5]};
            if (v[3] > 2 || cfg_for(p.tile, 1) < 0 || p.split_k < 1 || p.split_k > kMaxSplit) {
                rc = 1;
                break;
            }
            loaded[Key{v[0], v[1], v[2], (int)v[3]}] = p;
        }
    }
    std::fclose(f);
    if (rc == 0) {
        std::lock_guard<std::mutex> g(g_cache_mu);
        for (const auto& kv : loaded) g_cache[kv.first] = kv.second;
    }
    return rc;
}

maxmulsk_plan choose_plan(size_t M, size_t K, size_t N, int mode, int* source) {
    maxmulsk_plan p{};
    int src = 3;
    if (env_plan(&p)) {
        src = 0;
    } else if (cache_get(M, K, N, mode, &p)) {
        src = 1;
    } else if (selector_pick(M, K, N, mode, &p)) {
        src = 2;
    } else {
        p = rule_plan(M, K, N, mode);
    }
    if (source) *source = src;
    return p;
}

maxmulsk_plan choose_plan_no_selector(size_t M, size_t K, size_t N, int mode) {
    maxmulsk_plan p{};
    if (env_plan(&p) || cache_get(M, K, N, mode, &p)) return p;
    return rule_plan(M, K, N, mode);
}

}  // namespace mmsk
