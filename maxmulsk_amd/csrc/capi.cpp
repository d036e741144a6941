// extern "C" entry points of libmaxmulsk (declared in include/maxmulsk.h).
#include <hip/hip_runtime.h>

#include <cstring>

#include "kernels.hpp"
#include "lds_layout.hpp"
#include "maxmulsk.h"
#include "plan.hpp"
#include "runtime.hpp"
#include "trace.hpp"

using namespace mmsk;

namespace {

GemmCall base_call(const float* A, const float* B, float* C, size_t M, size_t K, size_t N) {
    GemmCall c{};
    c.A = A;
    c.B = B;
    c.C = C;
    c.M = M;
    c.K = K;
    c.N = N;
    c.mode = 0;
    c.stream = nullptr;
    return c;
}

}  // namespace

extern "C" maxmulsk_status maxmulsk_sgemm_v3(const float* A, const float* B, float* C, size_t M,
                                             size_t K, size_t N) {
    GemmCall c = base_call(A, B, C, M, K, N);
    c.mode = env_mode(0);
    c.rule_only = true;
    return run_gemm(c);
}

extern "C" maxmulsk_status maxmulsk_sgemm_v4(const float* A, const float* B, float* C, size_t M,
                                             size_t K, size_t N) {
    GemmCall c = base_call(A, B, C, M, K, N);
    c.mode = env_mode(0);
    return run_gemm(c);
}

extern "C" maxmulsk_status maxmulsk_sgemm_ex(const float* A, const float* B, float* C, size_t M,
                                             size_t K, size_t N, const maxmulsk_options* opt) {
    GemmCall c = base_call(A, B, C, M, K, N);
    if (opt) {
        c.mode = (int)opt->mode;
        if (opt->plan.tile != 0 || opt->plan.split_k != 0) {
            c.have_plan = true;
            c.plan = opt->plan;
        }
        c.stream = static_cast<hipStream_t>(opt->stream);
        c.async = opt->async != 0;
        c.ws = opt->workspace;
        c.ws_bytes = opt->workspace_bytes;
    }
    return run_gemm(c);
}

extern "C" size_t maxmulsk_workspace_bytes(size_t M, size_t K, size_t N, maxmulsk_mode mode,
                                           const maxmulsk_plan* plan) {
    if (M == 0 || K == 0 || N == 0) return 0;
    const maxmulsk_plan p = plan ? *plan : choose_plan(M, K, N, (int)mode, nullptr);
    Resolved r;
    if (!resolve(M, K, N, (int)mode, p, &r)) return 0;
    return r.ws_bytes;
}

extern "C" maxmulsk_status maxmulsk_sgemm_ws(const float* A, const float* B, float* C, size_t M,
                                             size_t K, size_t N, maxmulsk_mode mode,
                                             const maxmulsk_plan* plan, void* ws, size_t ws_bytes,
                                             void* stream) {
    GemmCall c = base_call(A, B, C, M, K, N);
    c.mode = (int)mode;
    if (plan) {
        c.have_plan = true;
        c.plan = *plan;
    }
    c.stream = static_cast<hipStream_t>(stream);
    c.ws = ws;
    c.ws_bytes = ws_bytes;
    c.ws_required = true;
    return run_gemm(c);
}

extern "C" maxmulsk_status maxmulsk_choose_plan(size_t M, size_t K, size_t N, maxmulsk_mode mode,
                                                maxmulsk_plan* out, int* source) {
    if (!out) return fail(MAXMULSK_ERR_UNSUPPORTED, "out is null");
    if ((int)mode < 0 || (int)mode > 2) return fail(MAXMULSK_ERR_UNSUPPORTED, "unknown mode");
    if (M == 0 || K == 0 || N == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "empty shape has no plan");
    *out = choose_plan(M, K, N, (int)mode, source);
    return MAXMULSK_OK;
}

extern "C" maxmulsk_plan maxmulsk_rule_plan(size_t M, size_t K, size_t N, maxmulsk_mode mode) {
    return rule_plan(M, K, N, (int)mode);
}

extern "C" int maxmulsk_num_candidates(void) { return kNumCandidates; }

extern "C" maxmulsk_plan maxmulsk_candidate(int i) { return candidate(i); }

extern "C" maxmulsk_status maxmulsk_cache_put(size_t M, size_t K, size_t N, maxmulsk_mode mode,
                                              maxmulsk_plan plan) {
    if ((int)mode < 0 || (int)mode > 2 || cfg_for(plan.tile, 1) < 0 || plan.split_k < 1 ||
        plan.split_k > kMaxSplit)
        return fail(MAXMULSK_ERR_UNSUPPORTED, "invalid cache entry");
    cache_put(M, K, N, (int)mode, plan);
    return MAXMULSK_OK;
}

extern "C" void maxmulsk_cache_clear(void) { cache_clear(); }

extern "C" size_t maxmulsk_cache_size(void) { return cache_size(); }

static std::string cache_device_name() {
    if (usable_devices() == 0) return "none";
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) {
        (void)hipGetLastError();
        d = 0;
    }
    return device_name(d);
}

extern "C" maxmulsk_status maxmulsk_cache_save(const char* path) {
    if (!path) return fail(MAXMULSK_ERR_UNSUPPORTED, "path is null");
    if (!cache_save(path, cache_device_name().c_str()))
        return fail(MAXMULSK_ERR_DEVICE, "cannot write %s", path);
    return MAXMULSK_OK;
}

extern "C" maxmulsk_status maxmulsk_cache_load(const char* path) {
    if (!path) return fail(MAXMULSK_ERR_UNSUPPORTED, "path is null");
    const int rc = cache_load(path, cache_device_name().c_str());
    if (rc == 1) return fail(MAXMULSK_ERR_UNSUPPORTED, "%s: wrong format, version or device", path);
    if (rc == 2) return fail(MAXMULSK_ERR_DEVICE, "cannot read %s", path);
    return MAXMULSK_OK;
}

extern "C" int maxmulsk_device_count(void) { return usable_devices(); }

extern "C" const char* maxmulsk_last_error(void) { return last_error(); }

extern "C" int maxmulsk_lds_chunk_offset(int mode, int op, int rows, int bk, int row, int c,
                                         int term) {
    if (mode < 0 || mode > 2 || (op != 0 && op != 1) || rows <= 0 || rows % 32 != 0 || bk <= 0)
        return -1;
    const int lu = mode_lu(mode);
    if (bk % (2 * lu) != 0) return -1;
    if (row < 0 || row >= rows || c < 0 || c >= bk / lu || term < 0 || term >= mode_terms(mode))
        return -1;
    return chunk_offset(mode, op, rows, bk, row, c, term);
}

extern "C" int maxmulsk_lds_epilogue_write(int wtn, int lane, int tj, int q) {
    if (wtn != 64 || lane < 0 || lane >= 64 || tj < 0 || tj >= wtn / 32 || q < 0 || q >= 4)
        return -1;
    return epi_write_offset(wtn, lane, tj, q);
}

extern "C" int maxmulsk_lds_epilogue_read(int wtn, int lane, int i, int* row, int* col) {
    if (wtn != 64 || lane < 0 || lane >= 64 || i < 0 || i >= wtn / 8 || !row || !col)
        return -1;
    *row = epi_read_row(wtn, lane, i);
    *col = epi_read_col(wtn, lane);
    return *row * epi_pitch(wtn) + *col;
}

extern "C" int maxmulsk_num_tile_geometries(void) { return kNumTileCfgs; }

extern "C" int maxmulsk_tile_info(int cfg, int mode, int* out) {
    if (cfg < 0 || cfg >= kNumTileCfgs || mode < 0 || mode > 2 || !out) return -1;
    const TileCfg& t = kTileCfgs[cfg];
    const int waves = cfg < 5 ? direct_wgm(mode, cfg) * direct_wgn(mode, cfg) : t.wgm * t.wgn;
    out[0] = t.bm;
    out[1] = t.bn;
    out[2] = cfg == kPackedCfg ? packed_bk(mode) : t.bk[mode];
    out[3] = 64 * waves;
    out[4] = b_unit_cols(mode, t.bn, t.bk[mode], waves);  // B staging unit width of the tile kernel
    out[5] = waves;
    return 0;
}

extern "C" int maxmulsk_num_phase_counters(void) { return kNumPhaseCounters; }

extern "C" const char* maxmulsk_phase_counter_name(int i) { return counter_name(i); }

extern "C" unsigned long long maxmulsk_phase_counter(int i) { return (unsigned long long)counter(i); }

extern "C" void maxmulsk_phase_counters_reset(void) { reset_counters(); }

extern "C" int maxmulsk_set_feed(int feed) {
    if (feed < MAXMULSK_FEED_DIRECT || feed > MAXMULSK_FEED_AUTO) return -1;
    const int prev = feed_mode();
    set_feed_mode(feed);
    return prev;
}

extern "C" int maxmulsk_get_feed(void) { return feed_mode(); }

extern "C" int maxmulsk_feed_variant(size_t M, size_t K, size_t N, maxmulsk_mode mode, const maxmulsk_plan* plan) {
    if (!plan) return -1;
    Resolved r{};
    if (!resolve(M, K, N, (int)mode, *plan, &r)) return -1;
    return feed_variant((int)mode, r.cfg, (long long)M, (long long)N, (long long)K,
                        (long long)r.tiles_m * r.tiles_n * r.splits, r.k_len, feed_mode());
}
