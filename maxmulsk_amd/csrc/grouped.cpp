// Host side of the grouped persistent launch: validate-all-first, per-problem
// plans, the flattened work list (most expensive items first), workspace for
// split-K partials, hipGraph capture, and the grouped autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <vector>

#include "kernels.hpp"
#include "maxmulsk.h"
#include "plan.hpp"
#include "runtime.hpp"
#include "trace.hpp"

using namespace mmsk;

struct maxmulsk_grouped_plan_s {
    int dev = -1;
    int mode = 0;
    size_t n = 0;
    int total_items = 0;
    int reduce_blocks = 0;
    int grid = 0;
    bool feed = false;  // split modes on the decoupled feed (grouped_feed_kernel)
    double flops = 0;
    std::vector<std::pair<float*, size_t>> zero_fills;  // K == 0 problems
    GroupProblem* d_problems = nullptr;
    int* d_item_problem = nullptr;
    int* d_block_problem = nullptr;
    unsigned int* d_counter = nullptr;
    void* d_ws = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Grouped cache: per shape class (M bucket, N, K, mode).
using ClassKey = std::tuple<int, unsigned long long, unsigned long long, int>;
std::mutex g_gmu;
std::map<ClassKey, maxmulsk_plan> g_group_cache;

int m_bucket(size_t M) { return M <= 64 ? 64 : (M <= 128 ? 128 : (M <= 256 ? 256 : 512)); }
ClassKey class_of(size_t M, size_t K, size_t N, int mode) { return ClassKey{m_bucket(M), N, K, mode}; }

maxmulsk_plan grouped_rule(size_t M, size_t K, size_t N) {
    (void)M;
    const int tile = N >= 256 ? 256 : (N >= 128 ? 128 : 64);
    int split = (int)((K + 4095) / 4096);
    if (split > kMaxSplit) split = kMaxSplit;
    if (split < 1) split = 1;
    return maxmulsk_plan{tile, split};
}

bool grouped_cached(size_t M, size_t K, size_t N, int mode, maxmulsk_plan* out) {
    std::lock_guard<std::mutex> g(g_gmu);
    auto it = g_group_cache.find(class_of(M, K, N, mode));
    if (it == g_group_cache.end()) return false;
    *out = it->second;
    return true;
}

maxmulsk_plan grouped_default(size_t M, size_t K, size_t N, int mode) {
    maxmulsk_plan p;
    return grouped_cached(M, K, N, mode, &p) ? p : grouped_rule(M, K, N);
}

// Default plans of a batch: the grouped cache's plan of the problem's class,
// else the rule. A launch whose tiles (one K slice each) number at least
// kBusyTiles keeps every CU busy on its own, and the longest-first order
// balances it, so slicing K there only adds partial sums and a reduce: the
// rule's problems run unsliced unless one tile of theirs is longer than the
// batch's work per CU. On the 256-problem batch (M ~ U{1..512}) unsliced
// tiles were 5% faster in split3 (228.4 -> 240.3 TFLOP/s), 4% in split6 and
// 2% in plain. Smaller launches keep the rule's slices. Entries of problems
// with M, N or K = 0 are unused.
constexpr double kBusyTiles = 4.0 * kNumCUs;

std::vector<maxmulsk_plan> batch_defaults(const maxmulsk_gemm* problems, size_t n, int mode, bool use_cache) {
    std::vector<maxmulsk_plan> out(n, maxmulsk_plan{0, 0});
    std::vector<double> tile_work(n, 0.0);
    std::vector<bool> ruled(n, false);
    double tiles = 0.0, work = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const maxmulsk_gemm& g = problems[i];
        if (g.M == 0 || g.N == 0 || g.K == 0) continue;
        if (!use_cache || !grouped_cached(g.M, g.K, g.N, mode, &out[i])) {
            out[i] = grouped_rule(g.M, g.K, g.N);
            ruled[i] = true;
        }
        Resolved r;
        if (!resolve(g.M, g.K, g.N, mode, maxmulsk_plan{out[i].tile, 1}, &r, kGroupedMaxCfg, grouped_min_cfg(mode)))
            continue;  // build() reports the problem
        const double t = (double)r.tiles_m * r.tiles_n;
        tile_work[i] = (double)r.bm * r.bn * (double)g.K;
        tiles += t;
        work += t * tile_work[i];
    }
    if (tiles >= kBusyTiles)
        for (size_t i = 0; i < n; ++i)
            if (ruled[i] && tile_work[i] > 0 && tile_work[i] <= work / kNumCUs) out[i].split_k = 1;
    return out;
}

void free_plan(maxmulsk_grouped_plan_s* p) {
    if (!p) return;
    if (p->dev >= 0) {
        DeviceGuard g(p->dev);
        if (p->exec) (void)hipGraphExecDestroy(p->exec);
        if (p->graph) (void)hipGraphDestroy(p->graph);
        (void)hipFree(p->d_problems);
        (void)hipFree(p->d_item_problem);
        (void)hipFree(p->d_block_problem);
        (void)hipFree(p->d_counter);
        (void)hipFree(p->d_ws);
    }
    delete p;
}

hipError_t enqueue(const maxmulsk_grouped_plan_s* p, hipStream_t s) {
    hipError_t e = launch_fill_zero_u32(p->d_counter, 1, s);
    for (const auto& z : p->zero_fills)
        if (e == hipSuccess) e = launch_fill_zero_f32(z.first, z.second, s);
    if (e == hipSuccess && p->total_items > 0) {
        GroupLaunch g{p->d_problems, p->d_item_problem, p->total_items, p->d_counter};
        e = p->feed ? launch_grouped_feed(p->mode, g, p->grid, s) : launch_grouped(p->mode, g, p->grid, s);
    }
    if (e == hipSuccess && p->reduce_blocks > 0)
        e = launch_grouped_reduce(p->d_problems, p->d_block_problem, p->reduce_blocks, s);
    return e;
}

template <class T>
bool upload(T** dst, const std::vector<T>& v) {
    if (v.empty()) return true;
    if (device_malloc(reinterpret_cast<void**>(dst), v.size() * sizeof(T)) != hipSuccess) return false;
    return hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

// Builds a grouped plan. capture: record the launch sequence into a graph.
maxmulsk_status build(const maxmulsk_gemm* problems, size_t n, int mode, const maxmulsk_plan* plans,
                      hipStream_t stream, bool capture, maxmulsk_grouped_plan_s** out) {
    if (!out) return fail(MAXMULSK_ERR_UNSUPPORTED, "out is null");
    *out = nullptr;
    if (n > 0 && !problems) return fail(MAXMULSK_ERR_UNSUPPORTED, "problems is null");
    if (mode < 0 || mode > 2) return fail(MAXMULSK_ERR_UNSUPPORTED, "unknown mode");
    if (n > (1u << 20)) return fail(MAXMULSK_ERR_UNSUPPORTED, "too many problems");

    const std::vector<maxmulsk_plan> defaults =
        plans ? std::vector<maxmulsk_plan>() : batch_defaults(problems, n, mode, true);

    // Split modes run on the decoupled feed under the current feed setting.
    const bool feed = grouped_feed(mode, feed_mode());

    // Validate every problem before anything is allocated or run.
    int dev = -1;
    std::vector<Resolved> res(n);
    std::vector<maxmulsk_plan> used_plan(n);
    std::vector<bool> active(n, false);
    for (size_t i = 0; i < n; ++i) {
        const maxmulsk_gemm& g = problems[i];
        if (g.M == 0 || g.N == 0) continue;
        if (!g.C) return fail(MAXMULSK_ERR_UNSUPPORTED, "problem %zu: C is null", i);
        if (g.M > 0x7fffffffULL || g.N > 0x7fffffffULL || g.K > 0x7fffffffULL)
            return fail(MAXMULSK_ERR_UNSUPPORTED, "problem %zu: shape too large", i);
        if (usable_devices() == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");
        int dc = -1;
        if (!is_device_ptr(g.C, &dc)) return fail(MAXMULSK_ERR_UNSUPPORTED, "problem %zu: C is not device memory", i);
        if (dev < 0) dev = dc;
        if (dc != dev) return fail(MAXMULSK_ERR_UNSUPPORTED, "problem %zu: another device", i);
        if (g.K == 0) continue;
        int da = -1, db = -1;
        if (!g.A || !g.B || !is_device_ptr(g.A, &da) || !is_device_ptr(g.B, &db) || da != dev || db != dev)
            return fail(MAXMULSK_ERR_UNSUPPORTED, "problem %zu: A/B null or not device memory on device %d", i, dev);
        const maxmulsk_plan pl = plans ? plans[i] : defaults[i];
        if (!resolve(g.M, g.K, g.N, mode, pl, &res[i], kGroupedMaxCfg, grouped_min_cfg(mode)))
            return fail(MAXMULSK_ERR_UNSUPPORTED, "problem %zu: plan {%d,%d} rejected", i, pl.tile, pl.split_k);
        used_plan[i] = pl;
        active[i] = true;
    }

    auto* p = new maxmulsk_grouped_plan_s();
    p->mode = mode;
    p->feed = feed;
    p->n = n;
    if (dev < 0) {  // nothing to run
        *out = p;
        return MAXMULSK_OK;
    }
    if (!device_usable(dev)) {
        delete p;
        return fail(MAXMULSK_ERR_UNSUPPORTED, "device is not gfx950");
    }
    p->dev = dev;
    DeviceGuard guard(dev);

    // Work list. In plain mode, and in the split modes on the decoupled feed, a
    // problem on 128-row tiles whose last tile would hold at most 64 rows runs
    // as two row ranges: the 128-row tiles of the first M - M % 128 rows and one
    // 64-row tile for the rest (same plan, so the same K slices and the same
    // products per element; only padding rows are dropped). On the 256-problem
    // batch (M ~ U{1..512}) that cut the padded MFMA work from 22% to 12%: the
    // plain launch went from 32.6 to 30.8 ms, the feed's split3 from 12.15 to
    // 11.88 ms and split6 from 17.84 to 17.55 ms. The direct split-mode kernel
    // keeps the padded tile: its eight-wave 64-row tiles split and stage more
    // operand bytes per product, and split3 measured 1.6% slower with ranges.
    // Row ranges are ordered by per-item cost (most expensive first).
    struct Part {
        const float* A;
        const float* B;
        float* C;
        size_t M, K, N;
        Resolved r;
    };
    std::vector<Part> parts;
    for (size_t i = 0; i < n; ++i) {
        const maxmulsk_gemm& g = problems[i];
        if (g.M == 0 || g.N == 0) continue;
        if (g.K == 0) {
            p->zero_fills.emplace_back(g.C, g.M * g.N);
            continue;
        }
        p->flops += 2.0 * (double)g.M * (double)g.N * (double)g.K;
        if (!active[i]) continue;
        const size_t tail = g.M % 128;
        Resolved rm, rt;
        if ((mode == 0 || feed) && res[i].bm == 128 && g.M > 128 && tail >= 1 && tail <= 64 &&
            resolve(g.M - tail, g.K, g.N, mode, used_plan[i], &rm, kGroupedMaxCfg, grouped_min_cfg(mode)) &&
            resolve(tail, g.K, g.N, mode, used_plan[i], &rt, kGroupedMaxCfg, grouped_min_cfg(mode)) && rm.bm == 128 &&
            rt.bm == 64 && rm.k_len == res[i].k_len && rt.k_len == res[i].k_len) {
            const size_t m0 = g.M - tail;
            parts.push_back(Part{g.A, g.B, g.C, m0, g.K, g.N, rm});
            parts.push_back(Part{g.A + m0 * g.K, g.B, g.C + m0 * g.N, tail, g.K, g.N, rt});
        } else {
            parts.push_back(Part{g.A, g.B, g.C, g.M, g.K, g.N, res[i]});
        }
    }
    auto item_cost = [&](const Part& x) { return (double)x.r.bm * x.r.bn * (double)x.r.k_len; };
    std::stable_sort(parts.begin(), parts.end(), [&](const Part& a, const Part& b) { return item_cost(a) > item_cost(b); });

    std::vector<GroupProblem> gp(parts.size());
    std::vector<int> item_problem, block_problem;
    size_t ws_bytes = 0;
    std::vector<size_t> ws_off(parts.size(), 0);
    long long items = 0;
    for (size_t j = 0; j < parts.size(); ++j) {
        const Part& g = parts[j];
        const Resolved& r = g.r;
        GroupProblem& q = gp[j];
        q.A = g.A;
        q.B = g.B;
        q.C = g.C;
        q.P = nullptr;
        q.M = (long long)g.M;
        q.N = (long long)g.N;
        q.K = (long long)g.K;
        q.k_len = r.k_len;
        q.splits = r.splits;
        q.cfg = r.cfg;
        q.tiles_m = r.tiles_m;
        q.tiles_n = r.tiles_n;
        q.item_begin = (int)items;
        const long long cnt = (long long)r.tiles_m * r.tiles_n * r.splits;
        items += cnt;
        if (items > 0x7fffffffLL) {
            delete p;
            return fail(MAXMULSK_ERR_UNSUPPORTED, "work list too long");
        }
        for (long long t = 0; t < cnt; ++t) item_problem.push_back((int)j);
        if (r.splits > 1) {
            ws_off[j] = ws_bytes;
            ws_bytes += (r.ws_bytes + 255) / 256 * 256;
            q.reduce_begin = (int)block_problem.size();
            const long long blocks = ((long long)g.M * g.N + 1023) / 1024;
            for (long long b = 0; b < blocks; ++b) block_problem.push_back((int)j);
        } else {
            q.reduce_begin = 0;
        }
    }
    p->total_items = (int)items;
    p->reduce_blocks = (int)block_problem.size();

    bool ok = true;
    count(kCntAllocation);  // the plan's device buffers
    if (ws_bytes) ok = device_malloc(&p->d_ws, ws_bytes) == hipSuccess;
    if (ok)
        for (size_t j = 0; j < gp.size(); ++j)
            if (gp[j].splits > 1) gp[j].P = reinterpret_cast<float*>(static_cast<char*>(p->d_ws) + ws_off[j]);
    ok = ok && device_malloc(reinterpret_cast<void**>(&p->d_counter), sizeof(unsigned int)) == hipSuccess;
    ok = ok && upload(&p->d_problems, gp) && upload(&p->d_item_problem, item_problem) &&
         upload(&p->d_block_problem, block_problem);
    if (!ok) {
        (void)hipGetLastError();
        free_plan(p);
        return fail(MAXMULSK_ERR_ALLOC, "grouped plan allocation failed");
    }
    hipDeviceProp_t prop;
    int cus = kNumCUs;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    p->grid = std::max(1, std::min(p->total_items,
                                   cus * (feed ? grouped_feed_blocks_per_cu(mode) : grouped_blocks_per_cu(mode))));

    if (capture) {
        // Capture needs a created stream; the replay may go to any stream.
        hipStream_t cs = nullptr;
        bool captured = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) == hipSuccess &&
                        hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (captured) {
            const hipError_t e = enqueue(p, cs);
            hipGraph_t graph = nullptr;
            const hipError_t e2 = hipStreamEndCapture(cs, &graph);
            captured = e == hipSuccess && e2 == hipSuccess && graph != nullptr;
            if (captured) {
                p->graph = graph;
                captured = hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0) == hipSuccess;
            } else if (graph) {
                (void)hipGraphDestroy(graph);
            }
        }
        if (cs) (void)hipStreamDestroy(cs);
        if (!captured) {
            (void)hipGetLastError();
            free_plan(p);
            return fail(MAXMULSK_ERR_DEVICE, "graph capture of the grouped launch failed");
        }
    }
    (void)stream;
    *out = p;
    return MAXMULSK_OK;
}

maxmulsk_status launch_plan(maxmulsk_grouped_plan_s* p, hipStream_t s) {
    if (p->dev < 0) return MAXMULSK_OK;
    DeviceGuard guard(p->dev);
    const hipError_t e = p->exec ? hipGraphLaunch(p->exec, s) : enqueue(p, s);
    count(kCntZeroFill, p->zero_fills.size());
    if (p->total_items > 0) count(kCntCompute);
    if (p->reduce_blocks > 0) count(kCntReduce);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "grouped launch failed: %s", hipGetErrorString(e));
    }
    return MAXMULSK_OK;
}

}  // namespace

extern "C" maxmulsk_status maxmulsk_grouped_create(const maxmulsk_gemm* problems, size_t n,
                                                   maxmulsk_mode mode, const maxmulsk_plan* plans,
                                                   void* stream, maxmulsk_grouped_plan* out) {
    maxmulsk_grouped_plan_s* p = nullptr;
    const maxmulsk_status st = build(problems, n, (int)mode, plans, static_cast<hipStream_t>(stream), true, &p);
    if (out) *out = st == MAXMULSK_OK ? p : nullptr;
    return st;
}

extern "C" maxmulsk_status maxmulsk_grouped_launch(maxmulsk_grouped_plan plan, void* stream) {
    if (!plan) return fail(MAXMULSK_ERR_UNSUPPORTED, "plan is null");
    return launch_plan(plan, static_cast<hipStream_t>(stream));
}

extern "C" maxmulsk_status maxmulsk_grouped_info(maxmulsk_grouped_plan plan, size_t* work_items,
                                                 double* flops, int* used_graph) {
    if (!plan) return fail(MAXMULSK_ERR_UNSUPPORTED, "plan is null");
    if (work_items) *work_items = (size_t)plan->total_items;
    if (flops) *flops = plan->flops;
    if (used_graph) *used_graph = plan->exec != nullptr;
    return MAXMULSK_OK;
}

extern "C" int maxmulsk_grouped_uses_feed(maxmulsk_grouped_plan plan) { return plan && plan->feed ? 1 : 0; }

extern "C" void maxmulsk_grouped_destroy(maxmulsk_grouped_plan plan) { free_plan(plan); }

extern "C" maxmulsk_plan maxmulsk_grouped_default_plan(size_t M, size_t K, size_t N, maxmulsk_mode mode) {
    return grouped_default(M, K, N, (int)mode);
}

extern "C" maxmulsk_status maxmulsk_grouped_plans(const maxmulsk_gemm* problems, size_t n, maxmulsk_mode mode,
                                                  maxmulsk_plan* out) {
    if (n > 0 && (!problems || !out)) return fail(MAXMULSK_ERR_UNSUPPORTED, "problems or out is null");
    if ((int)mode < 0 || (int)mode > 2) return fail(MAXMULSK_ERR_UNSUPPORTED, "unknown mode");
    const std::vector<maxmulsk_plan> p = batch_defaults(problems, n, (int)mode, true);
    for (size_t i = 0; i < n; ++i) out[i] = p[i];
    return MAXMULSK_OK;
}

extern "C" maxmulsk_status maxmulsk_sgemm_grouped(const maxmulsk_gemm* problems, size_t n,
                                                  const maxmulsk_options* opt) {
    const int mode = opt ? (int)opt->mode : 0;
    hipStream_t s = opt ? static_cast<hipStream_t>(opt->stream) : nullptr;
    maxmulsk_grouped_plan_s* p = nullptr;
    maxmulsk_status st = build(problems, n, mode, nullptr, s, false, &p);
    if (st != MAXMULSK_OK) return st;
    st = launch_plan(p, s);
    if (st == MAXMULSK_OK && p->dev >= 0) {
        DeviceGuard guard(p->dev);
        if (hipStreamSynchronize(s) != hipSuccess) {
            (void)hipGetLastError();
            st = fail(MAXMULSK_ERR_DEVICE, "grouped launch failed");
        }
    }
    free_plan(p);
    return st;
}

// ---- grouped autotuner -----------------------------------------------------------

extern "C" maxmulsk_status maxmulsk_grouped_autotune(const maxmulsk_gemm* problems, size_t n,
                                                     maxmulsk_mode mode, int max_classes, int reps,
                                                     maxmulsk_plan* plans_out, float* best_ms) {
    if (!problems || n == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "no problems");
    const int md = (int)mode;
    if (md < 0 || md > 2) return fail(MAXMULSK_ERR_UNSUPPORTED, "unknown mode");
    if (reps <= 0) reps = 3;

    // Shape classes and their total work.
    std::map<ClassKey, double> work;
    std::vector<ClassKey> cls(n);
    for (size_t i = 0; i < n; ++i) {
        const maxmulsk_gemm& g = problems[i];
        cls[i] = class_of(g.M, g.K, g.N, md);
        work[cls[i]] += 2.0 * (double)g.M * g.N * g.K;
    }
    std::vector<std::pair<double, ClassKey>> by_work;
    for (const auto& kv : work) by_work.emplace_back(kv.second, kv.first);
    std::sort(by_work.begin(), by_work.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    // Start from the batch's rule plans (the cache is not consulted), one per
    // class: the rule gives every problem of a class the same plan.
    std::map<ClassKey, maxmulsk_plan> choice;
    const std::vector<maxmulsk_plan> start = batch_defaults(problems, n, md, false);
    for (size_t i = 0; i < n; ++i)
        if (start[i].tile != 0 && !choice.count(cls[i])) choice[cls[i]] = start[i];
    for (const auto& kv : work)
        if (!choice.count(kv.first))
            choice[kv.first] = grouped_rule((size_t)std::get<0>(kv.first), std::get<2>(kv.first), std::get<1>(kv.first));

    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "stream/event creation failed");
    }
    std::vector<maxmulsk_plan> plans(n);
    auto evaluate = [&](float* ms_out) -> maxmulsk_status {
        for (size_t i = 0; i < n; ++i) plans[i] = choice[cls[i]];
        maxmulsk_grouped_plan_s* p = nullptr;
        maxmulsk_status st = build(problems, n, md, plans.data(), s, false, &p);
        if (st != MAXMULSK_OK) return st;
        std::vector<float> t;
        st = launch_plan(p, s);  // warm
        for (int r = 0; st == MAXMULSK_OK && r < reps; ++r) {
            if (hipEventRecord(e0, s) != hipSuccess) st = MAXMULSK_ERR_DEVICE;
            if (st == MAXMULSK_OK) st = launch_plan(p, s);
            float ms = 0;
            if (st == MAXMULSK_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                      hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                st = MAXMULSK_ERR_DEVICE;
            t.push_back(ms);
        }
        if (p && p->dev >= 0) {
            DeviceGuard g(p->dev);
            (void)hipStreamSynchronize(s);
        }
        free_plan(p);
        if (st != MAXMULSK_OK) {
            (void)hipGetLastError();
            return fail(st, "grouped autotune launch failed");
        }
        std::sort(t.begin(), t.end());
        *ms_out = t[t.size() / 2];
        return MAXMULSK_OK;
    };

    float best = 0;
    maxmulsk_status st = evaluate(&best);
    const int limit = max_classes > 0 ? std::min<int>(max_classes, (int)by_work.size()) : (int)by_work.size();
    for (int c = 0; st == MAXMULSK_OK && c < limit; ++c) {
        const ClassKey k = by_work[c].second;
        const maxmulsk_plan keep = choice[k];
        maxmulsk_plan win = keep;
        for (int i = 0; i < kNumCandidates; ++i) {
            const maxmulsk_plan cand = candidate(i);
            if (cand.tile == keep.tile && cand.split_k == keep.split_k) continue;
            Resolved r;
            if (!resolve((size_t)std::get<0>(k), std::get<2>(k), std::get<1>(k), md, cand, &r, kGroupedMaxCfg,
                         grouped_min_cfg(md)))
                continue;
            choice[k] = cand;
            float ms = 0;
            st = evaluate(&ms);
            if (st != MAXMULSK_OK) break;
            if (ms < best * 0.995f) {  // a win must beat the noise floor
                best = ms;
                win = cand;
            }
        }
        choice[k] = win;
    }
    (void)hipStreamDestroy(s);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != MAXMULSK_OK) return st;
    {
        std::lock_guard<std::mutex> g(g_gmu);
        for (const auto& kv : choice) g_group_cache[kv.first] = kv.second;
    }
    if (plans_out)
        for (size_t i = 0; i < n; ++i) plans_out[i] = choice[cls[i]];
    if (best_ms) *best_ms = best;
    return MAXMULSK_OK;
}
