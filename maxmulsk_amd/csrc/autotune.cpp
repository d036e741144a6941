// Two-stage, correctness-gated exhaustive autotuner over the fixed candidate
// set (counterpart of the blocking sweep
// bench/results/2026-09-13/source/bench_nb_sweep.cpp:4-18, 116-122, 160-212):
//   stage 1: every candidate, short, in rotated order over several passes,
//            median of the per-pass averages;
//   stage 2: the leaders (best three, anything within 3% of the best, and the
//            shape rule's plan), longer, rotated again;
//   gate:    a candidate whose result disagrees with the first valid one on a
//            strided sample of C is excluded from every ranking.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "maxmulsk.h"
#include "plan.hpp"
#include "runtime.hpp"

using namespace mmsk;

namespace {

float median(std::vector<float> v) {
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return n ? ((n % 2) ? v[n / 2] : 0.5f * (v[n / 2 - 1] + v[n / 2])) : -1.f;
}

struct Tuner {
    const float* A;
    const float* B;
    float* C;
    size_t M, K, N;
    int mode;
    void* ws;
    hipStream_t s;
    hipEvent_t e0, e1;

    hipError_t run(const Resolved& r) { return launch_resolved(A, B, C, M, K, N, mode, r, ws, s); }

    // Average ms per launch over `reps` back-to-back launches; -1 on failure.
    float time(const Resolved& r, int reps) {
        if (hipEventRecord(e0, s) != hipSuccess) return -1.f;
        for (int i = 0; i < reps; ++i)
            if (run(r) != hipSuccess) return -1.f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.f;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.f;
        return ms / reps;
    }
};

// Strided sample of C: at most 4096 values at a fixed stride, gathered on the
// device into `buf` and copied back in one piece.
std::vector<float> sample_c(const float* C, size_t M, size_t N, float* buf, hipStream_t s) {
    const size_t total = M * N;
    const size_t count = std::min<size_t>(total, 4096);
    const size_t stride = std::max<size_t>(1, total / count);
    std::vector<float> out(count);
    if (launch_gather_strided(C, stride, count, buf, s) != hipSuccess ||
        hipMemcpyAsync(out.data(), buf, count * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        out.clear();
    }
    return out;
}

bool agree(const std::vector<float>& a, const std::vector<float>& ref) {
    if (a.size() != ref.size() || a.empty()) return false;
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        if (!std::isfinite(a[i])) return false;
        const double d = (double)a[i] - ref[i];
        num += d * d;
        den += (double)ref[i] * ref[i];
    }
    return den == 0 ? num == 0 : std::sqrt(num / den) <= 1e-4;
}

}  // namespace

extern "C" maxmulsk_status maxmulsk_autotune(const float* A, const float* B, float* C, size_t M,
                                             size_t K, size_t N, maxmulsk_mode mode, int stage1_reps,
                                             int stage2_reps, float* times_ms, maxmulsk_plan* best) {
    if (!A || !B || !C) return fail(MAXMULSK_ERR_UNSUPPORTED, "null operand");
    if ((int)mode < 0 || (int)mode > 2) return fail(MAXMULSK_ERR_UNSUPPORTED, "unknown mode");
    if (M == 0 || K == 0 || N == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "empty shape");
    if (usable_devices() == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");
    int da = -1, db = -1, dc = -1;
    if (!is_device_ptr(A, &da) || !is_device_ptr(B, &db) || !is_device_ptr(C, &dc) || da != db ||
        db != dc)
        return fail(MAXMULSK_ERR_UNSUPPORTED, "autotune needs device pointers on one device");
    if (!device_usable(dc)) return fail(MAXMULSK_ERR_UNSUPPORTED, "device is not gfx950");
    const int s1 = stage1_reps > 0 ? stage1_reps : 3;
    const int s2 = stage2_reps > 0 ? stage2_reps : 10;

    std::vector<Resolved> res(kNumCandidates);
    std::vector<bool> valid(kNumCandidates);
    size_t ws_need = 0;
    for (int i = 0; i < kNumCandidates; ++i) {
        valid[i] = resolve(M, K, N, (int)mode, candidate(i), &res[i]);
        if (valid[i]) ws_need = std::max(ws_need, res[i].ws_bytes);
    }

    DeviceGuard guard(dc);
    DevicePools& pl = pools(dc);
    std::lock_guard<std::mutex> lock(pl.mu);
    if (ws_need && !grow(&pl.ws, &pl.ws_cap, ws_need)) return fail(MAXMULSK_ERR_ALLOC, "workspace allocation failed");
    // Earlier asynchronous users of the pooled workspace must be done.
    if (pl.ws_event_recorded && hipEventSynchronize(pl.ws_event) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "event wait failed");
    }

    if ((!pl.tune_stream && hipStreamCreateWithFlags(&pl.tune_stream, hipStreamNonBlocking) != hipSuccess) ||
        (!pl.tune_e0 && hipEventCreate(&pl.tune_e0) != hipSuccess) ||
        (!pl.tune_e1 && hipEventCreate(&pl.tune_e1) != hipSuccess) ||
        (!pl.tune_sample && hipMalloc(&pl.tune_sample, 4096 * sizeof(float)) != hipSuccess)) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "stream/event creation failed");
    }
    Tuner t{A, B, C, M, K, N, (int)mode, pl.ws, pl.tune_stream, pl.tune_e0, pl.tune_e1};
    // Wait for work the caller queued on the legacy stream (the inputs).
    (void)hipDeviceSynchronize();

    std::vector<float> stage1(kNumCandidates, -1.f), final_t(kNumCandidates, -1.f);
    std::vector<float> ref;
    // Correctness gate: first valid run defines the reference sample.
    for (int i = 0; i < kNumCandidates; ++i) {
        if (!valid[i]) continue;
        if (t.run(res[i]) != hipSuccess) {
            (void)hipGetLastError();
            valid[i] = false;
            continue;
        }
        std::vector<float> smp = sample_c(C, M, N, pl.tune_sample, t.s);
        if (ref.empty()) {
            bool finite = !smp.empty();
            for (float v : smp) finite = finite && std::isfinite(v);
            if (finite) ref = smp;
            else valid[i] = false;
        } else if (!agree(smp, ref)) {
            valid[i] = false;
        }
    }

    // Stage 1: rotated order, 2 passes.
    std::vector<std::vector<float>> passes(kNumCandidates);
    for (int pass = 0; pass < 2; ++pass) {
        for (int j = 0; j < kNumCandidates; ++j) {
            const int i = (j + pass * 7) % kNumCandidates;
            if (!valid[i]) continue;
            const float ms = t.time(res[i], s1);
            if (ms < 0) valid[i] = false;
            else passes[i].push_back(ms);
        }
    }
    float best1 = 1e30f;
    for (int i = 0; i < kNumCandidates; ++i) {
        if (!valid[i]) continue;
        stage1[i] = median(passes[i]);
        best1 = std::min(best1, stage1[i]);
    }
    if (best1 >= 1e30f) return fail(MAXMULSK_ERR_DEVICE, "no candidate ran correctly");

    // Stage 2: leaders.
    std::vector<int> order;
    for (int i = 0; i < kNumCandidates; ++i)
        if (valid[i]) order.push_back(i);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return stage1[a] < stage1[b]; });
    std::vector<bool> leader(kNumCandidates, false);
    for (size_t j = 0; j < order.size(); ++j)
        if (j < 3 || stage1[order[j]] <= best1 * 1.03f) leader[order[j]] = true;
    const int rule_i = candidate_index(rule_plan(M, K, N, (int)mode));
    if (rule_i >= 0 && valid[rule_i]) leader[rule_i] = true;
    std::vector<std::vector<float>> p2(kNumCandidates);
    for (int pass = 0; pass < 3; ++pass) {
        for (int j = 0; j < kNumCandidates; ++j) {
            const int i = (j + pass * 5) % kNumCandidates;
            if (!leader[i]) continue;
            const float ms = t.time(res[i], s2);
            if (ms >= 0) p2[i].push_back(ms);
        }
    }
    int best_i = -1;
    float best_t = 1e30f;
    for (int i = 0; i < kNumCandidates; ++i) {
        if (!valid[i]) continue;
        final_t[i] = (leader[i] && !p2[i].empty()) ? median(p2[i]) : stage1[i];
        // Only stage-2 times compete for the win.
        if (leader[i] && final_t[i] < best_t) {
            best_t = final_t[i];
            best_i = i;
        }
    }
    if (best_i < 0) return fail(MAXMULSK_ERR_DEVICE, "autotune produced no timing");
    if (times_ms)
        for (int i = 0; i < kNumCandidates; ++i) times_ms[i] = final_t[i];
    const maxmulsk_plan bp = candidate(best_i);
    if (best) *best = bp;
    cache_put(M, K, N, (int)mode, bp);
    return MAXMULSK_OK;
}
