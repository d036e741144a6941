#include "trace.hpp"

#include <dlfcn.h>

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace mmsk {

namespace {

std::atomic<uint64_t> g_counts[kNumPhaseCounters];
constexpr const char* kNames[kNumPhaseCounters] = {"allocation", "stage_in", "pack", "compute",
                                                   "reduce", "writeback", "zero_fill"};

using PushFn = int (*)(const char*);
using PopFn = int (*)();
std::once_flag g_once;
PushFn g_push = nullptr;
PopFn g_pop = nullptr;

// Loads the roctx library only when asked for, so the product has no link-time
// dependency on it.
void load_roctx() {
    const char* env = std::getenv("MAXMULSK_ROCTX");
    if (!env || std::strcmp(env, "1") != 0) return;
    for (const char* lib : {"librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so", "libroctx64.so.4"}) {
        void* h = dlopen(lib, RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        auto push = reinterpret_cast<PushFn>(dlsym(h, "roctxRangePushA"));
        auto pop = reinterpret_cast<PopFn>(dlsym(h, "roctxRangePop"));
        if (push && pop) {
            g_push = push;
            g_pop = pop;
            return;
        }
        dlclose(h);
    }
}

}  // namespace

void count(PhaseCounter c, uint64_t n) { g_counts[c].fetch_add(n, std::memory_order_relaxed); }

uint64_t counter(int i) {
    return i >= 0 && i < kNumPhaseCounters ? g_counts[i].load(std::memory_order_relaxed) : 0;
}

const char* counter_name(int i) { return i >= 0 && i < kNumPhaseCounters ? kNames[i] : nullptr; }

void reset_counters() {
    for (auto& c : g_counts) c.store(0, std::memory_order_relaxed);
}

Scope::Scope(const char* name) : on_(false) {
    std::call_once(g_once, load_roctx);
    if (g_push) {
        g_push(name);
        on_ = true;
    }
}

Scope::~Scope() {
    if (on_) g_pop();
}

}  // namespace mmsk
