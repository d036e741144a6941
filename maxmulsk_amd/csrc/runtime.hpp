// Host runtime shared by the C-ABI entry points: device discovery, per-device
// workspace and staging pools, pointer classification, error text.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <string>

#include "maxmulsk.h"
#include "plan.hpp"

namespace mmsk {

// Sets the thread's last-error text (printf style) and returns `st`.
maxmulsk_status fail(maxmulsk_status st, const char* fmt, ...);
const char* last_error();

// Number of visible gfx950 devices (0 when there is no GPU or HIP fails).
int usable_devices();
bool device_usable(int dev);
// "gfx950:<marketing name>" of a device, for cache files.
std::string device_name(int dev);

// True when p is device (or managed) memory; *dev receives its device.
bool is_device_ptr(const void* p, int* dev);
// True when p is pinned host memory (hipHostMalloc), which DMA reads directly.
bool is_pinned_host(const void* p);

// Makes `dev` current for the lifetime of the guard.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev);
    ~DeviceGuard();
};

// Per-device pools. The workspace pool serves split-K partials of calls that
// bring no workspace of their own; every use is stream-ordered behind the
// previous use through `ws_event`, so overlapping asynchronous calls on
// different streams never share partials. The staging pool holds device copies
// of host operands; it is only used by synchronous calls under `mu`.
struct DevicePools {
    std::mutex mu;
    void* ws = nullptr;
    size_t ws_cap = 0;
    hipEvent_t ws_event = nullptr;
    bool ws_event_recorded = false;   // ws_event covers the last use
    bool ws_used = false;             // the workspace has been handed out before
    hipStream_t ws_stream = nullptr;  // stream of the last use
    void* stage = nullptr;
    size_t stage_cap = 0;
    // Pinned bounce buffers between pageable caller memory and the staging
    // buffer, with the event of each one's last copy (synchronous calls only).
    void* bounce[2] = {nullptr, nullptr};
    hipEvent_t bounce_ev[2] = {nullptr, nullptr};
    bool bounce_used[2] = {false, false};
    bool bounce_failed = false;  // pinned allocation failed once: copy pageable memory directly
    // Stream and events of the autotuner, created on first use and kept (a
    // stream holds a hardware queue; creating one per tuning call is costly).
    hipStream_t tune_stream = nullptr;
    hipEvent_t tune_e0 = nullptr, tune_e1 = nullptr;
    float* tune_sample = nullptr;  // 4096 floats: gathered sample of C
    double* tune_ref = nullptr;    // 2 x 4096 doubles: FP64 value and |A||B| of the sample
};
DevicePools& pools(int dev);

// hipMalloc for every device buffer the library owns. For tests of the
// allocation-failure paths, MAXMULSK_TEST_ALLOC_LIMIT=<bytes> makes any larger
// request fail as hipMalloc would (hipErrorOutOfMemory), without allocating.
hipError_t device_malloc_bytes(void** p, size_t bytes);
template <class T>
hipError_t device_malloc(T** p, size_t bytes) {
    return device_malloc_bytes(reinterpret_cast<void**>(p), bytes);
}

// Grows a pool buffer to at least `bytes` (rounded up to 2 MiB). Waits for the
// device before freeing the old buffer. Returns false on allocation failure.
bool grow(void** buf, size_t* cap, size_t bytes);

// Stream-orders the pooled workspace (grown to `bytes`) behind its previous
// use and hands it out; caller holds pl.mu. After enqueuing the work that uses
// it, the caller records the release on the same stream.
maxmulsk_status pool_workspace_acquire(DevicePools& pl, size_t bytes, hipStream_t stream, void** ws);
hipError_t pool_workspace_release(DevicePools& pl, hipStream_t stream);

// The GEMM launch itself (no contract checks): all pointers are device
// pointers, `ws` holds r.ws_bytes when r.splits > 1.
hipError_t launch_resolved(const float* A, const float* B, float* C, size_t M, size_t K, size_t N,
                           int mode, const Resolved& r, void* ws, hipStream_t stream);

// Full entry point behind every GEMM symbol.
struct GemmCall {
    const float* A;
    const float* B;
    float* C;
    size_t M, K, N;
    int mode;
    bool have_plan;       // use `plan` instead of the planner
    maxmulsk_plan plan;
    bool rule_only;       // planner restricted to override + shape rule (v3)
    hipStream_t stream;
    bool async;
    void* ws;             // caller workspace (nullptr: pool)
    size_t ws_bytes;
    bool ws_required;     // caller promised a workspace (the _ws entry point)
    // Static-scratch entry point: the planner skips the selector tier (its
    // first use on a device allocates), and a workspace above ws_bytes is
    // refused as too large for the scratch.
    bool no_alloc = false;
};
maxmulsk_status run_gemm(const GemmCall& c);

}  // namespace mmsk
