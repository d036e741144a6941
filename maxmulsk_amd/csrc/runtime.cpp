#include "runtime.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "copy_workers.hpp"
#include "kernels.hpp"
#include "trace.hpp"

namespace mmsk {

namespace {
thread_local std::string t_err;
constexpr int kMaxDevices = 64;
DevicePools g_pools[kMaxDevices];
std::once_flag g_dev_once;
int g_dev_count = 0;
bool g_dev_ok[kMaxDevices] = {};

void discover() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    if (n > kMaxDevices) n = kMaxDevices;
    int ok = 0;
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) != hipSuccess) {
            (void)hipGetLastError();
            continue;
        }
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) {
            g_dev_ok[d] = true;
            ++ok;
        }
    }
    g_dev_count = ok;
}
}  // namespace

maxmulsk_status fail(maxmulsk_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return st;
}

const char* last_error() { return t_err.c_str(); }

int usable_devices() {
    std::call_once(g_dev_once, discover);
    return g_dev_count;
}

bool device_usable(int dev) {
    usable_devices();
    return dev >= 0 && dev < kMaxDevices && g_dev_ok[dev];
}

std::string device_name(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hipGetLastError();
        return "unknown";
    }
    std::string arch(prop.gcnArchName);
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    return arch + ":" + prop.name;
}

bool is_device_ptr(const void* p, int* dev) {
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof attr);
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) {
        *dev = attr.device;
        return true;
    }
    return false;
}

bool is_pinned_host(const void* p) {
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof attr);
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeHost;
}

DeviceGuard::DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) {
        (void)hipGetLastError();
        prev = -1;
    }
    if (prev != dev) (void)hipSetDevice(dev);
}

DeviceGuard::~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
}

DevicePools& pools(int dev) { return g_pools[dev]; }

hipError_t device_malloc_bytes(void** p, size_t bytes) {
    if (const char* lim = std::getenv("MAXMULSK_TEST_ALLOC_LIMIT")) {
        char* end = nullptr;
        const unsigned long long limit = std::strtoull(lim, &end, 10);
        if (end != lim && bytes > limit) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
    }
    return hipMalloc(p, bytes);
}

bool grow(void** buf, size_t* cap, size_t bytes) {
    if (*cap >= bytes && *buf) return true;
    const size_t want = (bytes + (2u << 20) - 1) / (2u << 20) * (2u << 20);
    if (*buf) {
        (void)hipDeviceSynchronize();
        (void)hipFree(*buf);
        *buf = nullptr;
        *cap = 0;
    }
    void* p = nullptr;
    Scope scope("maxmulsk/allocation");
    if (device_malloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    count(kCntAllocation);
    *buf = p;
    *cap = want;
    return true;
}

namespace {

CopyWorkers& copy_workers() {
    static CopyWorkers* w = [] {
        const unsigned hc = std::thread::hardware_concurrency();
        return new CopyWorkers(hc >= 8 ? 3 : (hc >= 2 ? (int)hc / 2 - 1 : 0));
    }();
    return *w;
}

constexpr size_t kBounceBytes = 8u << 20;  // per bounce buffer

// Pinned bounce buffers of the device (caller holds pl.mu). False when pinned
// memory is unavailable; callers then copy pageable memory directly.
bool ensure_bounce(DevicePools& pl) {
    if (pl.bounce_failed) return false;
    if (pl.bounce[0]) return true;
    bool ok = true;
    for (int b = 0; b < 2 && ok; ++b) {
        ok = hipHostMalloc(&pl.bounce[b], kBounceBytes, hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&pl.bounce_ev[b], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
        (void)hipGetLastError();
        for (int b = 0; b < 2; ++b) {
            if (pl.bounce[b]) (void)hipHostFree(pl.bounce[b]);
            if (pl.bounce_ev[b]) (void)hipEventDestroy(pl.bounce_ev[b]);
            pl.bounce[b] = nullptr;
            pl.bounce_ev[b] = nullptr;
        }
        pl.bounce_failed = true;
    }
    return ok;
}

// Host -> device copy of a caller operand on `stream`. Pageable memory goes
// through the two bounce buffers: the host fills one while the DMA of the
// other runs. Returns when the host has consumed `src` (DMAs may still run).
hipError_t stage_in(DevicePools& pl, void* dst, const void* src, size_t n, hipStream_t stream) {
    if (is_pinned_host(src) || !ensure_bounce(pl))
        return hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, stream);
    CopyWorkers& cw = copy_workers();
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    for (size_t off = 0, i = 0; off < n; off += kBounceBytes, ++i) {
        const int b = (int)(i & 1);
        const size_t len = std::min(kBounceBytes, n - off);
        if (pl.bounce_used[b]) {
            const hipError_t e = hipEventSynchronize(pl.bounce_ev[b]);
            if (e != hipSuccess) return e;
        }
        cw.copy(pl.bounce[b], s + off, len);
        hipError_t e = hipMemcpyAsync(d + off, pl.bounce[b], len, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipEventRecord(pl.bounce_ev[b], stream);
        if (e != hipSuccess) return e;
        pl.bounce_used[b] = true;
    }
    return hipSuccess;
}

// Device -> host copy of C on `stream` (after the work queued before it).
// Pageable memory: the DMA fills one bounce buffer while the host empties the
// other. Returns with `dst` complete.
hipError_t stage_out(DevicePools& pl, void* dst, const void* src, size_t n, hipStream_t stream) {
    if (is_pinned_host(dst) || !ensure_bounce(pl)) {
        const hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, stream);
        return e == hipSuccess ? hipStreamSynchronize(stream) : e;
    }
    CopyWorkers& cw = copy_workers();
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    const size_t chunks = (n + kBounceBytes - 1) / kBounceBytes;
    auto issue = [&](size_t i) {
        const int b = (int)(i & 1);
        const size_t off = i * kBounceBytes, len = std::min(kBounceBytes, n - off);
        hipError_t e = hipMemcpyAsync(pl.bounce[b], s + off, len, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipEventRecord(pl.bounce_ev[b], stream);
        pl.bounce_used[b] = true;
        return e;
    };
    hipError_t e = issue(0);
    if (e == hipSuccess && chunks > 1) e = issue(1);
    for (size_t i = 0; e == hipSuccess && i < chunks; ++i) {
        const int b = (int)(i & 1);
        const size_t off = i * kBounceBytes, len = std::min(kBounceBytes, n - off);
        e = hipEventSynchronize(pl.bounce_ev[b]);
        if (e != hipSuccess) break;
        cw.copy(d + off, pl.bounce[b], len);
        if (i + 2 < chunks) e = issue(i + 2);
    }
    return e;
}

}  // namespace

maxmulsk_status pool_workspace_acquire(DevicePools& pl, size_t bytes, hipStream_t stream, void** ws) {
    if (!grow(&pl.ws, &pl.ws_cap, bytes)) return fail(MAXMULSK_ERR_ALLOC, "workspace allocation of %zu bytes failed", bytes);
    // The event only orders device work between streams (hipStreamWaitEvent);
    // the host never inspects it. Without the system-scope fence its record
    // costs about 1.6 us of stream time instead of 3.
    if (!pl.ws_event &&
        hipEventCreateWithFlags(&pl.ws_event, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "event creation failed");
    }
    // A use on the stream of the previous use is ordered behind it by the
    // stream itself, so the event is only needed across streams, and it is
    // not free: on a 64 x 512 x 8192 split-K call its record cost about
    // 1.6 us of stream time, and the wait 6.5 us of host time and 3 us of
    // stream time. Uses on the device's legacy default stream therefore skip
    // both: that stream is one stream for every thread and is never
    // destroyed, so when a use on another stream follows, the event is
    // recorded on the default stream then, behind everything enqueued there
    // so far. Uses on created streams record at release and are always waited
    // for: hipStreamDestroy may return while a stream's work still runs, after
    // which a new stream can reuse the handle, and hipStreamPerThread names a
    // different stream in each thread.
    if (pl.ws_used && pl.ws_stream == nullptr) {
        if (stream != nullptr) {
            if (hipEventRecord(pl.ws_event, nullptr) != hipSuccess ||
                hipStreamWaitEvent(stream, pl.ws_event, 0) != hipSuccess) {
                (void)hipGetLastError();
                return fail(MAXMULSK_ERR_DEVICE, "stream wait failed");
            }
        }
    } else if (pl.ws_event_recorded && hipStreamWaitEvent(stream, pl.ws_event, 0) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "stream wait failed");
    }
    *ws = pl.ws;
    return MAXMULSK_OK;
}

hipError_t pool_workspace_release(DevicePools& pl, hipStream_t stream) {
    pl.ws_used = true;
    pl.ws_stream = stream;
    if (stream == nullptr) {
        pl.ws_event_recorded = false;  // recorded by the next use on another stream
        return hipSuccess;
    }
    const hipError_t e = hipEventRecord(pl.ws_event, stream);
    pl.ws_event_recorded = (e == hipSuccess);
    return e;
}

hipError_t launch_resolved(const float* A, const float* B, float* C, size_t M, size_t K, size_t N,
                           int mode, const Resolved& r, void* ws, hipStream_t stream) {
    char* wsb = static_cast<char*>(ws);
    float* partials = r.splits > 1 ? reinterpret_cast<float*>(wsb) : nullptr;
    hipError_t e = hipSuccess;
    if (r.packed) {
        char* pa = wsb + r.partial_bytes;
        char* pb = pa + r.pack_a_bytes;
        const int terms = mode == 0 ? 1 : (mode == 1 ? 2 : 3);
        const int rbA = r.tiles_m * (r.bm / 32), rbB = r.tiles_n * (r.bn / 32);
        // A rows are k-contiguous; B is read with k as the slow index.
        const PackJob ja{A, (long long)M, (long long)K, 1, pa, rbA};
        const PackJob jb{B, (long long)N, 1, (long long)N, pb, rbB};
        e = launch_pack2(mode, ja, jb, (long long)K, r.kgroups, stream);
        if (e != hipSuccess) return e;
        count(kCntPack);
        PackedParams p{};
        p.PA = pa;
        p.PB = pb;
        p.planeA = (long long)r.pack_a_bytes / terms;
        p.planeB = (long long)r.pack_b_bytes / terms;
        p.kgtA = p.kgtB = r.kgroups;
        p.M = (long long)M;
        p.N = (long long)N;
        p.ldc = (long long)N;
        p.kblocks = r.kblocks;
        p.kblocks_per_split = r.k_len / r.bk;
        p.splits = r.splits;
        p.tiles_m = r.tiles_m;
        p.tiles_n = r.tiles_n;
        if (r.splits > 1) {
            p.C = partials;
            p.c_split_stride = (long long)(M * N);
        } else {
            p.C = C;
            p.c_split_stride = 0;
        }
        e = launch_gemm_packed(mode, p, stream);
        count(kCntCompute);
    } else {
        GemmParams p{};
        p.A = A;
        p.B = B;
        p.M = (long long)M;
        p.N = (long long)N;
        p.K = (long long)K;
        p.lda = (long long)K;
        p.ldb = (long long)N;
        p.ldc = (long long)N;
        p.k_len = r.k_len;
        p.splits = r.splits;
        p.tiles_m = r.tiles_m;
        p.tiles_n = r.tiles_n;
        if (r.splits > 1) {
            p.C = partials;
            p.c_split_stride = (long long)(M * N);
        } else {
            p.C = C;
            p.c_split_stride = 0;
        }
        e = launch_gemm(mode, r.cfg, p, stream);
        count(kCntCompute);
    }
    if (e != hipSuccess || r.splits == 1) return e;
    count(kCntReduce);
    return launch_splitk_reduce(partials, C, (long long)(M * N), r.splits, (long long)(M * N), stream);
}

maxmulsk_status run_gemm(const GemmCall& c) {
    // Empty-shape contract of capi/maxmulsk.cpp:13-20: nothing to do when M or N
    // is 0; with K == 0 the product is the zero matrix.
    if (c.M == 0 || c.N == 0) return MAXMULSK_OK;
    if (!c.C) return fail(MAXMULSK_ERR_UNSUPPORTED, "C is null");
    if (c.mode < 0 || c.mode > 2) return fail(MAXMULSK_ERR_UNSUPPORTED, "unknown mode %d", c.mode);
    if (c.M > 0x7fffffffULL || c.N > 0x7fffffffULL || c.K > 0x7fffffffULL ||
        c.M * c.N > (size_t)1 << 40)
        return fail(MAXMULSK_ERR_UNSUPPORTED, "shape too large");
    // C is classified even without a usable device: device memory of another
    // GPU must not reach the host memset below.
    int cdev = -1;
    const bool c_dev = is_device_ptr(c.C, &cdev);
    if (c.K == 0) {
        if (!c_dev) {
            std::memset(c.C, 0, c.M * c.N * sizeof(float));
            return MAXMULSK_OK;
        }
        if (!device_usable(cdev)) return fail(MAXMULSK_ERR_UNSUPPORTED, "device %d is not gfx950", cdev);
        DeviceGuard g(cdev);
        count(kCntZeroFill);
        if (launch_fill_zero_f32(c.C, c.M * c.N, c.stream) != hipSuccess ||
            (!c.async && hipStreamSynchronize(c.stream) != hipSuccess)) {
            (void)hipGetLastError();
            return fail(MAXMULSK_ERR_DEVICE, "zero fill of C failed");
        }
        return MAXMULSK_OK;
    }
    if (!c.A || !c.B) return fail(MAXMULSK_ERR_UNSUPPORTED, "A or B is null");
    // An explicit plan is validated before anything else (classify runs before
    // any allocation in the reference, ncblock.cpp:627-636).
    Resolved r;
    if (c.have_plan && !resolve(c.M, c.K, c.N, c.mode, c.plan, &r))
        return fail(MAXMULSK_ERR_UNSUPPORTED, "plan {%d,%d} rejected for %zux%zux%zu", c.plan.tile,
                    c.plan.split_k, c.M, c.K, c.N);
    if (usable_devices() == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");

    int adev = -1, bdev = -1;
    const bool a_dev = is_device_ptr(c.A, &adev);
    const bool b_dev = is_device_ptr(c.B, &bdev);
    int dev = c_dev ? cdev : (a_dev ? adev : (b_dev ? bdev : -1));
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "hipGetDevice failed");
    }
    if ((a_dev && adev != dev) || (b_dev && bdev != dev))
        return fail(MAXMULSK_ERR_UNSUPPORTED, "A, B and C live on different devices");
    if (!device_usable(dev)) return fail(MAXMULSK_ERR_UNSUPPORTED, "device %d is not gfx950", dev);
    const bool staged = !(a_dev && b_dev && c_dev);
    if (staged && c.async) return fail(MAXMULSK_ERR_UNSUPPORTED, "async calls need device pointers");
    if (staged && c.ws_required) return fail(MAXMULSK_ERR_UNSUPPORTED, "the workspace entry point needs device pointers");

    // The operands' device is current from here on: the selector tier of the
    // planner runs its inference kernel there.
    DeviceGuard guard(dev);
    maxmulsk_plan plan = c.plan;
    bool resolved;
    {
        Scope scope("maxmulsk/plan");
        if (!c.have_plan) {
            if (c.rule_only) {
                if (!env_plan(&plan)) plan = rule_plan(c.M, c.K, c.N, c.mode);
            } else {
                plan = c.no_alloc ? choose_plan_no_selector(c.M, c.K, c.N, c.mode)
                                  : choose_plan(c.M, c.K, c.N, c.mode, nullptr);
            }
        }
        resolved = resolve(c.M, c.K, c.N, c.mode, plan, &r);
    }
    if (!resolved)
        return fail(MAXMULSK_ERR_UNSUPPORTED, "plan {%d,%d} rejected for %zux%zux%zu", plan.tile,
                    plan.split_k, c.M, c.K, c.N);

    if (c.no_alloc && r.ws_bytes > c.ws_bytes)
        return fail(MAXMULSK_ERR_ALLOC, "plan {%d,%d} for %zux%zux%zu needs %zu bytes of workspace, above the "
                    "%zu-byte static scratch", plan.tile, plan.split_k, c.M, c.K, c.N, r.ws_bytes, c.ws_bytes);
    if (c.ws_required || c.ws) {
        if (r.ws_bytes > 0 && (!c.ws || c.ws_bytes < r.ws_bytes || ((uintptr_t)c.ws & 255) != 0))
            return fail(MAXMULSK_ERR_ALLOC, "workspace null, short (%zu < %zu) or misaligned",
                        c.ws_bytes, r.ws_bytes);
    }

    // The pool lock covers what the call takes from the device pools: the
    // workspace (until the launch and its release event are enqueued) and, for
    // host operands, the staging and bounce buffers (until C is back on the
    // host). A synchronous call on device memory waits for its stream without
    // it, so callers on different streams overlap on the GPU.
    DevicePools& pl = pools(dev);
    std::unique_lock<std::mutex> lock(pl.mu);

    const float* A = c.A;
    const float* B = c.B;
    float* C = c.C;
    if (staged) {
        Scope scope("maxmulsk/stage_in");
        const size_t a_bytes = c.M * c.K * sizeof(float), b_bytes = c.K * c.N * sizeof(float),
                     c_bytes = c.M * c.N * sizeof(float);
        auto up = [](size_t x) { return (x + 255) / 256 * 256; };
        const size_t need = (a_dev ? 0 : up(a_bytes)) + (b_dev ? 0 : up(b_bytes)) + (c_dev ? 0 : up(c_bytes));
        if (!grow(&pl.stage, &pl.stage_cap, need)) return fail(MAXMULSK_ERR_ALLOC, "staging allocation of %zu bytes failed", need);
        char* s = static_cast<char*>(pl.stage);
        if (!a_dev) {
            if (stage_in(pl, s, c.A, a_bytes, c.stream) != hipSuccess) {
                (void)hipGetLastError();
                return fail(MAXMULSK_ERR_DEVICE, "H2D copy of A failed");
            }
            count(kCntStageIn);
            A = reinterpret_cast<const float*>(s);
            s += up(a_bytes);
        }
        if (!b_dev) {
            if (stage_in(pl, s, c.B, b_bytes, c.stream) != hipSuccess) {
                (void)hipGetLastError();
                return fail(MAXMULSK_ERR_DEVICE, "H2D copy of B failed");
            }
            count(kCntStageIn);
            B = reinterpret_cast<const float*>(s);
            s += up(b_bytes);
        }
        if (!c_dev) C = reinterpret_cast<float*>(s);
    }

    void* ws = c.ws;
    const bool pooled = r.ws_bytes > 0 && !ws;
    if (pooled) {
        const maxmulsk_status st = pool_workspace_acquire(pl, r.ws_bytes, c.stream, &ws);
        if (st != MAXMULSK_OK) return st;
    }

    hipError_t e;
    {
        Scope scope("maxmulsk/launch");
        e = launch_resolved(A, B, C, c.M, c.K, c.N, c.mode, r, ws, c.stream);
    }
    if (pooled) {
        // Also after a failed launch: whatever was enqueued before it must be
        // covered for the next user of the workspace.
        const hipError_t er = pool_workspace_release(pl, c.stream);
        if (e == hipSuccess) e = er;
    }
    if (e == hipSuccess && staged && !c_dev) {
        Scope scope("maxmulsk/writeback");
        e = stage_out(pl, c.C, C, c.M * c.N * sizeof(float), c.stream);
        count(kCntWriteback);
    }
    if (e == hipSuccess && staged && c_dev) {
        // Device C with staged operands: the staging buffer stays in use until
        // the GEMM has run.
        Scope scope("maxmulsk/sync");
        e = hipStreamSynchronize(c.stream);
    }
    lock.unlock();
    if (e == hipSuccess && !c.async) {
        Scope scope("maxmulsk/sync");
        e = hipStreamSynchronize(c.stream);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "GEMM launch failed: %s", hipGetErrorString(e));
    }
    return MAXMULSK_OK;
}

}  // namespace mmsk
