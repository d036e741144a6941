// maxmulsk_sgemm_noalloc: maxmulsk_sgemm_ws with the workspace in a static
// device array instead of caller or pool memory, for callers that cannot
// allocate (code under stream capture, tight serving loops, harnesses that
// forbid hipMalloc). No counterpart in the reference.
//
// The array is part of this library's code object: every GPU the library runs
// on holds its own copy, at its own address, until the process ends. Host code
// never uses the array's name as an address (on the host it is only a
// stand-in): the address is looked up once per device with
// hipGetSymbolAddress, with that device current, and that pointer is what the
// kernels receive. The workspace holds split-K partials and packed operands,
// which every call overwrites before reading, so nothing has to be cleared
// between calls. All calls on one GPU share the one array: calls that overlap
// on a GPU (different streams, or asynchronous use) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "maxmulsk.h"
#include "runtime.hpp"

static __device__ __align__(256) unsigned char scratch[MAXMULSK_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on device dev (looked up on the first call on
// that device, with the device current).
maxmulsk_status scratch_ptr(int dev, void** out) {
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        mmsk::DeviceGuard guard(dev);
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return mmsk::fail(MAXMULSK_ERR_DEVICE, "hipGetSymbolAddress of the static scratch failed on device %d: %s",
                              dev, hipGetErrorString(e));
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return MAXMULSK_OK;
}

}  // namespace

extern "C" maxmulsk_status maxmulsk_sgemm_noalloc(const float* A, const float* B, float* C, size_t M, size_t K,
                                                  size_t N, maxmulsk_mode mode, const maxmulsk_plan* plan,
                                                  void* stream) {
    if (M == 0 || N == 0) return MAXMULSK_OK;  // capi/maxmulsk.cpp:13-20, as every entry point
    int dev = -1;
    if (!C || !mmsk::is_device_ptr(C, &dev))
        return mmsk::fail(MAXMULSK_ERR_UNSUPPORTED, "maxmulsk_sgemm_noalloc needs device pointers");
    if (dev < 0 || dev >= kMaxDevices || !mmsk::device_usable(dev))
        return mmsk::fail(MAXMULSK_ERR_UNSUPPORTED, "device %d is not gfx950", dev);
    void* ws = nullptr;
    const maxmulsk_status st = scratch_ptr(dev, &ws);
    if (st != MAXMULSK_OK) return st;
    mmsk::GemmCall c{};
    c.A = A;
    c.B = B;
    c.C = C;
    c.M = M;
    c.K = K;
    c.N = N;
    c.mode = (int)mode;
    if (plan) {
        c.have_plan = true;
        c.plan = *plan;
    }
    c.stream = static_cast<hipStream_t>(stream);
    c.async = false;
    c.ws = ws;
    c.ws_bytes = MAXMULSK_NOALLOC_SCRATCH_BYTES;
    c.ws_required = true;
    c.no_alloc = true;
    return mmsk::run_gemm(c);
}
