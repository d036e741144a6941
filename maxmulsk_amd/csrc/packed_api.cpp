// Pack-once operands: maxmulsk_pack / maxmulsk_sgemm_packed /
// maxmulsk_packed_destroy. An operand is rewritten once into the fragment
// layout of the 256 x 256 packed GEMM (BF16 terms in the split modes) and then
// feeds any number of products without the per-call pack pass.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>

#include "kernels.hpp"
#include "maxmulsk.h"
#include "plan.hpp"
#include "runtime.hpp"
#include "trace.hpp"

using namespace mmsk;

struct maxmulsk_packed_s {
    int dev;
    int mode;
    int operand;      // 0: A (rows = M), 1: B (rows = N)
    size_t rows, K;
    int rblocks;      // 32-row fragment blocks, a multiple of 8 (one 256-row tile)
    int kgroups;      // fragment k-groups, kblocks * (packed_bk(mode) / packed_kgk(mode))
    long long kblocks;
    long long plane;  // bytes of one term plane
    void* buf;
};

namespace {

constexpr int kTile = 256;

int terms_of(int mode) { return mode == 0 ? 1 : (mode == 1 ? 2 : 3); }

}  // namespace

extern "C" maxmulsk_status maxmulsk_pack(const float* X, size_t rows, size_t K, int operand,
                                         maxmulsk_mode mode, void* stream, maxmulsk_packed* out) {
    if (!out) return fail(MAXMULSK_ERR_UNSUPPORTED, "out is null");
    *out = nullptr;
    const int md = (int)mode;
    if (md < 0 || md > 2) return fail(MAXMULSK_ERR_UNSUPPORTED, "unknown mode %d", md);
    if (operand != 0 && operand != 1) return fail(MAXMULSK_ERR_UNSUPPORTED, "operand must be 0 (A) or 1 (B)");
    if (!X || rows == 0 || K == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "empty or null operand");
    if (rows > 0x7fffffffULL || K > 0x7fffffffULL) return fail(MAXMULSK_ERR_UNSUPPORTED, "shape too large");
    if (usable_devices() == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");
    int dev = -1;
    if (!is_device_ptr(X, &dev)) return fail(MAXMULSK_ERR_UNSUPPORTED, "packing needs a device pointer");
    if (!device_usable(dev)) return fail(MAXMULSK_ERR_UNSUPPORTED, "device %d is not gfx950", dev);

    const long long kblocks = ((long long)K + packed_bk(md) - 1) / packed_bk(md);
    const int kgroups = (int)(kblocks * (packed_bk(md) / packed_kgk(md)));
    const int rblocks = (int)(((long long)rows + kTile - 1) / kTile * (kTile / 32));
    const long long plane = 1024LL * kgroups * rblocks;
    const size_t bytes = (size_t)plane * terms_of(md);

    DeviceGuard guard(dev);
    void* buf = nullptr;
    if (device_malloc(&buf, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_ALLOC, "allocation of %zu packed bytes failed", bytes);
    }
    count(kCntAllocation);
    count(kCntPack);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // A is rows x K row-major (row stride K); B is K x rows row-major, read
    // with k as the slow index (row stride 1, k stride rows).
    hipError_t e = operand == 0
                       ? launch_pack(md, X, (long long)rows, (long long)K, (long long)K, 1, buf, rblocks, kgroups, s)
                       : launch_pack(md, X, (long long)rows, (long long)K, 1, (long long)rows, buf, rblocks, kgroups, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(buf);
        return fail(MAXMULSK_ERR_DEVICE, "pack launch failed: %s", hipGetErrorString(e));
    }
    auto* p = new (std::nothrow) maxmulsk_packed_s{dev, md, operand, rows, K, rblocks, kgroups, kblocks, plane, buf};
    if (!p) {
        (void)hipFree(buf);
        return fail(MAXMULSK_ERR_ALLOC, "handle allocation failed");
    }
    *out = p;
    return MAXMULSK_OK;
}

extern "C" maxmulsk_status maxmulsk_sgemm_packed(maxmulsk_packed A, maxmulsk_packed B, float* C,
                                                 const maxmulsk_options* opt) {
    if (!A || !B || !C) return fail(MAXMULSK_ERR_UNSUPPORTED, "null packed operand or C");
    if (A->operand != 0 || B->operand != 1) return fail(MAXMULSK_ERR_UNSUPPORTED, "operands are not (A, B)");
    if (A->mode != B->mode || A->K != B->K || A->dev != B->dev)
        return fail(MAXMULSK_ERR_UNSUPPORTED, "packed operands differ in mode, K or device");
    int cdev = -1;
    if (!is_device_ptr(C, &cdev) || cdev != A->dev)
        return fail(MAXMULSK_ERR_UNSUPPORTED, "C must be device memory on the operands' device");
    const size_t M = A->rows, N = B->rows;
    if (M * N > (size_t)1 << 40) return fail(MAXMULSK_ERR_UNSUPPORTED, "shape too large");
    int split = opt && opt->plan.split_k > 0 ? opt->plan.split_k : 1;
    if (split > kMaxSplit) return fail(MAXMULSK_ERR_UNSUPPORTED, "split_k %d above %d", split, kMaxSplit);
    const long long per = (A->kblocks + split - 1) / split;
    const int splits = (int)((A->kblocks + per - 1) / per);
    const int tiles_m = (int)((M + kTile - 1) / kTile), tiles_n = (int)((N + kTile - 1) / kTile);
    if ((long long)tiles_m * tiles_n * splits > 0x7fffffffLL) return fail(MAXMULSK_ERR_UNSUPPORTED, "grid too large");
    const size_t partial_bytes = splits > 1 ? ((size_t)splits * M * N * sizeof(float) + 255) / 256 * 256 : 0;
    hipStream_t s = opt ? static_cast<hipStream_t>(opt->stream) : nullptr;
    const bool async = opt && opt->async;

    void* ws = opt ? opt->workspace : nullptr;
    if (ws && partial_bytes > 0 && (opt->workspace_bytes < partial_bytes || ((uintptr_t)ws & 255) != 0))
        return fail(MAXMULSK_ERR_ALLOC, "workspace short (%zu < %zu) or misaligned", opt->workspace_bytes,
                    partial_bytes);

    DeviceGuard guard(A->dev);
    DevicePools& pl = pools(A->dev);
    std::unique_lock<std::mutex> lock(pl.mu);
    const bool pooled = partial_bytes > 0 && !ws;
    if (pooled) {
        const maxmulsk_status st = pool_workspace_acquire(pl, partial_bytes, s, &ws);
        if (st != MAXMULSK_OK) return st;
    }
    PackedParams p{};
    p.PA = static_cast<const char*>(A->buf);
    p.PB = static_cast<const char*>(B->buf);
    p.planeA = A->plane;
    p.planeB = B->plane;
    p.kgtA = A->kgroups;
    p.kgtB = B->kgroups;
    p.M = (long long)M;
    p.N = (long long)N;
    p.ldc = (long long)N;
    p.kblocks = A->kblocks;
    p.kblocks_per_split = per;
    p.splits = splits;
    p.tiles_m = tiles_m;
    p.tiles_n = tiles_n;
    float* partials = static_cast<float*>(ws);
    if (splits > 1) {
        p.C = partials;
        p.c_split_stride = (long long)(M * N);
    } else {
        p.C = C;
        p.c_split_stride = 0;
    }
    hipError_t e = launch_gemm_packed(A->mode, p, s);
    count(kCntCompute);
    if (e == hipSuccess && splits > 1) {
        e = launch_splitk_reduce(partials, C, (long long)(M * N), splits, (long long)(M * N), s);
        count(kCntReduce);
    }
    if (pooled) {
        // Also after a failed launch: whatever was enqueued before it must be
        // covered for the next user of the workspace.
        const hipError_t er = pool_workspace_release(pl, s);
        if (e == hipSuccess) e = er;
    }
    lock.unlock();  // the wait below needs nothing from the pools
    if (e == hipSuccess && !async) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "packed GEMM launch failed: %s", hipGetErrorString(e));
    }
    return MAXMULSK_OK;
}

extern "C" void maxmulsk_packed_destroy(maxmulsk_packed p) {
    if (!p) return;
    {
        DeviceGuard guard(p->dev);
        (void)hipDeviceSynchronize();
        (void)hipFree(p->buf);
    }
    delete p;
}
