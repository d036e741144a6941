// MFMA throughput probe (introspection, not on the GEMM path): every wave
// issues back-to-back MFMAs into eight independent accumulators, with operands
// held in registers and no other work in the loop, so nothing but the matrix
// cores limits the rate. With random operands the eight MFMAs of an iteration
// take eight different operand pairs of random bits (full-entropy mantissas and
// signs, exponents in a narrow finite range), so the operand bits change on
// every instruction; with constant operands all eight use one pair. The gap
// between the two rates is what the data values themselves cost (power and
// clock); the random rate is the ceiling the GEMM kernels meet on random data.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gemm_tile.hpp"
#include "kernels.hpp"
#include "maxmulsk.h"
#include "plan.hpp"
#include "runtime.hpp"

namespace mmsk {
namespace {

constexpr int kProbeThreads = 256;     // four waves per workgroup
constexpr int kProbeBlocksPerCU = 2;   // eight waves per CU, two per SIMD

__device__ __forceinline__ uint32_t xorshift(uint32_t& s) {
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}

// A BF16 pair with random signs and mantissas and exponents 120..127.
__device__ __forceinline__ uint32_t bf16_pair_bits(uint32_t r) { return (r & 0x83ff83ffu) | 0x3c003c00u; }
// An FP32 value with a random sign and mantissa and exponent 120..127.
__device__ __forceinline__ float f32_bits(uint32_t r) { return __uint_as_float((r & 0x83ffffffu) | 0x3c000000u); }

template <int MODE, bool RANDOM>
__global__ __launch_bounds__(kProbeThreads) void mfma_probe_kernel(int iters, float* out) {
    constexpr int NACC = 8;
    uint32_t s = 0x9e3779b9u * (uint32_t)(blockIdx.x * kProbeThreads + threadIdx.x + 1);
    f32x16 acc[NACC];
#pragma unroll
    for (int j = 0; j < NACC; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[j][e] = 0.f;
    if constexpr (MODE == 1) {
        bf16x8 a[NACC], b[NACC];
#pragma unroll
        for (int j = 0; j < NACC; ++j) {
            if (RANDOM || j == 0) {
                u32x4 x, y;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    x[q] = bf16_pair_bits(xorshift(s));
                    y[q] = bf16_pair_bits(xorshift(s));
                }
                a[j] = __builtin_bit_cast(bf16x8, x);
                b[j] = __builtin_bit_cast(bf16x8, y);
            } else {
                a[j] = a[0];
                b[j] = b[0];
            }
        }
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int j = 0; j < NACC; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], b[j], acc[j], 0, 0, 0);
    } else {
        float a[NACC], b[NACC];
#pragma unroll
        for (int j = 0; j < NACC; ++j) {
            a[j] = (RANDOM || j == 0) ? f32_bits(xorshift(s)) : a[0];
            b[j] = (RANDOM || j == 0) ? f32_bits(xorshift(s)) : b[0];
        }
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int j = 0; j < NACC; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[j], b[j], acc[j], 0, 0, 0);
    }
    float t = 0.f;
#pragma unroll
    for (int j = 0; j < NACC; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) t += acc[j][e];
    out[blockIdx.x * kProbeThreads + threadIdx.x] = t;
}

// Streaming floor of a direct-tile launch: the GEMM's grid (tiles x K slices,
// same XCD remap and tile order), each workgroup reading exactly the A rows and
// B columns of its BM x BN tile over its K slice with dwordx4 loads, summing
// them, and writing its block of split-K partials; no split, LDS or MFMA. The
// empty kernel of the same grid gives the launch floor. K and N must be
// multiples of 4 (the loads are dwordx4). The unit loops step their (row,
// chunk) position instead of dividing.
template <int BM, int BN>
__global__ __launch_bounds__(256) void stream_probe_kernel(GemmParams p) {
    const int ntiles = p.tiles_m * p.tiles_n;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int split = bid / ntiles;
    int tm, tn;
    grouped_tile(bid - split * ntiles, p.tiles_m, p.tiles_n, tm, tn);
    const long long kbeg = split * p.k_len;
    const long long kend = min(p.K, kbeg + p.k_len);
    const long long m0 = (long long)tm * BM, n0 = (long long)tn * BN;
    const int kq = (int)((kend - kbeg) / 4);  // float4 per A row of the slice
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    {
        // A: BM rows x kq chunks, unit u = row * kq + chunk, u = tid + 256 i.
        // The loop runs on u (a trip count the compiler can unroll against,
        // so the loads issue back to back); (row, chunk) steps along. Rows
        // past M re-read row M - 1 (the sum is discarded), so every load is
        // unconditional.
        int r = (int)threadIdx.x / kq, c = (int)threadIdx.x % kq;
        const int dr = 256 / kq, dc = 256 % kq;
#pragma unroll 8
        for (int u = threadIdx.x; u < BM * kq; u += 256) {
            const long long m = min(m0 + r, p.M - 1);
            acc += *reinterpret_cast<const f32x4u*>(p.A + m * p.lda + kbeg + 4 * c);
            r += dr;
            c += dc;
            const bool wrap = c >= kq;
            c -= wrap ? kq : 0;
            r += wrap ? 1 : 0;
        }
    }
    constexpr int NQ = BN / 4;  // float4 per B row of the tile
    const int krows = (int)(kend - kbeg);
    const long long n = n0 + 4 * ((int)threadIdx.x % NQ);
    if (n < p.N) {
#pragma unroll 8
        for (int k = (int)threadIdx.x / NQ; k < krows; k += 256 / NQ)
            acc += *reinterpret_cast<const f32x4u*>(p.B + (kbeg + k) * p.ldb + n);
    }
    float* dst = p.C + split * p.c_split_stride;
    for (int u = threadIdx.x; u < BM * NQ; u += 256) {
        const long long m = m0 + u / NQ, nn = n0 + 4 * (u % NQ);
        if (m < p.M && nn < p.N) *reinterpret_cast<f32x4u*>(dst + m * p.ldc + nn) = acc;
    }
}

__global__ __launch_bounds__(256) void empty_probe_kernel() {}

template <int MODE, bool RANDOM>
hipError_t launch_probe(int blocks, int iters, float* out, hipStream_t s) {
    hipLaunchKernelGGL((mfma_probe_kernel<MODE, RANDOM>), dim3(blocks), dim3(kProbeThreads), 0, s, iters, out);
    return hipGetLastError();
}

}  // namespace
}  // namespace mmsk

using namespace mmsk;

extern "C" maxmulsk_status maxmulsk_probe_mfma(int mode, int random_operands, int iterations, double* tflops) {
    if ((mode != 0 && mode != 1) || iterations <= 0 || !tflops)
        return fail(MAXMULSK_ERR_UNSUPPORTED, "probe: mode 0 or 1, iterations > 0, tflops not null");
    int dev = -1;
    if (usable_devices() == 0 || hipGetDevice(&dev) != hipSuccess || !device_usable(dev)) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "device properties");
    }
    const int blocks = prop.multiProcessorCount * kProbeBlocksPerCU;
    float* out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = device_malloc(&out, (size_t)blocks * kProbeThreads * sizeof(float));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto launch = [&](int it) {
        if (mode == 1) return random_operands ? launch_probe<1, true>(blocks, it, out, s)
                                              : launch_probe<1, false>(blocks, it, out, s);
        return random_operands ? launch_probe<0, true>(blocks, it, out, s) : launch_probe<0, false>(blocks, it, out, s);
    };
    float ms = 0.f;
    if (e == hipSuccess) e = launch(iterations / 8 + 1);  // warm-up: clocks up, code loaded
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    if (e == hipSuccess) e = launch(iterations);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    if (s) (void)hipStreamDestroy(s);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "probe launch failed: %s", hipGetErrorString(e));
    }
    const double flop_per_mfma = mode == 1 ? 2.0 * 32 * 32 * 16 : 2.0 * 32 * 32 * 2;
    const double waves = (double)blocks * (kProbeThreads / 64);
    *tflops = waves * 8.0 * iterations * flop_per_mfma / (ms * 1e-3) / 1e12;
    return MAXMULSK_OK;
}

extern "C" maxmulsk_status maxmulsk_probe_stream(size_t M, size_t K, size_t N, const maxmulsk_plan* plan, int calls,
                                                 void* stream, double* us) {
    Resolved r{};
    if (!plan || !us || calls <= 0 || K % 4 != 0 || N % 4 != 0 ||
        !resolve(M, K, N, MAXMULSK_MODE_SPLIT3, *plan, &r) || r.cfg == kPackedCfg)
        return fail(MAXMULSK_ERR_UNSUPPORTED, "probe_stream: direct-tile plan, K and N multiples of 4, calls > 0");
    int dev = -1;
    if (usable_devices() == 0 || hipGetDevice(&dev) != hipSuccess || !device_usable(dev)) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");
    }
    const size_t mn = M * N;
    float *A = nullptr, *B = nullptr, *P = nullptr, *C = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipError_t e = device_malloc(&A, M * K * sizeof(float));
    if (e == hipSuccess) e = device_malloc(&B, K * N * sizeof(float));
    if (e == hipSuccess) e = device_malloc(&P, (size_t)r.splits * mn * sizeof(float));
    if (e == hipSuccess) e = device_malloc(&C, mn * sizeof(float));
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
    if (e == hipSuccess) e = hipMemsetAsync(A, 0, M * K * sizeof(float), s);
    if (e == hipSuccess) e = hipMemsetAsync(B, 0, K * N * sizeof(float), s);
    GemmParams p{};
    p.A = A;
    p.B = B;
    p.C = r.splits > 1 ? P : C;
    p.M = (long long)M;
    p.N = (long long)N;
    p.K = (long long)K;
    p.lda = (long long)K;
    p.ldb = (long long)N;
    p.ldc = (long long)N;
    p.k_len = r.k_len;
    p.c_split_stride = (long long)mn;
    p.splits = r.splits;
    p.tiles_m = r.tiles_m;
    p.tiles_n = r.tiles_n;
    const unsigned grid = (unsigned)((long long)r.tiles_m * r.tiles_n * r.splits);
    auto empty = [&]() {
        hipLaunchKernelGGL(empty_probe_kernel, dim3(grid), dim3(256), 0, s);
        return hipGetLastError();
    };
    auto streaming = [&]() {
        switch (r.cfg) {
            case 0: hipLaunchKernelGGL((stream_probe_kernel<64, 64>), dim3(grid), dim3(256), 0, s, p); break;
            case 1: hipLaunchKernelGGL((stream_probe_kernel<64, 128>), dim3(grid), dim3(256), 0, s, p); break;
            case 2: hipLaunchKernelGGL((stream_probe_kernel<64, 256>), dim3(grid), dim3(256), 0, s, p); break;
            case 3: hipLaunchKernelGGL((stream_probe_kernel<128, 128>), dim3(grid), dim3(256), 0, s, p); break;
            default: hipLaunchKernelGGL((stream_probe_kernel<128, 256>), dim3(grid), dim3(256), 0, s, p); break;
        }
        return hipGetLastError();
    };
    auto stream_reduce = [&]() {
        hipError_t x = streaming();
        if (x == hipSuccess && r.splits > 1) x = launch_splitk_reduce(P, C, (long long)mn, r.splits, (long long)mn, s);
        return x;
    };
    float ms[3] = {0.f, 0.f, 0.f};
    auto timed = [&](auto&& fn, float& out) {
        for (int i = 0; i < 3 && e == hipSuccess; ++i) e = fn();  // warm-up
        if (e == hipSuccess) e = hipEventRecord(ev[0], s);
        for (int i = 0; i < calls && e == hipSuccess; ++i) e = fn();
        if (e == hipSuccess) e = hipEventRecord(ev[1], s);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&out, ev[0], ev[1]);
    };
    timed(empty, ms[0]);
    timed(streaming, ms[1]);
    timed(stream_reduce, ms[2]);
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (float* x : {A, B, P, C})
        if (x) (void)hipFree(x);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "probe_stream failed: %s", hipGetErrorString(e));
    }
    for (int i = 0; i < 3; ++i) us[i] = 1e3 * ms[i] / calls;
    return MAXMULSK_OK;
}
