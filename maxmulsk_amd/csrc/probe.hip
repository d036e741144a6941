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
#include "maxmulsk.h"
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
