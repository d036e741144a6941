// One-tile-per-workgroup GEMM kernels for every (mode, tile geometry), the
// split-K reduce, and their host launchers.
#include <hip/hip_runtime.h>

#include "gemm_tile.hpp"
#include "kernels.hpp"
#include "splitk_sum.hpp"

namespace mmsk {

template <int MODE, int BM, int BN, int BK, int WGM, int WGN>
__global__ __launch_bounds__(64 * WGM * WGN) void gemm_kernel(GemmParams p) {
    using T = Tile<MODE, BM, BN, BK, WGM, WGN>;
    __shared__ __attribute__((aligned(16))) char smem[T::SMEM];
    const int ntiles = p.tiles_m * p.tiles_n;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int split = bid / ntiles;
    int tm, tn;
    grouped_tile(bid - split * ntiles, p.tiles_m, p.tiles_n, tm, tn);
    const long long kbeg = split * p.k_len;
    const long long kend = min(p.K, kbeg + p.k_len);
    T t;
    t.run(p.A, p.B, p.C + split * p.c_split_stride, p.M, p.N, p.lda, p.ldb, p.ldc,
          (long long)tm * BM, (long long)tn * BN, kbeg, kend, smem);
}

// Sums the split-K partials in slice order (deterministic).
template <int NT>
__global__ __launch_bounds__(NT) void splitk_reduce_kernel(const float* __restrict__ P,
                                                           float* __restrict__ C, long long mn,
                                                           int splits, long long stride) {
    const long long step = (long long)gridDim.x * NT * 4;
    for (long long i = ((long long)blockIdx.x * NT + threadIdx.x) * 4; i < mn; i += step) {
        if (i + 3 < mn) {
            *reinterpret_cast<f32x4u*>(C + i) = sum_splits4(P + i, stride, splits);
        } else {
            for (long long j = i; j < i + 4 && j < mn; ++j) {
                float s = P[j];
                for (int k = 1; k < splits; ++k) s += P[k * stride + j];
                C[j] = s;
            }
        }
    }
}

template <int MODE, int CFG>
static hipError_t launch_one(const GemmParams& p, hipStream_t stream) {
    constexpr TileCfg c = kTileCfgs[CFG];
    constexpr int BK = c.bk[MODE];
    constexpr int WGM = direct_wgm(MODE, CFG), WGN = direct_wgn(MODE, CFG);
    const long long grid = (long long)p.tiles_m * p.tiles_n * p.splits;
    if (grid <= 0 || grid > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((gemm_kernel<MODE, c.bm, c.bn, BK, WGM, WGN>), dim3((unsigned)grid),
                       dim3(64 * WGM * WGN), 0, stream, p);
    return hipGetLastError();
}

template <int MODE, int BK>
static hipError_t launch_deep(const GemmParams& p, hipStream_t stream) {
    constexpr TileCfg c = kTileCfgs[0];
    const long long grid = (long long)p.tiles_m * p.tiles_n * p.splits;
    hipLaunchKernelGGL((gemm_kernel<MODE, c.bm, c.bn, BK, c.wgm, c.wgn>), dim3((unsigned)grid),
                       dim3(64 * c.wgm * c.wgn), 0, stream, p);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_mode(int cfg, const GemmParams& p, hipStream_t stream) {
    switch (cfg) {
        case 0:
            if (const int fv = feed_variant(MODE, 0, p.M, p.N, p.K, (long long)p.tiles_m * p.tiles_n * p.splits,
                                            p.k_len, feed_mode()))
                return launch_gemm_feed(fv, p, stream);
            if constexpr (kDeepBK[MODE] > 0) {
                const long long grid = (long long)p.tiles_m * p.tiles_n * p.splits;
                if (grid > 0 && grid <= kDeepMaxGrid) return launch_deep<MODE, kDeepBK[MODE]>(p, stream);
                if constexpr (kDeep2BK[MODE] > 0)
                    if (grid > 0 && grid <= 2 * kDeepMaxGrid) return launch_deep<MODE, kDeep2BK[MODE]>(p, stream);
            }
            return launch_one<MODE, 0>(p, stream);
        case 1: return launch_one<MODE, 1>(p, stream);
        case 2: return launch_one<MODE, 2>(p, stream);
        case 3:
            if (const int fv = feed_variant(MODE, 3, p.M, p.N, p.K, (long long)p.tiles_m * p.tiles_n * p.splits,
                                            p.k_len, feed_mode()))
                return launch_gemm_feed(fv, p, stream);
            if constexpr (kDeep128BK[MODE] > 0) {
                const long long grid = (long long)p.tiles_m * p.tiles_n * p.splits;
                if (grid > 0 && grid <= kDeepMaxGrid) {
                    constexpr int WGM = direct_wgm(MODE, 3), WGN = direct_wgn(MODE, 3);
                    hipLaunchKernelGGL((gemm_kernel<MODE, 128, 128, kDeep128BK[MODE], WGM, WGN>), dim3((unsigned)grid),
                                       dim3(64 * WGM * WGN), 0, stream, p);
                    return hipGetLastError();
                }
            }
            return launch_one<MODE, 3>(p, stream);
        case 4: return launch_one<MODE, 4>(p, stream);
        case 5:
            // 256 x 256 in the split modes on whole tiles and k-blocks
            // (packed_path); every other call of geometry 5 runs packed.
            if constexpr (MODE != 0) return launch_one<MODE, 5>(p, stream);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gemm(int mode, int cfg, const GemmParams& p, hipStream_t stream) {
    switch (mode) {
        case 0: return launch_mode<0>(cfg, p, stream);
        case 1: return launch_mode<1>(cfg, p, stream);
        case 2: return launch_mode<2>(cfg, p, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_splitk_reduce(const float* P, float* C, long long mn, int splits,
                                long long stride, hipStream_t stream) {
    // Outputs too small to give every CU a 256-thread workgroup run one-wave
    // workgroups (the same four values per thread, the same slice order):
    // more CUs pull partials at once. Measured 64 x 512 x 8192 split3 10.4 ->
    // 10.0 us per call, 128 x 512 x 32768 27.4 -> 27.0 us.
    long long blocks = (mn + 1023) / 1024;
    if (blocks < 256) {
        blocks = (mn + 255) / 256;
        if (blocks < 1) blocks = 1;
        hipLaunchKernelGGL(splitk_reduce_kernel<64>, dim3((unsigned)blocks), dim3(64), 0, stream, P, C, mn,
                           splits, stride);
        return hipGetLastError();
    }
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(splitk_reduce_kernel<256>, dim3((unsigned)blocks), dim3(256), 0, stream, P, C, mn,
                       splits, stride);
    return hipGetLastError();
}

}  // namespace mmsk

namespace mmsk {

__global__ __launch_bounds__(256) void fill_zero_u32_kernel(unsigned int* p, size_t n) {
    const size_t step = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += step) p[i] = 0u;
}

// Graph-safe zero fill (a kernel node instead of a memset node).
hipError_t launch_fill_zero_u32(unsigned int* p, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_zero_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launch_fill_zero_f32(float* p, size_t n, hipStream_t stream) {
    return launch_fill_zero_u32(reinterpret_cast<unsigned int*>(p), n, stream);
}

__global__ __launch_bounds__(256) void gather_strided_kernel(const float* __restrict__ src, size_t stride,
                                                             size_t count, float* __restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[i] = src[i * stride];
}

hipError_t launch_gather_strided(const float* src, size_t stride, size_t count, float* dst,
                                 hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_strided_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, src,
                       stride, count, dst);
    return hipGetLastError();
}

}  // namespace mmsk

namespace mmsk {

// One wave per sampled output: lanes take k = lane, lane + 64, ... and the
// partial sums are combined by a fixed butterfly, all in FP64.
__global__ __launch_bounds__(256) void sample_dot_f64_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                             size_t K, size_t N, size_t stride, size_t count,
                                                             double* __restrict__ ref, double* __restrict__ absref) {
    const size_t i = (size_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (i >= count) return;
    const size_t idx = i * stride, r = idx / N, c = idx % N;
    const float* a = A + r * K;
    double s = 0.0, t = 0.0;
    for (size_t k = lane; k < K; k += 64) {
        const double x = (double)a[k], y = (double)B[k * N + c];
        s = fma(x, y, s);
        t = fma(fabs(x), fabs(y), t);
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        t += __shfl_xor(t, o);
    }
    if (lane == 0) {
        ref[i] = s;
        absref[i] = t;
    }
}

hipError_t launch_sample_dot_f64(const float* A, const float* B, size_t K, size_t N, size_t stride,
                                 size_t count, double* ref, double* absref, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(sample_dot_f64_kernel, dim3((unsigned)((count + 3) / 4)), dim3(256), 0, stream, A, B, K, N,
                       stride, count, ref, absref);
    return hipGetLastError();
}

}  // namespace mmsk
