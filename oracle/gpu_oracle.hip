/*
 * ORACLE — test infrastructure only. Nothing in the product library links,
 * loads or calls this file; only tests/ use it, as the checker.
 *
 * FP64 GEMM ground truth on the GPU, for checking every element of C at sizes
 * where the CPU oracle (oracle.c / numpy dgemm) would take minutes. It computes
 * the same two products as oracle_gemm_f64 and oracle_absgemm_f64 (the
 * reference checks against cblas_dgemm, docs/BENCHMARKS.md:112-121): every
 * FP32 input is widened exactly to FP64 and accumulated with FP64 FMAs, so the
 * result differs from the CPU oracle only by FP64 rounding
 * (<= 2 K 2^-53 |A||B|, pinned in tests/test_gpu_oracle.py against the CPU
 * oracle on whole small products and on sampled rows and columns of large
 * ones).
 *
 * In the same pass it compares a candidate FP32 C with that ground truth and
 * reduces the error figures of oracle.py's errors() per workgroup:
 *   bound_ratio = max |C - C64| / (K 2^-24 |A||B|)
 *   max_rel     = max |C - C64| / max(|C64|, 2^-24 |A||B|)
 *   rel_frob    = ||C - C64||_F / ||C64||_F   (check_nb.cpp:13-17)
 *   max_abs     = max |C - C64|
 *   nonfinite   = count of non-finite C entries (their ratios count as +inf)
 * The host sums the per-workgroup partials in workgroup order.
 *
 * Plain LDS-tiled FP64 FMA: 64 x 64 outputs per 256-thread workgroup, 4 x 4 per
 * thread, k-tiles of 16. Speed is not the point; exactness of the widening and
 * of the bounds is.
 */
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstddef>
#include <vector>

namespace {

constexpr int TM = 64, TN = 64, TK = 16, NT = 256;
constexpr double U32 = 1.0 / 16777216.0;  // 2^-24

struct Part {
    double ratio, rel, abs, d2, c2, nonfinite;
};

__device__ double warp_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
__device__ double warp_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// C64 / ABS may be null (not stored); Cg may be null (no comparison, part unused).
__global__ __launch_bounds__(NT) void f64_gemm_check_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                            const float* __restrict__ Cg, double* __restrict__ C64,
                                                            double* __restrict__ ABS, long long M, long long N,
                                                            long long K, Part* __restrict__ part) {
    __shared__ float sA[TK][TM];
    __shared__ float sB[TK][TN];
    __shared__ double red[NT / 64][6];
    const int tid = threadIdx.x;
    const int tx = tid % 16, ty = tid / 16;
    const long long m0 = (long long)blockIdx.y * TM, n0 = (long long)blockIdx.x * TN;
    double c[4][4], ab[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) c[i][j] = ab[i][j] = 0.0;

    for (long long k0 = 0; k0 < K; k0 += TK) {
        for (int i = 0; i < (TM * TK) / NT; ++i) {
            const int idx = tid + NT * i;
            const int r = idx / TK, kk = idx % TK;
            const long long gm = m0 + r, gk = k0 + kk;
            sA[kk][r] = (gm < M && gk < K) ? A[gm * K + gk] : 0.f;
        }
        for (int i = 0; i < (TN * TK) / NT; ++i) {
            const int idx = tid + NT * i;
            const int kk = idx / TN, cc = idx % TN;
            const long long gk = k0 + kk, gn = n0 + cc;
            sB[kk][cc] = (gk < K && gn < N) ? B[gk * N + gn] : 0.f;
        }
        __syncthreads();
        for (int kk = 0; kk < TK; ++kk) {
            double av[4], bv[4];
            for (int i = 0; i < 4; ++i) av[i] = (double)sA[kk][ty + 16 * i];
            for (int j = 0; j < 4; ++j) bv[j] = (double)sB[kk][tx + 16 * j];
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    c[i][j] = fma(av[i], bv[j], c[i][j]);
                    ab[i][j] = fma(fabs(av[i]), fabs(bv[j]), ab[i][j]);
                }
        }
        __syncthreads();
    }

    double p_ratio = 0.0, p_rel = 0.0, p_abs = 0.0, p_d2 = 0.0, p_c2 = 0.0, p_nf = 0.0;
    for (int i = 0; i < 4; ++i) {
        const long long m = m0 + ty + 16 * i;
        if (m >= M) continue;
        for (int j = 0; j < 4; ++j) {
            const long long n = n0 + tx + 16 * j;
            if (n >= N) continue;
            const long long e = m * N + n;
            if (C64) C64[e] = c[i][j];
            if (ABS) ABS[e] = ab[i][j];
            if (!Cg) continue;
            const double g = (double)Cg[e];
            const double scale = fmax(ab[i][j], DBL_MIN);
            p_c2 += c[i][j] * c[i][j];
            if (!isfinite(g)) {
                p_nf += 1.0;
                p_ratio = p_rel = p_abs = INFINITY;
                continue;
            }
            const double d = fabs(g - c[i][j]);
            p_ratio = fmax(p_ratio, d / ((double)K * U32 * scale));
            p_rel = fmax(p_rel, d / fmax(fabs(c[i][j]), U32 * scale));
            p_abs = fmax(p_abs, d);
            p_d2 += d * d;
        }
    }
    if (!Cg) return;
    p_ratio = warp_max(p_ratio);
    p_rel = warp_max(p_rel);
    p_abs = warp_max(p_abs);
    p_d2 = warp_sum(p_d2);
    p_c2 = warp_sum(p_c2);
    p_nf = warp_sum(p_nf);
    const int wave = tid / 64;
    if ((tid & 63) == 0) {
        red[wave][0] = p_ratio;
        red[wave][1] = p_rel;
        red[wave][2] = p_abs;
        red[wave][3] = p_d2;
        red[wave][4] = p_c2;
        red[wave][5] = p_nf;
    }
    __syncthreads();
    if (tid == 0) {
        Part q = {0, 0, 0, 0, 0, 0};
        for (int w = 0; w < NT / 64; ++w) {
            q.ratio = fmax(q.ratio, red[w][0]);
            q.rel = fmax(q.rel, red[w][1]);
            q.abs = fmax(q.abs, red[w][2]);
            q.d2 += red[w][3];
            q.c2 += red[w][4];
            q.nonfinite += red[w][5];
        }
        part[(long long)blockIdx.y * gridDim.x + blockIdx.x] = q;
    }
}

int launch(const float* A, const float* B, const float* Cg, double* C64, double* ABS, size_t M, size_t K,
           size_t N, double* out) {
    if (M == 0 || N == 0 || K == 0) return 1;
    const size_t tm = (M + TM - 1) / TM, tn = (N + TN - 1) / TN;
    if (tm > 65535 || tn > 0x7fffffff) return 1;
    Part* part = nullptr;
    const size_t nb = tm * tn;
    if (Cg && hipMalloc(&part, nb * sizeof(Part)) != hipSuccess) return 2;
    hipLaunchKernelGGL(f64_gemm_check_kernel, dim3((unsigned)tn, (unsigned)tm), dim3(NT), 0, 0, A, B, Cg, C64,
                       ABS, (long long)M, (long long)N, (long long)K, part);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && Cg) {
        std::vector<Part> h(nb);
        e = hipMemcpy(h.data(), part, nb * sizeof(Part), hipMemcpyDeviceToHost);
        Part q = {0, 0, 0, 0, 0, 0};
        for (const Part& p : h) {
            q.ratio = fmax(q.ratio, p.ratio);
            q.rel = fmax(q.rel, p.rel);
            q.abs = fmax(q.abs, p.abs);
            q.d2 += p.d2;
            q.c2 += p.c2;
            q.nonfinite += p.nonfinite;
        }
        out[0] = q.ratio;
        out[1] = q.rel;
        out[2] = q.c2 > 0 ? std::sqrt(q.d2) / std::sqrt(q.c2) : 0.0;
        out[3] = q.abs;
        out[4] = q.nonfinite;
    }
    if (part) (void)hipFree(part);
    return e == hipSuccess ? 0 : 3;
}

}  // namespace

extern "C" {

/* Errors of the FP32 product C (device, M x N) against FP64 A*B (device A
 * M x K, B K x N, all row-major): out[0..4] = bound_ratio, max_rel, rel_frob,
 * max_abs, nonfinite. Synchronous. 0 = ok, 1 = bad shape, 2 = allocation,
 * 3 = HIP error. */
int oracle_gpu_check(const float* A, const float* B, const float* C, size_t M, size_t K, size_t N, double* out) {
    return launch(A, B, C, nullptr, nullptr, M, K, N, out);
}

/* C64 = A*B and ABS = |A|*|B| in FP64 (device buffers, M x N each; either may
 * be null). Synchronous. */
int oracle_gpu_gemm_f64(const float* A, const float* B, double* C64, double* ABS, size_t M, size_t K, size_t N) {
    return launch(A, B, nullptr, C64, ABS, M, K, N, nullptr);
}

}  // extern "C"
