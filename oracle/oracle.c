/*
 * ORACLE — test infrastructure only. Nothing in the product library links,
 * loads or calls this file; tests/, the smoke check and bench.py's CPU-baseline
 * leg use it as the checker.
 *
 * CPU restatement of what the reference's path computes and decides:
 *   - FP64 GEMM (the ground truth the reference checks against, cblas_dgemm in
 *     docs/BENCHMARKS.md:112-121) and |A|*|B| for the error bound;
 *   - the portable FP32 loop Utils::multiply_scalar (common/utils.hpp:50-61);
 *   - BF16 round-to-nearest-even and the hi/lo (hi/mid/lo) split, and the exact
 *     FP64 value of the 3- and 6-product split GEMM (no reference counterpart);
 *   - the reference's plan rules: SMEKernels1x4V4C::choose
 *     (sme/v4/sme-1x4-v4c-dispatch.cpp:17-22), tuning::select_from_table for the
 *     1x4 kernel (sme/support/gemm_tuning.cpp:75,91-94,138-151), and
 *     NB::classify / capacity / counts
 *     (sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:410-466).
 * Pinning of each function is documented in oracle/oracle.py.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

/* C64 = A * B in double (inputs FP32, row-major, A MxK, B KxN). */
void oracle_gemm_f64(const float* A, const float* B, double* C, size_t M, size_t K, size_t N) {
    for (size_t i = 0; i < M; ++i) {
        double* c = C + i * N;
        for (size_t j = 0; j < N; ++j) c[j] = 0.0;
        for (size_t k = 0; k < K; ++k) {
            const double a = (double)A[i * K + k];
            const float* b = B + k * N;
            for (size_t j = 0; j < N; ++j) c[j] += a * (double)b[j];
        }
    }
}

/* |A| * |B| in double: the scale of the FP32 rounding bound. */
void oracle_absgemm_f64(const float* A, const float* B, double* C, size_t M, size_t K, size_t N) {
    for (size_t i = 0; i < M; ++i) {
        double* c = C + i * N;
        for (size_t j = 0; j < N; ++j) c[j] = 0.0;
        for (size_t k = 0; k < K; ++k) {
            const double a = fabs((double)A[i * K + k]);
            const float* b = B + k * N;
            for (size_t j = 0; j < N; ++j) c[j] += a * fabs((double)b[j]);
        }
    }
}

/*
 * Utils::multiply_scalar (common/utils.hpp:50-61): ijk loop, one float
 * accumulator per output, C overwritten. Argument order (M, N, K) as there.
 * The reference builds it with clang, whose default -ffp-contract=on fuses
 * `sum += a * b` into one FMA; fmaf states that rounding explicitly.
 */
void oracle_sgemm_portable(const float* A, const float* B, float* C, size_t M, size_t N, size_t K) {
    for (size_t i = 0; i < M; ++i) {
        for (size_t j = 0; j < N; ++j) {
            float sum = 0.0f;
            for (size_t k = 0; k < K; ++k) sum = fmaf(A[i * K + k], B[k * N + j], sum);
            C[i * N + j] = sum;
        }
    }
}

/* Round FP32 to BF16 (nearest, ties to even), returned as the FP32 value. */
float oracle_bf16_rn(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) { /* inf / nan: truncate, keep nan quiet */
        if (u & 0x007fffffu) u |= 0x00400000u;
        u &= 0xffff0000u;
    } else {
        const uint32_t lsb = (u >> 16) & 1u;
        u += 0x7fffu + lsb;
        u &= 0xffff0000u;
    }
    float y;
    memcpy(&y, &u, 4);
    return y;
}

/* Split x[i] into `terms` BF16 values: t0 = RN(x), t1 = RN(x - t0),
 * t2 = RN(x - t0 - t1); out is terms x n (term-major). */
void oracle_split(const float* x, size_t n, int terms, float* out) {
    for (size_t i = 0; i < n; ++i) {
        float r = x[i];
        for (int t = 0; t < terms; ++t) {
            const float h = oracle_bf16_rn(r);
            out[(size_t)t * n + i] = h;
            r = r - h;
        }
    }
}

/*
 * Exact FP64 value of the split-precision product that the BF16 MFMA path
 * approximates: sum over the mode's cross products of term_a(i,k) * term_b(k,j)
 * (every BF16 x BF16 product is exact in FP32 and in FP64).
 * mode 1: hi*hi + hi*lo + lo*hi; mode 2: adds hi*mid, mid*hi, mid*mid and uses
 * the third term as lo (hi*lo + lo*hi).
 */
void oracle_split_gemm_f64(const float* At, const float* Bt, double* C, size_t M, size_t K,
                           size_t N, int mode) {
    /* At: terms x M x K, Bt: terms x K x N (from oracle_split). */
    static const int pa3[3] = {1, 0, 0}, pb3[3] = {0, 1, 0};
    static const int pa6[6] = {2, 0, 1, 1, 0, 0}, pb6[6] = {0, 2, 1, 0, 1, 0};
    const int np = mode == 1 ? 3 : 6;
    const int* pa = mode == 1 ? pa3 : pa6;
    const int* pb = mode == 1 ? pb3 : pb6;
    for (size_t i = 0; i < M * N; ++i) C[i] = 0.0;
    for (int p = 0; p < np; ++p) {
        const float* a = At + (size_t)pa[p] * M * K;
        const float* b = Bt + (size_t)pb[p] * K * N;
        for (size_t i = 0; i < M; ++i)
            for (size_t k = 0; k < K; ++k) {
                const double av = (double)a[i * K + k];
                for (size_t j = 0; j < N; ++j) C[i * N + j] += av * (double)b[k * N + j];
            }
    }
}

/* ---- the reference's plan rules ----------------------------------------- */

/* SMEKernels1x4V4C::choose: out = {Mc, Nc, Kc}. */
void oracle_v4c_choose(size_t M, size_t K, size_t N, size_t* out) {
    (void)K;
    if (M >= 2048 && N > 512) {
        out[0] = 32; out[1] = 1024; out[2] = 2048;
    } else {
        out[0] = 16; out[1] = 512; out[2] = 2048;
    }
}

/* tuning::select_from_table for the 1x4 kernel: out = {M_tile, N_tile, Kc, ir_outer}. */
void oracle_v3_select_from_table(size_t M, size_t K, size_t N, size_t* out) {
    out[0] = 1024; out[1] = 64; out[2] = 2048; out[3] = 0;
    if (M != K || K != N) return;
    if (M >= 4096) {
        out[0] = 2048; out[1] = 512; out[2] = 4096; out[3] = 1;
    } else if (M >= 2048) {
        out[0] = 1024; out[1] = 512; out[2] = 2048; out[3] = 1;
    }
}

/* NB::classify: 0 Native, 1 Unsupported (empty), 2 BadBlocking. */
int oracle_nb_classify(size_t M, size_t K, size_t N, size_t Mc, size_t Nc, size_t Kc) {
    if (M == 0 || K == 0 || N == 0) return 1;
    if (Mc == 0 || Mc % 16 != 0) return 2;
    if (Nc == 0 || Nc % 64 != 0) return 2;
    if (Kc == 0 || Kc % 64 != 0) return 2;
    return 0;
}

static size_t min_sz(size_t a, size_t b) { return a < b ? a : b; }

/* NB::capacity: packed_A and packed_B bytes after the clamps. */
void oracle_nb_capacity(size_t M, size_t K, size_t N, size_t Mc, size_t Nc, size_t Kc, size_t* out) {
    const size_t kc = min_sz(Kc, K);
    const size_t nc = min_sz(Nc, (N + 63) / 64 * 64);
    const size_t mc = min_sz(Mc, (M + 15) / 16 * 16);
    const size_t kc_pad = (kc + 63) / 64 * 64;
    out[0] = mc * kc_pad * 4;
    out[1] = nc * kc * 4;
}

/* NB::counts: {kc_slices, nc_blocks, mc_blocks, a_panel_packs, b_panel_packs,
 * microkernel_calls}. */
void oracle_nb_counts(size_t M, size_t K, size_t N, size_t Mc, size_t Nc, size_t Kc, size_t* out) {
    const size_t kc = min_sz(Kc, K);
    const size_t nc = min_sz(Nc, (N + 63) / 64 * 64);
    const size_t mc = min_sz(Mc, (M + 15) / 16 * 16);
    out[0] = (K + kc - 1) / kc;
    out[1] = (N + nc - 1) / nc;
    out[2] = (M + mc - 1) / mc;
    out[4] = out[0] * (N / 64);
    out[3] = out[0] * out[1] * (M / 16);
    out[5] = out[0] * (N / 64) * (M / 16);
}
