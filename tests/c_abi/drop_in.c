/* A reference caller, in C, of the two symbols of capi/maxmulsk.h: built with
 * gcc against include/maxmulsk.h and linked with libmaxmulsk.so, with no other
 * dependency. Exit status 0 on success.
 *   drop_in contract  - empty shapes and a missing device answer like the
 *                       reference's C ABI (runs without a GPU)
 *   drop_in gemm      - host arrays through maxmulsk_sgemm_v4/v3, checked
 *                       against an FP64 loop within K * 2^-24 * |A||B|
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "maxmulsk.h"

static float frand(unsigned* s) {
    *s = *s * 1664525u + 1013904223u;
    return (float)((*s >> 8) & 0xffffff) / 16777216.0f * 2.0f - 1.0f;
}

static int contract(void) {
    float a[4] = {1, 2, 3, 4}, b[4] = {5, 6, 7, 8}, c[4] = {9, 9, 9, 9};
    /* M == 0 or N == 0: OK and C untouched (capi/maxmulsk.cpp:13-20) */
    if (maxmulsk_sgemm_v4(a, b, c, 0, 2, 2) != MAXMULSK_OK) return 1;
    if (maxmulsk_sgemm_v3(a, b, c, 2, 2, 0) != MAXMULSK_OK) return 2;
    for (int i = 0; i < 4; ++i)
        if (c[i] != 9.0f) return 3;
    /* K == 0: C becomes the zero matrix */
    if (maxmulsk_sgemm_v4(a, b, c, 2, 0, 2) != MAXMULSK_OK) return 4;
    for (int i = 0; i < 4; ++i)
        if (c[i] != 0.0f) return 5;
    /* without a gfx950 device a real product is refused, C untouched */
    if (maxmulsk_device_count() == 0) {
        c[0] = 7.0f;
        if (maxmulsk_sgemm_v4(a, b, c, 2, 2, 2) != MAXMULSK_ERR_UNSUPPORTED) return 6;
        if (c[0] != 7.0f) return 7;
    }
    return 0;
}

static int gemm(void) {
    const size_t M = 67, K = 1031, N = 130;
    float* A = malloc(M * K * sizeof(float));
    float* B = malloc(K * N * sizeof(float));
    float* C = malloc(M * N * sizeof(float));
    if (!A || !B || !C) return 10;
    unsigned s = 12345u;
    for (size_t i = 0; i < M * K; ++i) A[i] = frand(&s);
    for (size_t i = 0; i < K * N; ++i) B[i] = frand(&s);
    for (int entry = 0; entry < 2; ++entry) {
        for (size_t i = 0; i < M * N; ++i) C[i] = 3.75f;
        const maxmulsk_status st = entry == 0 ? maxmulsk_sgemm_v4(A, B, C, M, K, N) : maxmulsk_sgemm_v3(A, B, C, M, K, N);
        if (st != MAXMULSK_OK) {
            fprintf(stderr, "status %d: %s\n", (int)st, maxmulsk_last_error());
            return 11;
        }
        for (size_t i = 0; i < M; ++i)
            for (size_t j = 0; j < N; ++j) {
                double ref = 0, mag = 0;
                for (size_t k = 0; k < K; ++k) {
                    ref += (double)A[i * K + k] * (double)B[k * N + j];
                    mag += fabs((double)A[i * K + k] * (double)B[k * N + j]);
                }
                if (fabs((double)C[i * N + j] - ref) > (double)K * ldexp(1.0, -24) * mag) {
                    fprintf(stderr, "entry %d: C[%zu][%zu] = %.9g, FP64 %.9g\n", entry, i, j, C[i * N + j], ref);
                    return 12;
                }
            }
    }
    free(A);
    free(B);
    free(C);
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 1 && strcmp(argv[1], "gemm") == 0) return gemm();
    return contract();
}
