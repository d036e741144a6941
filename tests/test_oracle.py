"""Pins the oracle (CPU): golden fixtures, closed forms, brute force, identities,
and planted mistakes that the checks must catch."""
import os

import numpy as np
import pytest

from oracle import oracle as orc

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz"))


def brute_f64(A, B):
    M, K = A.shape
    N = B.shape[1]
    out = np.zeros((M, N))
    for i in range(M):
        for j in range(N):
            s = 0.0
            for k in range(K):
                s += float(A[i, k]) * float(B[k, j])
            out[i, j] = s
    return out


def test_prng_golden():
    assert np.array_equal(orc.uniform(1, (64,)), FIX["prng_uniform_seed1"])
    assert np.array_equal(orc.log_uniform(3, (64,)), FIX["prng_logu_seed3"])
    assert np.array_equal(orc._splitmix64(1, 16), FIX["prng_raw_seed1"])


def test_prng_ranges():
    u = orc.uniform(5, (100000,))
    assert u.dtype == np.float32 and u.min() >= -1.0 and u.max() < 1.0
    assert abs(float(u.mean())) < 0.01
    # exact 2^-23 grid
    assert np.all(np.round((u.astype(np.float64) + 1.0) * 2 ** 23) == (u.astype(np.float64) + 1.0) * 2 ** 23)
    g = orc.log_uniform(6, (100000,))
    e = np.log2(np.abs(g.astype(np.float64)))
    assert e.min() >= -20.0001 and e.max() <= 20.0001
    assert 0.45 < float(np.mean(g < 0)) < 0.55


@pytest.mark.parametrize("n", range(5))
def test_gemm_f64_matches_exact_fixture(n):
    A, B, Cx = FIX[f"gemm{n}_A"], FIX[f"gemm{n}_B"], FIX[f"gemm{n}_C_exact"]
    K = A.shape[1]
    for fn in (orc.gemm_f64, orc.gemm_f64_c):
        C64 = fn(A, B)
        # FP64 accumulation error only: K * 2^-53 * |A||B|
        bound = K * 2.0 ** -52 * orc.absgemm_f64(A, B) + 1e-300
        assert np.all(np.abs(C64 - Cx) <= bound)


def test_gemm_f64_logu_fixture():
    A, B, Cx = FIX["gemm_logu_A"], FIX["gemm_logu_B"], FIX["gemm_logu_C_exact"]
    C64 = orc.gemm_f64_c(A, B)
    bound = A.shape[1] * 2.0 ** -52 * orc.absgemm_f64(A, B)
    assert np.all(np.abs(C64 - Cx) <= bound)


def test_gemm_f64_brute_and_c_agree():
    A = orc.uniform(11, (9, 23))
    B = orc.uniform(12, (23, 7))
    bf = brute_f64(A, B)
    assert np.allclose(orc.gemm_f64(A, B), bf, rtol=0, atol=1e-13)
    assert np.allclose(orc.gemm_f64_c(A, B), bf, rtol=0, atol=1e-13)
    assert np.allclose(orc.absgemm_f64_c(A, B), brute_f64(np.abs(A), np.abs(B)), rtol=0, atol=1e-13)


def test_identities_and_exact_integers():
    A = orc.uniform(13, (31, 31))
    I = np.eye(31, dtype=np.float32)
    assert np.array_equal(orc.gemm_f64_c(A, I), A.astype(np.float64))
    assert np.array_equal(orc.gemm_f64_c(I, A), A.astype(np.float64))
    # small integers: every partial sum exact in FP32, any order is bit-exact
    Ai = orc.small_ints(14, (40, 64), 16)
    Bi = orc.small_ints(15, (64, 24), 16)
    exact = Ai.astype(np.int64) @ Bi.astype(np.int64)
    assert np.array_equal(orc.sgemm_portable(Ai, Bi), exact.astype(np.float32))
    assert np.array_equal(orc.gemm_f64_c(Ai, Bi), exact.astype(np.float64))


def test_portable_path_within_fp32_bound():
    A = orc.uniform(16, (48, 256))
    B = orc.uniform(17, (256, 40))
    e = orc.errors(orc.sgemm_portable(A, B), orc.gemm_f64(A, B), orc.absgemm_f64(A, B), 256)
    assert e["bound_ratio"] <= 1.0
    assert e["bound_ratio"] > 0.0  # FP32 accumulation is not exact on random inputs


def test_portable_path_is_an_fma_chain():
    # Inputs on a coarse grid keep every FMA result exact in FP64, so a float64
    # emulation of fmaf(a, b, sum) rounded to FP32 per step is exact.
    A = (orc.small_ints(18, (6, 50), 64) / 64).astype(np.float32)
    B = (orc.small_ints(19, (50, 5), 64) / 8).astype(np.float32)
    ref = np.zeros((6, 5), np.float32)
    for i in range(6):
        for j in range(5):
            s = np.float32(0)
            for k in range(50):
                s = np.float32(float(A[i, k]) * float(B[k, j]) + float(s))
            ref[i, j] = s
    assert np.array_equal(orc.sgemm_portable(A, B), ref)


def test_checker_catches_planted_mistakes():
    A = orc.uniform(20, (32, 128))
    B = orc.uniform(21, (128, 32))
    C64 = orc.gemm_f64(A, B)
    ab = orc.absgemm_f64(A, B)
    good = orc.sgemm_portable(A, B)
    assert orc.errors(good, C64, ab, 128)["bound_ratio"] <= 1.0
    # transposed operand
    bad = orc.sgemm_portable(np.ascontiguousarray(A), np.ascontiguousarray(B.T.copy().reshape(128, 32)))
    assert orc.errors(bad, C64, ab, 128)["bound_ratio"] > 1.0
    # dropped last k term
    dropped = orc.sgemm_portable(A[:, :-1].copy(), B[:-1].copy())
    assert orc.errors(dropped, C64, ab, 128)["bound_ratio"] > 1.0
    # wrong sign on one element
    flipped = good.copy()
    flipped[3, 4] = -flipped[3, 4]
    assert orc.errors(flipped, C64, ab, 128)["bound_ratio"] > 1.0
    # off-by-one column shift
    shifted = np.roll(good, 1, axis=1)
    assert orc.errors(shifted, C64, ab, 128)["bound_ratio"] > 1.0


def test_bf16_rounding_fixture():
    x = FIX["split_x"]
    got = np.array([orc.bf16_rn(float(v)) for v in x])
    assert np.array_equal(got, FIX["split_bf16"])


def test_bf16_exhaustive_round_trip():
    # Every finite BF16 value is a fixed point of rounding.
    bits = np.arange(0, 1 << 16, dtype=np.uint32) << np.uint32(16)
    vals = bits.view(np.float32)
    fin = np.isfinite(vals)
    for v in vals[fin][::97]:
        assert orc.bf16_rn(float(v)) == float(v)
    sp = orc.split(vals[fin], 2)
    assert np.array_equal(sp[0], vals[fin]) and np.all(sp[1] == 0)


def test_bf16_ties_to_even():
    one = 1.0
    assert orc.bf16_rn(one + 2.0 ** -8) == 1.0            # tie -> even (down)
    assert orc.bf16_rn(one + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie -> even (up)
    assert orc.bf16_rn(one + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7  # above tie


def test_split_fixture_and_bounds():
    x = FIX["split_x"]
    assert np.array_equal(orc.split(x, 2).T.astype(np.float64), FIX["split_terms2"])
    assert np.array_equal(orc.split(x, 3).T.astype(np.float64), FIX["split_terms3"])
    # representation bounds on every exponent in a wide range
    xs = np.concatenate([orc.log_uniform(30, (20000,), -100, 100), orc.uniform(31, (20000,))])
    xs = xs[xs != 0]
    x64 = xs.astype(np.float64)
    s2 = orc.split(xs, 2).astype(np.float64)
    s3 = orc.split(xs, 3).astype(np.float64)
    assert np.all(np.abs(x64 - s2.sum(0)) <= 2.0 ** -16 * np.abs(x64))
    assert np.all(np.abs(x64 - s3.sum(0)) <= 2.0 ** -24 * np.abs(x64))
    # 3 terms of 8 bits cover all 24 bits of an FP32 significand: exact
    assert np.array_equal(s3.sum(0), x64)


def test_split_gemm_numpy_vs_c_and_exact_cases():
    A = orc.uniform(40, (20, 36))
    B = orc.uniform(41, (36, 12))
    for mode in (1, 2):
        assert np.allclose(orc.split_gemm_f64(A, B, mode), orc.split_gemm_f64_c(A, B, mode),
                           rtol=0, atol=1e-12)
    # operands already in BF16: every lo term is 0, split == exact product
    Ab = orc.split(A, 1)[0]
    Bb = orc.split(B, 1)[0]
    for mode in (1, 2):
        assert np.allclose(orc.split_gemm_f64(Ab, Bb, mode), orc.gemm_f64(Ab, Bb), rtol=0, atol=1e-13)
    # split6 keeps FP32-level accuracy of every product; split3 drops lo*lo only
    C64 = orc.gemm_f64(A, B)
    ab = orc.absgemm_f64(A, B)
    assert np.all(np.abs(orc.split_gemm_f64(A, B, 2) - C64) <= 4 * 2.0 ** -24 * ab)
    assert np.all(np.abs(orc.split_gemm_f64(A, B, 1) - C64) <= 3 * 2.0 ** -16 * ab)
    # planted: dropping the hi*lo product leaves a visible error
    terms = orc.split(A, 2).astype(np.float64), orc.split(B, 2).astype(np.float64)
    partial = terms[0][1] @ terms[1][0] + terms[0][0] @ terms[1][0]
    assert np.max(np.abs(partial - C64) / ab) > 2.0 ** -12


def test_reference_plan_rules():
    # v4c: v4b exactly when M >= 2048 and N > 512 (sme-1x4-v4c-dispatch.cpp:17-22)
    assert orc.v4c_choose(4096, 4096, 4096) == (32, 1024, 2048)
    assert orc.v4c_choose(2048, 2048, 2048) == (32, 1024, 2048)
    assert orc.v4c_choose(4096, 1536, 24576) == (32, 1024, 2048)
    assert orc.v4c_choose(4096, 4096, 256) == (16, 512, 2048)   # N <= 512
    for M, K, N in [(256, 256, 256), (1024, 1024, 1024), (64, 8192, 512), (128, 32768, 512),
                    (64, 1536, 24576), (1023, 999, 1025), (67, 8191, 513)]:
        assert orc.v4c_choose(M, K, N) == (16, 512, 2048)
    # v3 table: squares >= 2048 only (gemm_tuning.cpp:91-94)
    assert orc.v3_select_from_table(4096, 4096, 4096) == (2048, 512, 4096, 1)
    assert orc.v3_select_from_table(2048, 2048, 2048) == (1024, 512, 2048, 1)
    assert orc.v3_select_from_table(1024, 1024, 1024) == (1024, 64, 2048, 0)
    assert orc.v3_select_from_table(64, 8192, 512) == (1024, 64, 2048, 0)
    assert orc.v51_choose(1, 2, 3) == (64, 1024, 1024)
    # classify: bad blockings of check_nb.cpp:69-79 and empty shapes
    assert orc.nb_classify(256, 256, 256, 24, 256, 2048) == 2
    assert orc.nb_classify(256, 256, 256, 16, 100, 2048) == 2
    assert orc.nb_classify(256, 256, 256, 16, 256, 100) == 2
    assert orc.nb_classify(0, 256, 256, 16, 256, 2048) == 1
    assert orc.nb_classify(256, 256, 256, 16, 256, 2048) == 0
    # capacity at 4096^3 with v4b: packed_A 32*2048*4, packed_B 1024*2048*4
    assert orc.nb_capacity(4096, 4096, 4096, 32, 1024, 2048) == (32 * 2048 * 4, 1024 * 2048 * 4)
    # 256^3 with v4a: Nc clamps to 256, Kc to 256
    assert orc.nb_capacity(256, 256, 256, 16, 512, 2048) == (16 * 256 * 4, 256 * 256 * 4)
    # counts for 64x8192(K)x512 with v4a: 4 Kc slices, 1 Nc block, 4 Mc blocks
    c = orc.nb_counts(64, 8192, 512, 16, 512, 2048)
    assert c[:3] == (4, 1, 4)
    assert c[5] == 4 * (512 // 64) * (64 // 16)
