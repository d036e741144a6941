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


REF = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_plan_rows.npz"))


def _mib3(nbytes: int) -> float:
    """Bytes as the reference's sweeps print them: MiB with three decimals."""
    return float(f"{nbytes / 2.0 ** 20:.3f}")


def walk_nb_loops(M, K, N, Mc, Nc, Kc):
    """Counts events by walking the v4 loop nest itself, independently of the
    closed forms: run_multiplication's clamps
    (sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:638-654), then run_streaming's
    Kc -> Nc -> Mc -> 64-column -> 16-row loops (:483-595): B panels packed on
    mm == 0, A panels packed on nr == 0, one micro-kernel call per 16 x 64 tile.
    Returns (kc_slices, nc_blocks, mc_blocks, a_panel_packs, b_panel_packs,
    microkernel_calls) for microtile-aligned shapes (M % 16 == 0, N % 64 == 0),
    the case NB::counts describes (:463-465)."""
    Kc = min(Kc, K)
    Nc = min(Nc, (N + 63) // 64 * 64)
    Mc = min(Mc, (M + 15) // 16 * 16)
    kcs = ncs = mcs = a = b = mk = 0
    for kk in range(0, K, Kc):
        kcs += 1
        for nn in range(0, N, Nc):
            if kk == 0:
                ncs += 1
            ncl = min(Nc, N - nn)
            for mm in range(0, M, Mc):
                if kk == 0 and nn == 0:
                    mcs += 1
                mcl = min(Mc, M - mm)
                for nr in range(0, ncl, 64):
                    if mm == 0:
                        b += 1
                    for ir in range(0, mcl, 16):
                        if nr == 0:
                            a += 1
                        mk += 1
    return kcs, ncs, mcs, a, b, mk


def test_nb_counts_and_capacity_on_every_sweep_row():
    """All 7,862 native rows of nb_sweep.csv + nb_sweep_ext.csv: a_packs,
    b_packs, mk_calls and both packed-buffer sizes as the reference printed
    them, for every blocking of the sweep."""
    rows, mib = REF["nb"], REF["nb_mib"]
    assert len(rows) == 7862 and rows[:, 9].all()
    for r, (pa, pb) in zip(rows.tolist(), mib.tolist()):
        M, K, N, Mc, Nc, Kc, a_packs, b_packs, mk = r[:9]
        assert orc.nb_classify(M, K, N, Mc, Nc, Kc) == 0, r
        c = orc.nb_counts(M, K, N, Mc, Nc, Kc)
        assert c[3:] == (a_packs, b_packs, mk), (r, c)
        ca, cb = orc.nb_capacity(M, K, N, Mc, Nc, Kc)
        assert (_mib3(ca), _mib3(cb)) == (pa, pb), (r, ca, cb)


def test_nb_counts_all_six_fields_against_the_loop_nest():
    """Every field of nb_counts, including the block counts the sweep files do
    not print, against the loop walker, on every sweep shape x a spread of the
    sweep's blockings (including Kc > K and Nc > N clamps)."""
    shapes = sorted({tuple(r[:3]) for r in REF["nb"].tolist()})
    assert len(shapes) == 35
    blockings = [(16, 64, 256), (16, 512, 2048), (32, 1024, 2048), (64, 256, 1024),
                 (128, 4096, 16384), (128, 64, 512), (48, 192, 320)]
    for (M, K, N) in shapes:
        for (Mc, Nc, Kc) in blockings:
            assert orc.nb_counts(M, K, N, Mc, Nc, Kc) == walk_nb_loops(M, K, N, Mc, Nc, Kc), (M, K, N, Mc, Nc, Kc)


def test_loop_walker_catches_planted_mistakes():
    """The walker distinguishes the fields a swapped or dropped term would
    confuse: a_panel_packs (per Kc slice and Nc block) from b_panel_packs (per
    Kc slice only), and nc_blocks from mc_blocks."""
    c = walk_nb_loops(128, 4096, 1024, 32, 256, 1024)
    assert c == (4, 4, 4, 4 * 4 * 8, 4 * 16, 4 * 16 * 8)
    swapped = c[:3] + (c[4], c[3]) + c[5:]
    assert swapped != c


def test_v4c_choose_on_every_v4c3_row():
    """v4c3.csv: the profile (v4a/v4b) and blocking the v4c rule picked for each
    of its 35 shapes, and that blocking's packed-buffer sizes."""
    rows, mib = REF["v4c"], REF["v4c_mib"]
    assert len(rows) == 35 and set(rows[:, 3].tolist()) == {0, 1}
    for r, (pa, pb) in zip(rows.tolist(), mib.tolist()):
        M, K, N, prof, Mc, Nc, Kc = r
        assert orc.v4c_choose(M, K, N) == (Mc, Nc, Kc), r
        assert (Mc, Nc, Kc) == ((32, 1024, 2048) if prof else (16, 512, 2048))
        ca, cb = orc.nb_capacity(M, K, N, Mc, Nc, Kc)
        assert (_mib3(ca), _mib3(cb)) == (pa, pb), r


def test_v3_table_and_v4_capacities_on_every_final5_row():
    """final5.csv: the v3 tuning table's (M_tile, N_tile, Kc) for 35 shapes, and
    NB::capacity of both v4c blockings (v4a and v4b) for each."""
    rows, mib = REF["final5"], REF["final5_mib"]
    assert len(rows) == 35
    for r, m in zip(rows.tolist(), mib.tolist()):
        M, K, N, mt, nt, kc = r
        assert orc.v3_select_from_table(M, K, N)[:3] == (mt, nt, kc), r
        got = orc.nb_capacity(M, K, N, 16, 512, 2048) + orc.nb_capacity(M, K, N, 32, 1024, 2048)
        assert tuple(_mib3(x) for x in got) == tuple(m), (r, got)


def test_v3_table_on_every_headline_row():
    """headline_run*.csv and tuned_run*.csv: the blocking column, including the
    ir_outer ("swap") flag, on all 66 MaxMulSK rows."""
    rows = REF["headline"]
    assert len(rows) == 66
    for r in rows.tolist():
        assert orc.v3_select_from_table(*r[:3]) == tuple(r[3:]), r


def test_nb_classify_bad_blockings():
    # bad blockings of check_nb.cpp:69-79 (refused, C untouched) and empty shapes
    # (ncblock.cpp:411); every sweep row above is classified native.
    assert orc.nb_classify(256, 256, 256, 24, 256, 2048) == 2
    assert orc.nb_classify(256, 256, 256, 16, 100, 2048) == 2
    assert orc.nb_classify(256, 256, 256, 16, 256, 100) == 2
    assert orc.nb_classify(0, 256, 256, 16, 256, 2048) == 1
    assert orc.nb_classify(256, 0, 256, 16, 256, 2048) == 1
    assert orc.nb_classify(256, 256, 0, 16, 256, 2048) == 1
