"""GPU parity of the FP32-MFMA and split-BF16 GEMM paths against the oracle,
through the C ABI. Shapes are M x N x K as in the workload list; the ABI takes
(M, K, N)."""
import numpy as np
import pytest

import maxmulsk_amd as mm
from oracle import oracle as orc
from tests.gpu_util import SENTINEL, Guarded, full_check, gpu_gemm, gpu_gemm_checked, report, torch_cuda

pytestmark = pytest.mark.gpu

ALL_PLANS = [(t, s) for t in (64, 128, 256) for s in (1, 2, 4, 8, 16, 32)]


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("shape", [(64, 512, 256), (67, 200, 130), (200, 96, 257), (1, 40, 1)])
def test_small_integers_bit_exact_every_plan(mode, shape):
    """Integers |x| <= 8 are exact in BF16 and every partial sum is exact in FP32,
    so every plan, tile edge and split must reproduce the integer product exactly."""
    M, K, N = shape
    A = orc.small_ints(1000 + M, (M, K), 8)
    B = orc.small_ints(2000 + N, (K, N), 8)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
    for plan in ALL_PLANS:
        Cg = gpu_gemm(A, B, mode=mode, plan=plan)
        bad = np.argwhere(Cg != exact)
        assert bad.size == 0, f"mode {mode} plan {plan}: {len(bad)} wrong, first {bad[:4].tolist()}"


def _up256(x):
    return (x + 255) // 256 * 256


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_every_split_count_reduces_exactly(mode):
    """Plan (tile, 32) on a K of s k-blocks runs exactly s slices, so the reduce
    sees every split count 1..32 (direct 64-row tiles: k-blocks of 32; packed
    256 x 256 tiles: k-blocks of 32, or 16 in split6). Small integers keep every
    partial sum exact, so the result must equal the integer product bit for bit."""
    packed_depth = mm.tile_info(5, mode)["bk"]
    assert packed_depth == (16 if mode == 2 else 32)
    for s in range(1, 33):
        for M, N, tile, depth in ((67, 93, 64, 32), (300, 260, 256, packed_depth)):
            K = depth * s
            extra = mm.workspace_bytes(M, K, N, mode, (tile, 32)) - (
                mm.workspace_bytes(M, K, N, mode, (tile, 1)))
            assert extra == (_up256(s * M * N * 4) if s > 1 else 0), (s, tile, extra)
            A = orc.small_ints(3000 + s, (M, K), 8)
            B = orc.small_ints(4000 + s, (K, N), 8)
            exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
            Cg = gpu_gemm(A, B, mode=mode, plan=(tile, 32))
            bad = np.argwhere(Cg != exact)
            assert bad.size == 0, f"mode {mode} splits {s} tile {tile}: {len(bad)} wrong, first {bad[:4].tolist()}"


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("shape,plan", [((64, 8192, 512), (64, 16)), ((128, 32768, 512), (64, 16)),
                                        ((100, 999, 300), (64, 4)), ((512, 4096, 512), (64, 4)),
                                        ((67, 8191, 513), (64, 16)), ((1000, 1000, 1000), (64, 2))],
                         ids=["64x512x8192", "128x512x32768", "100x300x999", "512x512x4096",
                              "67x513x8191", "1000x1000x1000"])
def test_deep_small_grid_tile_matches_regular_tile(mode, shape, plan):
    """A 64-tile launch of at most 256 workgroups runs the deep-k-block variant,
    one of at most 512 the 64-deep one; the grouped kernel always runs its own
    k-block depth. Same slices, same summation order: the two must agree bit for
    bit on random data."""
    torch = torch_cuda()
    M, K, N = shape
    tiles = -(-M // 64) * -(-N // 64)
    assert tiles * plan[1] <= 512  # one of the deep variants is the one launched
    A = orc.uniform(7000 + M, (M, K))
    B = orc.uniform(8000 + N, (K, N))
    single = gpu_gemm(A, B, mode=mode, plan=plan)
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL)
    gp = mm.GroupedPlan([(dA.ptr, dB.ptr, dC.ptr, M, K, N)], mode=mode, plans=[plan])
    gp.launch()
    torch.cuda.synchronize()
    gp.close()
    grouped = dC.numpy((M, N))
    assert np.array_equal(single, grouped), int((single != grouped).sum())
    assert full_check(A, B, single)["bound_ratio"] <= 1.0


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("shape,plan", [((300, 999, 1025), (128, 2)), ((256, 4096, 512), (128, 1)),
                                        ((1000, 640, 700), (128, 4)), ((100, 2000, 1536), (256, 2)),
                                        ((1024, 1536, 1024), (128, 1)), ((512, 4095, 1000), (128, 2))],
                         ids=["300x1025x999", "256x512x4096", "1000x700x640", "100x1536x2000",
                              "1024x1024x1536", "512x1000x4095"])
def test_eight_wave_128_tile_matches_grouped_tile(mode, shape, plan):
    """Direct split-mode launches of the 128 x 128 and 128 x 256 tiles run
    eight waves (split3 128 x 128 launches of at most 256 workgroups with
    64-deep k-blocks), and the grouped kernel runs its own eight-wave tiles.
    Both compute the same products in the same order as the four-wave layout,
    so the results must agree bit for bit."""
    torch = torch_cuda()
    M, K, N = shape
    A = orc.uniform(9000 + M, (M, K))
    B = orc.uniform(9500 + N, (K, N))
    single = gpu_gemm(A, B, mode=mode, plan=plan)
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL)
    gp = mm.GroupedPlan([(dA.ptr, dB.ptr, dC.ptr, M, K, N)], mode=mode, plans=[plan])
    gp.launch()
    torch.cuda.synchronize()
    gp.close()
    grouped = dC.numpy((M, N))
    assert np.array_equal(single, grouped), int((single != grouped).sum())
    assert full_check(A, B, single)["bound_ratio"] <= 1.0


# Plain FP32 MFMA on the measured shapes (M x N x K), checked against FP64 within
# |C - C64| <= K * 2^-24 * |A||B| (gamma = 1).
PLAIN_SHAPES = [
    (1024, 1024, 1024), (2048, 2048, 2048), (4096, 4096, 4096),
    (64, 512, 8192), (128, 512, 32768), (64, 24576, 1536), (64, 32768, 512),
    (1023, 1025, 999), (67, 513, 8191),
]


@pytest.mark.parametrize("mnk", PLAIN_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_plain_within_fp32_bound(mnk):
    """Every element of C against FP64 (the GPU oracle), through the
    reference's own entry point; small shapes also against the CPU oracle."""
    M, N, K = mnk
    A = orc.uniform(11, (M, K))
    B = orc.uniform(12, (K, N))
    Cg, e = gpu_gemm_checked(A, B, mode=0, entry="v4")
    assert not np.any(Cg == np.float32(SENTINEL))
    if M * N * K <= 2 ** 30:
        ec = full_check(A, B, Cg)
        assert ec["bound_ratio"] <= 1.0, ec
        assert abs(ec["bound_ratio"] - e["bound_ratio"]) <= 1e-5 * ec["bound_ratio"], (e, ec)
    assert e["nonfinite"] == 0
    assert e["bound_ratio"] <= 1.0, e
    plan, _ = mm.choose_plan(M, K, N, 0)
    report("accuracy", shape_mnk=[M, N, K], mode="plain", dist="uniform", plan=list(plan), **e)
    print(f"plain {M}x{N}x{K}: {e}")


@pytest.mark.parametrize("plan", ALL_PLANS)
def test_plain_ragged_every_plan(plan):
    M, N, K = 1023, 1025, 999
    A = orc.uniform(21, (M, K))
    B = orc.uniform(22, (K, N))
    Cg = gpu_gemm(A, B, mode=0, plan=plan)
    e = full_check(A, B, Cg)
    assert e["bound_ratio"] <= 1.0, (plan, e)


SPLIT_SHAPES = [(2048, 2048, 2048), (4096, 4096, 4096), (64, 7168, 16384), (4096, 24576, 1536),
                (1023, 1025, 999), (67, 513, 8191)]


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("dist", ["uniform", "logu"])
@pytest.mark.parametrize("mnk", SPLIT_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_split_within_fp32_bound(mode, dist, mnk):
    """Every element of C against FP64 (the GPU oracle) on the split-precision
    workload shapes, both distributions, beside plain FP32 MFMA (mode 0) on
    the same inputs."""
    M, N, K = mnk
    gen = orc.uniform if dist == "uniform" else orc.log_uniform
    A = gen(31, (M, K))
    B = gen(32, (K, N))
    Cg, e = gpu_gemm_checked(A, B, mode=mode)
    if M * N * K <= 2 ** 30:
        ec = full_check(A, B, Cg)
        assert abs(ec["bound_ratio"] - e["bound_ratio"]) <= 1e-5 * ec["bound_ratio"], (e, ec)
    assert e["nonfinite"] == 0
    plan, _ = mm.choose_plan(M, K, N, mode)
    report("accuracy", shape_mnk=[M, N, K], mode=["plain", "split3", "split6"][mode], dist=dist,
           plan=list(plan), **e)
    print(f"mode {mode} {dist} {M}x{N}x{K}: {e}")
    # split6 is FP32-level on any input; split3 drops lo*lo (<= 2^-16 |a||b| per
    # product), inside the K*u32 bound for these K on both distributions
    assert e["bound_ratio"] <= 1.0, e


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("plan", [(64, 1), (128, 4), (256, 1), (256, 16)])
def test_split_matches_exact_split_emulation(mode, plan):
    """The GPU result must equal the exact FP64 sum of the mode's BF16 cross
    products up to FP32 accumulation: |C - S64| <= P*K*u32*|A||B|, P products."""
    M, N, K = 200, 300, 640
    A = orc.uniform(41, (M, K))
    B = orc.log_uniform(42, (K, N), -4, 4)
    Cg = gpu_gemm(A, B, mode=mode, plan=plan)
    S64 = orc.split_gemm_f64(A, B, mode)
    ab = orc.absgemm_f64(A, B)
    P = 3 if mode == 1 else 6
    d = np.abs(Cg.astype(np.float64) - S64)
    assert np.all(d <= P * K * orc.U32 * ab), float(np.max(d / ab))
    # and it is closer to the split value than the split value is to FP64 for split3
    if mode == 1:
        C64 = orc.gemm_f64(A, B)
        assert np.max(np.abs(S64 - C64) / ab) > 0  # split3 is not exact


def test_v3_v4_and_host_pointers_agree():
    M, N, K = 300, 257, 515
    A = orc.uniform(51, (M, K))
    B = orc.uniform(52, (K, N))
    torch_cuda()
    C3 = gpu_gemm(A, B, entry="v3")
    C4 = gpu_gemm(A, B, entry="v4")
    Ch = np.full((M, N), SENTINEL, np.float32)
    assert mm.sgemm_v4(A, B, Ch, M, K, N) == mm.OK, mm.last_error()
    Ch3 = np.full((M, N), SENTINEL, np.float32)
    assert mm.sgemm_v3(A, B, Ch3, M, K, N) == mm.OK, mm.last_error()
    # same plan, same kernel: bit-identical
    assert np.array_equal(C4, Ch)
    assert np.array_equal(C3, Ch3)
    for Cx in (C3, C4, Ch):
        assert full_check(A, B, Cx)["bound_ratio"] <= 1.0


@pytest.mark.parametrize("name,mode", [("split3", 1), ("split6", 2)])
def test_mode_environment_switches_the_reference_entry_points(monkeypatch, name, mode):
    """MAXMULSK_MODE switches maxmulsk_sgemm_v4 and v3 to a split mode, read on
    every call: v4 then equals maxmulsk_sgemm_ex in that mode with the same
    planner bit for bit, v3 equals sgemm_ex with the shape rule's plan, and
    both differ from plain on data that is not exact in BF16."""
    M, N, K = 256, 384, 2048
    A = orc.uniform(61, (M, K))
    B = orc.uniform(62, (K, N))
    torch_cuda()
    plain = gpu_gemm(A, B, entry="v4")
    monkeypatch.setenv("MAXMULSK_MODE", name)
    C4 = gpu_gemm(A, B, entry="v4")
    C3 = gpu_gemm(A, B, entry="v3")
    monkeypatch.delenv("MAXMULSK_MODE")
    Cex = gpu_gemm(A, B, mode=mode)
    Crule = gpu_gemm(A, B, mode=mode, plan=mm.rule_plan(M, K, N, mode))
    assert np.array_equal(C4, Cex)
    assert np.array_equal(C3, Crule)
    assert not np.array_equal(C4, plain)
    for Cx in (C3, C4):
        assert full_check(A, B, Cx)["bound_ratio"] <= 1.0
    # an unknown name leaves plain mode
    monkeypatch.setenv("MAXMULSK_MODE", "split4")
    assert np.array_equal(gpu_gemm(A, B, entry="v4"), plain)


@pytest.mark.parametrize("plan", [(128, 1), (256, 2)])
def test_host_operands_larger_than_the_staging_chunks(plan):
    """Host operands that span several 8 MiB staging chunks, with ragged sizes,
    in pageable and pinned host memory, mixed with device operands: every
    combination gives the device-pointer result bit for bit and leaves the
    device operands and C's guards intact."""
    torch = torch_cuda()
    M, N, K = 2600, 1300, 1700  # A 17.7 MB, B 8.8 MB, C 13.5 MB
    A = orc.uniform(71, (M, K))
    B = orc.uniform(72, (K, N))
    dA, dB = Guarded(M * K, A), Guarded(K * N, B)
    dC = Guarded(M * N, SENTINEL)
    assert mm.sgemm_ex(dA.ptr, dB.ptr, dC.ptr, M, K, N, mode=1, plan=plan) == mm.OK, mm.last_error()
    torch.cuda.synchronize()
    ref = dC.numpy((M, N)).copy()
    assert full_check(A, B, ref)["bound_ratio"] <= 1.0
    pA = torch.from_numpy(A).pin_memory()
    pB = torch.from_numpy(B).pin_memory()
    pC = torch.empty((M, N), dtype=torch.float32).pin_memory()
    hC = np.empty((M, N), np.float32)
    cases = [("pageable A, B, C", A, B, hC), ("pinned A, pageable B, pinned C", pA, B, pC),
             ("pageable A, device B, pageable C", A, dB.ptr, hC), ("pageable A, B, device C", A, B, dC),
             ("device A, pinned B, device C", dA.ptr, pB, dC)]
    for name, a, b, c in cases:
        if c is dC:
            dC.view.fill_(SENTINEL)
            torch.cuda.synchronize()
        elif isinstance(c, np.ndarray):
            c.fill(SENTINEL)
        else:
            c.fill_(SENTINEL)
        assert mm.sgemm_ex(a, b, c.ptr if c is dC else c, M, K, N, mode=1, plan=plan) == mm.OK, (name, mm.last_error())
        got = dC.numpy((M, N)) if c is dC else (c if isinstance(c, np.ndarray) else c.numpy())
        assert np.array_equal(got, ref), name
    assert dA.guards_intact() and dB.guards_intact() and dC.guards_intact()
    assert np.array_equal(dA.numpy((M, K)), A) and np.array_equal(dB.numpy((K, N)), B)


def test_repeat_call_with_new_B_same_buffers():
    """check_nb.cpp:61-66: nothing stale may survive between calls."""
    torch = torch_cuda()
    M, N, K = 130, 300, 2000
    A = orc.uniform(61, (M, K))
    B1 = orc.uniform(62, (K, N))
    B2 = orc.uniform(63, (K, N))
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B1), Guarded(M * N, SENTINEL)
    for plan in [(64, 16), (128, 4), (256, 1)]:
        for B in (B1, B2):
            dB.view.copy_(torch.from_numpy(B.reshape(-1)))
            assert mm.sgemm_ex(dA.ptr, dB.ptr, dC.ptr, M, K, N, plan=plan) == mm.OK
            torch.cuda.synchronize()
            assert full_check(A, B, dC.numpy((M, N)))["bound_ratio"] <= 1.0
    assert dC.guards_intact()


def test_caller_workspace_entry_point():
    torch = torch_cuda()
    M, N, K = 64, 512, 8192
    A = orc.uniform(71, (M, K))
    B = orc.uniform(72, (K, N))
    plan = (64, 16)
    need = mm.workspace_bytes(M, K, N, 0, plan)
    assert need == 16 * M * N * 4
    ws = torch.empty(need // 4 + 64, dtype=torch.float32, device="cuda")
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL)
    # short workspace: ERR_ALLOC, C untouched
    assert mm.sgemm_ws(dA.ptr, dB.ptr, dC.ptr, M, K, N, 0, plan, ws.data_ptr(), need - 4) == mm.ERR_ALLOC
    torch.cuda.synchronize()
    assert np.all(dC.numpy((M, N)) == np.float32(SENTINEL))
    # misaligned workspace: ERR_ALLOC
    assert mm.sgemm_ws(dA.ptr, dB.ptr, dC.ptr, M, K, N, 0, plan, ws.data_ptr() + 4, need) == mm.ERR_ALLOC
    assert mm.sgemm_ws(dA.ptr, dB.ptr, dC.ptr, M, K, N, 0, plan, ws.data_ptr(), need) == mm.OK
    torch.cuda.synchronize()
    assert full_check(A, B, dC.numpy((M, N)))["bound_ratio"] <= 1.0
    # a plan without split-K needs no workspace
    assert mm.workspace_bytes(M, K, N, 0, (64, 1)) == 0
    assert mm.sgemm_ws(dA.ptr, dB.ptr, dC.ptr, M, K, N, 0, (64, 1), None, 0) == mm.OK
    torch.cuda.synchronize()
    assert full_check(A, B, dC.numpy((M, N)))["bound_ratio"] <= 1.0


def test_async_on_a_stream_and_k_zero():
    torch = torch_cuda()
    M, N, K = 256, 256, 256
    A = orc.uniform(81, (M, K))
    B = orc.uniform(82, (K, N))
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    assert mm.sgemm_ex(dA.ptr, dB.ptr, dC.ptr, M, K, N, stream=s, async_=True) == mm.OK
    s.synchronize()
    assert full_check(A, B, dC.numpy((M, N)))["bound_ratio"] <= 1.0
    # K == 0 on device memory: C becomes zero
    assert mm.sgemm_v4(dA.ptr, dB.ptr, dC.ptr, M, 0, N) == mm.OK
    torch.cuda.synchronize()
    assert np.all(dC.numpy((M, N)) == 0.0) and dC.guards_intact()
    # host pointers cannot be asynchronous
    Ch = np.zeros((M, N), np.float32)
    assert mm.sgemm_ex(A, B, Ch, M, K, N, async_=True) == mm.ERR_UNSUPPORTED


def test_phase_counters_follow_the_launch_structure():
    """Launch counts per call, as NB::counts states them for the reference's
    loop nest: direct tiles launch one GEMM (plus one reduce with split-K), the
    256 x 256 path packs both operands in one launch first, host operands are
    staged in and written back, and a warm pool allocates nothing."""
    torch = torch_cuda()
    M, N, K = 320, 384, 1024
    A = orc.uniform(131, (M, K))
    B = orc.uniform(132, (K, N))
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL)
    torch.cuda.synchronize()
    expect = {(64, 1): (0, 1, 0), (64, 4): (0, 1, 1), (256, 1): (1, 1, 0), (256, 2): (1, 1, 1)}
    for plan, (packs, gemms, reduces) in expect.items():
        assert mm.sgemm_ex(dA.ptr, dB.ptr, dC.ptr, M, K, N, mode=1, plan=plan) == mm.OK  # warms the pool
        mm.reset_phase_counters()
        assert mm.sgemm_ex(dA.ptr, dB.ptr, dC.ptr, M, K, N, mode=1, plan=plan) == mm.OK
        c = mm.phase_counters()
        assert (c["pack"], c["compute"], c["reduce"]) == (packs, gemms, reduces), (plan, c)
        assert c["allocation"] == 0 and c["stage_in"] == 0 and c["writeback"] == 0, (plan, c)
    Ch = np.full((M, N), SENTINEL, np.float32)
    assert mm.sgemm_v4(A, B, Ch, M, K, N) == mm.OK
    mm.reset_phase_counters()
    assert mm.sgemm_v4(A, B, Ch, M, K, N) == mm.OK
    c = mm.phase_counters()
    assert c["stage_in"] == 2 and c["writeback"] == 1 and c["compute"] == 1 and c["allocation"] == 0, c
    mm.reset_phase_counters()
    assert mm.sgemm_v4(dA.ptr, dB.ptr, dC.ptr, M, 0, N) == mm.OK
    assert mm.phase_counters()["zero_fill"] == 1
