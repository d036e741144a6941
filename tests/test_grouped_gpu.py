"""Grouped persistent launch on the GPU: per-problem parity with the oracle,
bit-identity with standalone calls of the same plan and geometry, graph
replays, and the validate-all-first contract of run_batch
(sme/v4/sme-1x4-v4c-onestream.cpp:683-694)."""
import numpy as np
import pytest

import maxmulsk_amd as mm
from oracle import oracle as orc
from tests.gpu_util import SENTINEL, Guarded, full_check, torch_cuda

pytestmark = pytest.mark.gpu

# (M, K, N) in ABI order, including empty and K == 0 problems
SPECS = [(1, 512, 512), (37, 1536, 2112), (64, 7168, 512), (130, 512, 4096), (300, 999, 1025),
         (512, 16384, 512), (5, 0, 7), (0, 10, 10), (200, 64, 96), (129, 2048, 256), (64, 513, 67)]


def build(specs, seed, gen=orc.uniform):
    probs, host = [], []
    for i, (M, K, N) in enumerate(specs):
        A = gen(seed + 2 * i, (M, K)) if M * K else np.zeros((M, K), np.float32)
        B = gen(seed + 2 * i + 1, (K, N)) if K * N else np.zeros((K, N), np.float32)
        dA = Guarded(max(M * K, 1), A.reshape(-1) if M * K else 0.0)
        dB = Guarded(max(K * N, 1), B.reshape(-1) if K * N else 0.0)
        dC = Guarded(max(M * N, 1), SENTINEL)
        probs.append((dA, dB, dC, M, K, N))
        host.append((A, B))
    return probs, host


def as_args(probs):
    return [(a.ptr, b.ptr, c.ptr, M, K, N) for a, b, c, M, K, N in probs]


def standalone_plan(M, plan):
    """Plan of a standalone call that computes every element as a grouped
    launch of `plan` does: a 256-wide tile on more than 128 rows runs on
    128-row tiles there, so the standalone call uses 128-wide tiles (the same
    K slices)."""
    return (128, plan[1]) if plan[0] == 256 and M > 128 else plan


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_grouped_matches_oracle_and_standalone(mode):
    torch = torch_cuda()
    probs, host = build(SPECS, 500)
    torch.cuda.synchronize()
    assert mm.sgemm_grouped(as_args(probs), mode=mode) == mm.OK, mm.last_error()
    torch.cuda.synchronize()
    plans = mm.grouped_plans(as_args(probs), mode)  # what the launch ran
    for (dA, dB, dC, M, K, N), (A, B), gplan in zip(probs, host, plans):
        assert dA.guards_intact() and dB.guards_intact() and dC.guards_intact()
        if M == 0 or N == 0:
            assert np.all(dC.numpy((-1,)) == np.float32(SENTINEL))
            continue
        Cg = dC.numpy((M, N))
        if K == 0:
            assert np.all(Cg == 0.0)
            continue
        assert full_check(A, B, Cg)["bound_ratio"] <= 1.0, (M, K, N)
        assert gplan == mm.grouped_default_plan(M, K, N, mode)  # a small launch: the per-problem plan
        plan = standalone_plan(M, gplan)
        # same slices, same products per element, same reduce order (the
        # split-mode grouped kernels run 64-tile plans on 64 x 128 tiles with
        # eight waves: other wave tiles, same products)
        ref = Guarded(M * N, SENTINEL)
        assert mm.sgemm_ex(dA.ptr, dB.ptr, ref.ptr, M, K, N, mode=mode, plan=plan) == mm.OK
        torch.cuda.synchronize()
        assert np.array_equal(ref.numpy((M, N)), Cg), (M, K, N, plan)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_grouped_tail_rows_match_standalone(mode):
    """In plain mode and on the split modes' decoupled feed (their default),
    problems on 128-row tiles whose last tile holds at most 64 rows run as two
    row ranges (128-row tiles, then one 64-row tile); the direct split-mode
    kernel keeps the padded tile. Either way every row gets the standalone
    call's K slices and products: bit-identical results, the bound, and
    nothing written outside C."""
    torch = torch_cuda()
    specs = [(129, 700, 256), (192, 1536, 130), (193, 300, 200), (255, 513, 260), (320, 2048, 512),
             (130, 4100, 384)]
    plans = [(256, 2), (128, 1), (128, 4), (256, 1), (256, 4), (128, 16)]
    probs, host = build(specs, 900)
    gp = mm.GroupedPlan(as_args(probs), mode=mode, plans=plans)
    gp.launch()
    torch.cuda.synchronize()
    gp.close()
    for (dA, dB, dC, M, K, N), (A, B), plan in zip(probs, host, plans):
        assert dC.guards_intact(), (M, K, N)
        Cg = dC.numpy((M, N))
        assert full_check(A, B, Cg)["bound_ratio"] <= 1.0, (M, K, N)
        ref = Guarded(M * N, SENTINEL)
        assert mm.sgemm_ex(dA.ptr, dB.ptr, ref.ptr, M, K, N, mode=mode, plan=standalone_plan(M, plan)) == mm.OK
        torch.cuda.synchronize()
        assert np.array_equal(ref.numpy((M, N)), Cg), (mode, M, K, N, plan)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_grouped_reduce_every_split_count(mode):
    """Thirty-two problems whose K spans 1..32 k-blocks of 32 under plan
    (64, 32): one grouped reduce serves every split count. Small-integer inputs
    make the products exact."""
    torch = torch_cuda()
    specs = [(40 + s, 32 * s, 70 + 3 * s) for s in range(1, 33)]
    probs, host = build(specs, 800, gen=lambda seed, shape: orc.small_ints(seed, shape, 8))
    gp = mm.GroupedPlan(as_args(probs), mode=mode, plans=[(64, 32)] * len(specs))
    gp.launch()
    torch.cuda.synchronize()
    gp.close()
    for (dA, dB, dC, M, K, N), (A, B) in zip(probs, host):
        exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
        got = dC.numpy((M, N))
        assert np.array_equal(got, exact), f"mode {mode} K {K}: {int((got != exact).sum())} wrong"
        assert dC.guards_intact()


def test_graph_replays_are_identical():
    torch = torch_cuda()
    probs, host = build(SPECS, 600)
    gp = mm.GroupedPlan(as_args(probs), mode=1)
    info = gp.info()
    assert info["graph"] and info["work_items"] > 0
    assert info["flops"] == sum(2.0 * M * K * N for (M, K, N) in SPECS)
    results = []
    for rep in range(3):
        for dA, dB, dC, M, K, N in probs:
            dC.view.fill_(-1.0)  # eager work between replays
        torch.cuda.synchronize()
        gp.launch()
        torch.cuda.synchronize()
        results.append([dC.numpy((-1,)).copy() for (_, _, dC, _, _, _) in probs])
    for rep in (1, 2):
        for a, b, (M, K, N) in zip(results[0], results[rep], SPECS):
            if M * N:
                assert np.array_equal(a, b), (rep, M, K, N)
    for (dA, dB, dC, M, K, N), (A, B) in zip(probs, host):
        if M * N * K:
            assert full_check(A, B, dC.numpy((M, N)))["bound_ratio"] <= 1.0
    gp.close()


def test_validate_all_first_touches_nothing():
    torch = torch_cuda()
    probs, host = build(SPECS[:6], 700)
    args = as_args(probs)
    M, K, N = SPECS[3]
    hostA = np.ones((M, K), np.float32)
    bad = list(args)
    bad[3] = (hostA.ctypes.data, args[3][1], args[3][2], M, K, N)  # host pointer
    torch.cuda.synchronize()
    assert mm.sgemm_grouped(bad) == mm.ERR_UNSUPPORTED
    torch.cuda.synchronize()
    for dA, dB, dC, *_ in probs:
        assert np.all(dC.numpy((-1,)) == np.float32(SENTINEL))
    # an invalid explicit plan is refused before anything runs
    with pytest.raises(mm.MaxMulSKError):
        mm.GroupedPlan(args, plans=[(64, 1)] * 5 + [(32, 1)])
    torch.cuda.synchronize()
    for dA, dB, dC, *_ in probs:
        assert np.all(dC.numpy((-1,)) == np.float32(SENTINEL))


def test_config4_style_batch_sampled():
    """64 problems drawn like the 256-problem workload (M ~ U{1..512},
    N in {512, 2112, 4096, 7168}, K in {512, 1536, 7168, 16384}); inputs alias
    slices of one random pool; 24 sampled outputs per problem against FP64."""
    torch = torch_cuda()
    rng = np.random.default_rng(4)
    specs = [(int(rng.integers(1, 513)), int(rng.choice([512, 1536, 7168, 16384])),
              int(rng.choice([512, 2112, 4096, 7168]))) for _ in range(64)]
    pool_n = 16384 * 7168 + 1024
    pool = torch.from_numpy(orc.uniform(77, (pool_n,))).cuda()
    outs = []
    args = []
    meta = []
    for M, K, N in specs:
        oa = int(rng.integers(0, pool_n - M * K)) & ~3
        ob = int(rng.integers(0, pool_n - K * N)) & ~3
        dC = torch.full((M * N,), SENTINEL, dtype=torch.float32, device="cuda")
        outs.append(dC)
        args.append((pool.data_ptr() + 4 * oa, pool.data_ptr() + 4 * ob, dC.data_ptr(), M, K, N))
        meta.append((oa, ob))
    torch.cuda.synchronize()
    for mode in (0, 1):
        gp = mm.GroupedPlan(args, mode=mode)
        gp.launch()
        torch.cuda.synchronize()
        poolh = pool.cpu().numpy()
        for (M, K, N), (oa, ob), dC in zip(specs, meta, outs):
            Ch = dC.cpu().numpy().reshape(M, N)
            A = poolh[oa:oa + M * K].reshape(M, K)
            ii = rng.integers(0, M, 24)
            jj = rng.integers(0, N, 24)
            Bc = poolh[ob:ob + K * N].reshape(K, N)[:, jj].astype(np.float64)
            Ar = A[ii].astype(np.float64)
            ref = np.sum(Ar * Bc.T, axis=1)
            absr = np.sum(np.abs(Ar) * np.abs(Bc.T), axis=1)
            d = np.abs(Ch[ii, jj].astype(np.float64) - ref)
            assert np.all(d <= K * orc.U32 * absr), (mode, M, K, N)
        gp.close()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_config4_batch_every_element(mode):
    """The 256-problem batch bench.py times (same seed, same aliasing of one
    input pool): every element of every problem's C against FP64 through the
    GPU oracle, within K * 2^-24 * |A||B|, after two graph replays."""
    torch = torch_cuda()
    import bench
    from maxmulsk_amd import inputs
    from oracle import gpu_oracle as go

    specs, pool, outs, probs = bench.config4_batch(torch, inputs)
    for c in outs:
        c.fill_(SENTINEL)
    gp = mm.GroupedPlan(probs, mode=mode)
    gp.launch()
    gp.launch()
    torch.cuda.synchronize()
    gp.close()
    worst = 0.0
    for (a, b, c, M, K, N) in probs:
        e = go.check(a, b, c, M, K, N)
        assert e["nonfinite"] == 0 and e["bound_ratio"] <= 1.0, (mode, M, K, N, e)
        worst = max(worst, e["bound_ratio"])
    print(f"mode {mode}: 256 problems, worst bound ratio {worst:.3g}")
