"""The autotuners on the GPU: every candidate timed and gated, the winner cached
and used by the planner; the grouped autotuner's class plans run correctly."""
import numpy as np
import pytest

import maxmulsk_amd as mm
from oracle import oracle as orc
from tests.gpu_util import SENTINEL, Guarded, full_check, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1])
def test_autotune_times_every_candidate_and_caches_the_winner(mode):
    torch = torch_cuda()
    lib = mm.lib()
    lib.maxmulsk_cache_clear()
    M, K, N = 96, 3000, 640
    A = orc.uniform(91, (M, K))
    B = orc.uniform(92, (K, N))
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL)
    torch.cuda.synchronize()
    times, best = mm.autotune(dA.ptr, dB.ptr, dC.ptr, M, K, N, mode, 2, 4)
    assert len(times) == 18 and all(t > 0 for t in times), times
    assert best in mm.candidates()
    # the winner is at least as fast as every candidate re-timed in stage 2
    assert times[mm.candidates().index(best)] <= min(times) * 1.03 + 1e-3
    assert mm.choose_plan(M, K, N, mode) == (best, 1)
    assert dA.guards_intact() and dB.guards_intact() and dC.guards_intact()
    # the cached plan now drives the reference entry point (plain mode)
    if mode == 0:
        dC.view.fill_(SENTINEL)
        torch.cuda.synchronize()
        assert mm.sgemm_v4(dA.ptr, dB.ptr, dC.ptr, M, K, N) == mm.OK
        torch.cuda.synchronize()
        assert full_check(A, B, dC.numpy((M, N)))["bound_ratio"] <= 1.0
    lib.maxmulsk_cache_clear()


def test_autotune_rejects_host_pointers():
    torch_cuda()
    A = np.ones((8, 8), np.float32)
    with pytest.raises(mm.MaxMulSKError):
        mm.autotune(A, A, A, 8, 8, 8)


def test_grouped_autotune_small_batch():
    torch = torch_cuda()
    specs = [(17, 1536, 2112), (300, 512, 512), (64, 7168, 512), (500, 1536, 4096)] * 3
    probs, host = [], []
    for i, (M, K, N) in enumerate(specs):
        A = orc.uniform(300 + i, (M, K))
        B = orc.uniform(400 + i, (K, N))
        probs.append((Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL), M, K, N))
        host.append((A, B))
    args = [(a.ptr, b.ptr, c.ptr, M, K, N) for a, b, c, M, K, N in probs]
    torch.cuda.synchronize()
    plans, best_ms = mm.grouped_autotune(args, mode=0, max_classes=2, reps=2)
    assert best_ms > 0 and len(plans) == len(specs)
    # problems of one shape class share a plan, and the class choice is cached
    for (M, K, N), p in zip(specs, plans):
        assert mm.grouped_default_plan(M, K, N, 0) == p
    gp = mm.GroupedPlan(args, mode=0)
    gp.launch()
    torch.cuda.synchronize()
    for (dA, dB, dC, M, K, N), (A, B) in zip(probs, host):
        assert full_check(A, B, dC.numpy((M, N)))["bound_ratio"] <= 1.0
    gp.close()


@pytest.mark.parametrize("reject", [0, 7])
def test_autotune_gate_excludes_a_wrong_candidate(reject, monkeypatch):
    """The gate checks every candidate's sampled outputs against FP64 dot
    products, not against another candidate: a candidate whose sample reads
    wrong (test hook: a NaN in its sample) is excluded even when it is the
    first one, and every other candidate still qualifies."""
    torch = torch_cuda()
    lib = mm.lib()
    lib.maxmulsk_cache_clear()
    M, K, N = 80, 2000, 520
    A = orc.uniform(93, (M, K))
    B = orc.uniform(94, (K, N))
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL)
    torch.cuda.synchronize()
    monkeypatch.setenv("MAXMULSK_TEST_GATE_REJECT", str(reject))
    times, best = mm.autotune(dA.ptr, dB.ptr, dC.ptr, M, K, N, 1, 2, 3)
    monkeypatch.delenv("MAXMULSK_TEST_GATE_REJECT")
    assert times[reject] < 0, times
    assert all(t > 0 for i, t in enumerate(times) if i != reject), times
    assert best != mm.candidates()[reject]
    lib.maxmulsk_cache_clear()


def test_autotune_gate_accepts_split3_on_wide_dynamic_range():
    """The gate's bound follows the mode: split3 on log-uniform inputs (2^-20..2^20)
    stays eligible, with every candidate accepted."""
    torch = torch_cuda()
    lib = mm.lib()
    lib.maxmulsk_cache_clear()
    M, K, N = 130, 4000, 300
    A = orc.log_uniform(95, (M, K))
    B = orc.log_uniform(96, (K, N))
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL)
    torch.cuda.synchronize()
    times, best = mm.autotune(dA.ptr, dB.ptr, dC.ptr, M, K, N, 1, 2, 3)
    assert all(t > 0 for t in times), times
    lib.maxmulsk_cache_clear()
