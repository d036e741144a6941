"""Edge cases of the C ABI on the GPU: non-finite inputs, C offsets past 2^32
bytes and 2^31 elements, a very long K, and concurrent callers on several
streams sharing the library's per-device pools."""
import threading
import time

import numpy as np
import pytest

import maxmulsk_amd as mm
from oracle import oracle as orc
from tests.gpu_util import SENTINEL, full_check, gpu_gemm, torch_cuda

pytestmark = pytest.mark.gpu


def _classes(x):
    return np.isnan(x), np.isposinf(x), np.isneginf(x)


@pytest.mark.parametrize("plan", [(64, 1), (128, 4), (256, 1)])
def test_nonfinite_inputs(plan):
    """Plain FP32 MFMA is IEEE: NaN, +Inf and -Inf land exactly where the FP64
    product has them (Inf + finite and Inf - Inf do not depend on the
    summation order). The split modes turn an infinite input into NaN terms
    (hi - x is not finite), so they are checked for non-finite exactly where the
    FP64 product is non-finite; every other entry stays finite and inside the
    FP32 bound."""
    M, N, K = 300, 260, 320
    A = orc.uniform(91, (M, K))
    B = orc.uniform(92, (K, N))
    A[3, 10] = np.inf
    A[7, 20] = np.nan
    A[250, 40] = -np.inf
    B[30, 5] = -np.inf
    B[41, 200] = np.nan
    with np.errstate(invalid="ignore", over="ignore"):
        C64 = A.astype(np.float64) @ B.astype(np.float64)
    bad = ~np.isfinite(C64)
    Af, Bf = np.where(np.isfinite(A), A, 0).astype(np.float32), np.where(np.isfinite(B), B, 0).astype(np.float32)
    ok_rows = ~np.any(~np.isfinite(A), axis=1)
    ok_cols = ~np.any(~np.isfinite(B), axis=0)
    for mode in (0, 1, 2):
        Cg = gpu_gemm(A, B, mode=mode, plan=plan)
        if mode == 0:
            for g, r in zip(_classes(Cg), _classes(C64)):
                assert np.array_equal(g, r), (mode, plan)
        else:
            assert np.array_equal(~np.isfinite(Cg), bad), (mode, plan)
        sub = np.ix_(ok_rows, ok_cols)
        e = orc.errors(Cg[sub], orc.gemm_f64(Af, Bf)[sub], orc.absgemm_f64(Af, Bf)[sub], K)
        assert e["nonfinite"] == 0 and e["bound_ratio"] <= 1.0, (mode, plan, e)


@pytest.mark.parametrize("mode", [0, 1])
def test_c_past_4_gib_and_2_31_elements(mode):
    """48000 x 48000 C (9.2 GB, 2.3e9 elements): 64-bit row offsets in every
    kernel; rows and columns at both ends are checked exactly (small integers)."""
    torch = torch_cuda()
    M = N = 48000
    K = 4
    A = orc.small_ints(101, (M, K), 4)
    B = orc.small_ints(102, (K, N), 4)
    dA = torch.from_numpy(A).cuda()
    dB = torch.from_numpy(B).cuda()
    dC = torch.full((M, N), SENTINEL, dtype=torch.float32, device="cuda")
    assert mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=mode) == mm.OK, mm.last_error()
    torch.cuda.synchronize()
    A64, B64 = A.astype(np.int64), B.astype(np.int64)
    for r in (0, 1, 23999, 44739, 47998, 47999):
        assert np.array_equal(dC[r].cpu().numpy(), (A64[r] @ B64).astype(np.float32)), r
    for c in (0, 31, 44739, 47999):
        assert np.array_equal(dC[:, c].cpu().numpy(), (A64 @ B64[:, c]).astype(np.float32)), c
    del dC
    torch.cuda.empty_cache()


@pytest.mark.parametrize("plan", [(64, 1), (64, 16), (256, 4)])
def test_very_long_k(plan):
    """K = 2^21 + 13: long k loops and slices with a ragged tail; integers in
    [-2, 2] keep every partial sum below 2^24, so the result is bit-exact for
    any summation order."""
    M, N, K = 8, 12, (1 << 21) + 13
    A = orc.small_ints(111, (M, K), 2)
    B = orc.small_ints(112, (K, N), 2)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
    for mode in (0, 1):
        assert np.array_equal(gpu_gemm(A, B, mode=mode, plan=plan), exact), (mode, plan)


def test_concurrent_callers_on_several_streams():
    """Four threads, each on its own stream with split-K plans that use the
    shared workspace pool, plus one thread on host pointers (staging pool)."""
    torch = torch_cuda()
    M, N, K = 192, 320, 4096
    errors = []

    def device_worker(seed):
        try:
            A = orc.uniform(seed, (M, K))
            B = orc.uniform(seed + 1, (K, N))
            dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
            dC = torch.empty((M, N), dtype=torch.float32, device="cuda")
            s = torch.cuda.Stream()
            for it in range(6):
                plan = [(64, 4), (128, 8), (256, 2)][it % 3]
                st = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=1 + it % 2,
                                 plan=plan, stream=s)
                if st != mm.OK:
                    errors.append((seed, it, st))
                    return
                e = full_check(A, B, dC.cpu().numpy())
                if e["bound_ratio"] > 1.0:
                    errors.append((seed, it, e))
        except Exception as ex:  # reported below
            errors.append((seed, repr(ex)))

    def host_worker():
        try:
            A = orc.uniform(500, (M, K))
            B = orc.uniform(501, (K, N))
            for _ in range(4):
                Ch = np.full((M, N), SENTINEL, np.float32)
                st = mm.sgemm_v4(A, B, Ch, M, K, N)
                if st != mm.OK or full_check(A, B, Ch)["bound_ratio"] > 1.0:
                    errors.append(("host", st))
        except Exception as ex:
            errors.append(("host", repr(ex)))

    threads = [threading.Thread(target=device_worker, args=(300 + 10 * i,)) for i in range(4)]
    threads.append(threading.Thread(target=host_worker))
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors


def test_async_calls_alternating_streams_share_the_pool():
    """Asynchronous split-K calls alternate between two streams with no host
    synchronisation in between; all of them use the library's one pooled
    workspace. Each call must wait for the previous user of the pool on the
    other stream, or partials would be overwritten while still being reduced."""
    torch = torch_cuda()
    M, N, K = 128, 1024, 8192
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    calls = []
    for i in range(8):
        A = orc.uniform(900 + 2 * i, (M, K))
        B = orc.uniform(901 + 2 * i, (K, N))
        dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        dC = torch.full((M, N), SENTINEL, dtype=torch.float32, device="cuda")
        calls.append((A, B, dA, dB, dC))
    torch.cuda.synchronize()
    for i, (A, B, dA, dB, dC) in enumerate(calls):
        st = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=1 + i % 2,
                         plan=(64, 16), stream=streams[i % 2], async_=True)
        assert st == mm.OK, mm.last_error()
    torch.cuda.synchronize()
    for i, (A, B, dA, dB, dC) in enumerate(calls):
        assert full_check(A, B, dC.cpu().numpy())["bound_ratio"] <= 1.0, i


def test_async_calls_mixing_the_default_stream_and_a_created_one():
    """Consecutive calls on the legacy default stream skip the pool's event
    wait (the stream orders them); a call that switches between it and a
    created stream must still wait. Each call on the created stream sits
    behind a bounded GPU sleep, so a missing wait would let the next
    default-stream call overwrite partials that are still being reduced."""
    torch = torch_cuda()
    M, N, K = 128, 1024, 8192
    side = torch.cuda.Stream()
    pattern = [0, 0, 1, 0, 0, 1, 1, 0, 1, 0]
    calls = []
    for i in range(len(pattern)):
        A = orc.uniform(950 + 2 * i, (M, K))
        B = orc.uniform(951 + 2 * i, (K, N))
        dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        dC = torch.full((M, N), SENTINEL, dtype=torch.float32, device="cuda")
        calls.append((A, B, dA, dB, dC))
    torch.cuda.synchronize()
    for i, (A, B, dA, dB, dC) in enumerate(calls):
        stream = 0
        if pattern[i]:
            with torch.cuda.stream(side):
                torch.cuda._sleep(200000)  # bounded spin, about 0.1 ms
            stream = side.cuda_stream
        st = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=1 + i % 2,
                         plan=(64, 16), stream=stream, async_=True)
        assert st == mm.OK, mm.last_error()
    torch.cuda.synchronize()
    for i, (A, B, dA, dB, dC) in enumerate(calls):
        assert full_check(A, B, dC.cpu().numpy())["bound_ratio"] <= 1.0, i


def test_synchronous_calls_on_different_streams_overlap():
    """A synchronous call waits for its own stream only: thread A's call sits
    behind a bounded GPU sleep on stream A while thread B's call on stream B
    starts later and returns first (the per-device pool lock is not held
    across the wait). Both results stay correct."""
    torch = torch_cuda()
    M, N, K = 96, 640, 2048
    A = orc.uniform(701, (M, K))
    B = orc.uniform(702, (K, N))
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    cA = torch.full((M, N), SENTINEL, dtype=torch.float32, device="cuda")
    cB = torch.full((M, N), SENTINEL, dtype=torch.float32, device="cuda")
    sA, sB = torch.cuda.Stream(), torch.cuda.Stream()
    plan = (64, 1)  # no split-K: neither call takes the pooled workspace
    assert mm.workspace_bytes(M, K, N, 1, plan) == 0
    # warm up: kernels loaded, selector and pools initialised
    for s, c in ((sA, cA), (sB, cB)):
        assert mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), c.data_ptr(), M, K, N, mode=1, plan=plan, stream=s) == mm.OK
    torch.cuda.synchronize()
    done, status = {}, {}
    entered = threading.Event()

    def call_a():
        with torch.cuda.stream(sA):
            torch.cuda._sleep(int(2e9))  # about a second of GPU time on stream A
        entered.set()
        status["a"] = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), cA.data_ptr(), M, K, N, mode=1, plan=plan, stream=sA)
        done["a"] = time.perf_counter()

    def call_b():
        entered.wait(30)
        time.sleep(0.1)  # A is inside its call, waiting on stream A
        status["b"] = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), cB.data_ptr(), M, K, N, mode=1, plan=plan, stream=sB)
        done["b"] = time.perf_counter()

    ta, tb = threading.Thread(target=call_a), threading.Thread(target=call_b)
    ta.start()
    tb.start()
    ta.join(60)
    tb.join(60)
    assert not ta.is_alive() and not tb.is_alive()
    assert status == {"a": mm.OK, "b": mm.OK}, (status, mm.last_error())
    assert done["b"] < done["a"], "B waited for A's stream"
    torch.cuda.synchronize()
    for c in (cA, cB):
        assert full_check(A, B, c.cpu().numpy())["bound_ratio"] <= 1.0
