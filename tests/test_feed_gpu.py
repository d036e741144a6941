"""The decoupled producer/consumer feed of the 64 x 64 tile and of the split
modes' 128 x 128 tile (the GPU analogue of the reference's Trio engine,
sme/trio/trio_sme.cpp:274-334, whose result is bit-identical to the v5.1 kernel
it feeds, README.md "MaxMulSK Trio"): every variant gives the direct tile's C
bit for bit (-0 may become +0) on full and ragged tiles, with and without
split-K; exact on integer inputs; the automatic rule launches it where it
measured faster; a handshake that runs out of polls ends the kernel with NaN
in C instead of hanging."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import maxmulsk_amd as mm
from oracle import oracle as orc
from tests.gpu_util import full_check, gpu_gemm, gpu_gemm_checked, torch_cuda

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def feed_setting():
    torch_cuda()
    prev = mm.get_feed()
    yield
    mm.set_feed(prev)


# (M, N, K, plan): one workgroup per CU with long and short K slices (variants
# 1, 2, 3), more than one per CU (4, 5, 6), ragged M, N and K, K below one
# k-block, K not a multiple of 4, split-K partials on small and large grids.
CASES = [
    (1, 1, 1, (64, 1)),
    (5, 7, 3, (64, 1)),
    (64, 64, 32, (64, 1)),
    (33, 65, 129, (64, 2)),
    (96, 160, 300, (64, 1)),
    (192, 320, 1000, (64, 1)),
    (256, 256, 4096, (64, 1)),
    (70, 90, 1023, (64, 2)),
    (67, 513, 8191, (64, 4)),
    (67, 513, 8191, (64, 16)),
    (64, 512, 8192, (64, 32)),
    (1100, 1030, 600, (64, 1)),
]


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("M,N,K,plan", CASES)
def test_forced_feed_matches_direct_tile(feed_setting, mode, M, N, K, plan):
    A = orc.uniform(11 + M, (M, K))
    B = orc.uniform(12 + N, (K, N))
    mm.set_feed(mm.FEED_DIRECT)
    assert mm.feed_variant(M, K, N, mode, plan) == 0
    Cd = gpu_gemm(A, B, mode=mode, plan=plan)
    mm.set_feed(mm.FEED_DECOUPLED)
    v = mm.feed_variant(M, K, N, mode, plan)
    assert 1 <= v <= 6, v
    Cf = gpu_gemm(A, B, mode=mode, plan=plan)  # guard bands around A, B and C checked
    assert not np.isnan(Cf).any()
    assert np.array_equal(Cf, Cd), (mode, M, N, K, plan, v, float(np.abs(Cf - Cd).max()))


# 128 x 128 tiles of the split modes: full tiles and k-blocks (the automatic
# rule's case), ragged M, N and K, a partial last k-block, split-K.
CASES_128 = [
    (256, 256, 2048, (128, 1)),
    (384, 256, 640, (128, 2)),
    (200, 300, 1000, (128, 1)),
    (130, 260, 4099, (128, 4)),
    (1000, 1100, 3000, (128, 1)),
    (129, 129, 33, (128, 1)),
]


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("M,N,K,plan", CASES_128)
def test_forced_feed_128_matches_direct_tile(feed_setting, mode, M, N, K, plan):
    A = orc.uniform(41 + M, (M, K))
    B = orc.uniform(42 + N, (K, N))
    mm.set_feed(mm.FEED_DIRECT)
    Cd = gpu_gemm(A, B, mode=mode, plan=plan)
    mm.set_feed(mm.FEED_DECOUPLED)
    v = mm.feed_variant(M, K, N, mode, plan)
    assert v == (7 if mode == 1 else 8), v
    Cf = gpu_gemm(A, B, mode=mode, plan=plan)
    assert not np.isnan(Cf).any()
    assert np.array_equal(Cf, Cd), (mode, M, N, K, plan, float(np.abs(Cf - Cd).max()))
    # plain mode keeps the direct 128 x 128 tile even when the feed is forced
    assert mm.feed_variant(M, K, N, 0, plan) == 0


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_forced_feed_exact_on_integers(feed_setting, mode):
    """Small integers: every product and partial sum is exact in FP32 (and in
    BF16 hi terms), so C must equal the integer product whatever the order."""
    rng = np.random.default_rng(5 + mode)
    mm.set_feed(mm.FEED_DECOUPLED)
    cases = [(96, 160, 1100, (64, 1)), (130, 70, 4099, (64, 8)), (256, 256, 2048, (64, 1)), (1100, 300, 700, (64, 2))]
    if mode != 0:
        cases += [(256, 384, 2048, (128, 1)), (200, 300, 1000, (128, 2))]
    for M, N, K, plan in cases:
        A = rng.integers(-8, 9, (M, K)).astype(np.float32)
        B = rng.integers(-8, 9, (K, N)).astype(np.float32)
        assert mm.feed_variant(M, K, N, mode, plan) > 0
        Cg = gpu_gemm(A, B, mode=mode, plan=plan)
        ref = A.astype(np.int64) @ B.astype(np.int64)
        assert np.array_equal(Cg.astype(np.int64), ref), (mode, M, N, K, plan)


# The automatic rule, shape by shape, against the variants measured faster.
AUTO = [
    ((1024, 1024, 1024, (64, 1)), {0: 0, 1: 2, 2: 3}),
    ((768, 768, 4096, (64, 1)), {0: 0, 1: 1, 2: 3}),
    ((512, 512, 8192, (64, 4)), {0: 0, 1: 2, 2: 3}),
    ((64, 512, 8192, (64, 32)), {0: 0, 1: 0, 2: 3}),      # split3: slices of 256, direct
    ((64, 512, 8192, (64, 16)), {0: 0, 1: 0, 2: 3}),      # split3: slices of 512
    ((512, 384, 2500, (64, 1)), {0: 0, 1: 2, 2: 3}),
    ((67, 513, 8191, (64, 16)), {0: 0, 1: 0, 2: 0}),      # 288 workgroups: direct
    ((64, 7168, 16384, (64, 8)), {0: 0, 1: 0, 2: 0}),
    ((4096, 4096, 4096, (256, 1)), {0: 0, 1: 0, 2: 0}),   # packed path
    ((1024, 1024, 1024, (128, 1)), {0: 0, 1: 7, 2: 8}),   # 128 x 128, full tiles and k-blocks
    ((1920, 1920, 2000, (128, 1)), {0: 0, 1: 0, 2: 0}),   # partial last k-block, slices < 2048
    ((1000, 1100, 3000, (128, 1)), {0: 0, 1: 7, 2: 0}),   # ragged: split3 from slices of 2048
]


def test_auto_rule_and_results(feed_setting):
    mm.set_feed(mm.FEED_AUTO)
    for (M, N, K, plan), want in AUTO:
        for mode, v in want.items():
            assert mm.feed_variant(M, K, N, mode, plan) == v, (M, N, K, plan, mode)
    # an automatically fed launch gives the direct tile's C, inside the FP32 bound
    for mode, (M, N, K, plan) in ((1, (640, 640, 3000, (64, 1))), (1, (512, 768, 4500, (64, 1))),
                                  (2, (200, 300, 1200, (64, 2))), (2, (128, 1024, 4096, (64, 8))),
                                  (1, (512, 384, 2048, (128, 1))), (2, (384, 512, 1024, (128, 1))),
                                  (1, (300, 500, 2500, (128, 1))), (2, (130, 70, 300, (64, 1)))):
        A = orc.uniform(21, (M, K))
        B = orc.uniform(22, (K, N))
        mm.set_feed(mm.FEED_AUTO)
        assert mm.feed_variant(M, K, N, mode, plan) > 0
        Ca, e = gpu_gemm_checked(A, B, mode=mode, plan=plan)
        assert e["bound_ratio"] <= 1.0, e
        mm.set_feed(mm.FEED_DIRECT)
        Cd = gpu_gemm(A, B, mode=mode, plan=plan)
        assert np.array_equal(Ca, Cd)


def test_planner_calls_under_the_feed(feed_setting):
    """The planner's plan and the reference entry point, with the feed forced
    and automatic: same C as the direct tile."""
    M, N, K = 300, 260, 1536
    A = orc.uniform(31, (M, K))
    B = orc.uniform(32, (K, N))
    mm.set_feed(mm.FEED_DIRECT)
    ref = {mode: gpu_gemm(A, B, mode=mode, plan=(64, 2)) for mode in (0, 1, 2)}
    for feed in (mm.FEED_DECOUPLED, mm.FEED_AUTO):
        mm.set_feed(feed)
        for mode in (0, 1, 2):
            assert np.array_equal(gpu_gemm(A, B, mode=mode, plan=(64, 2)), ref[mode])
        Cv4 = gpu_gemm(A, B, entry="v4")
        assert full_check(A, B, Cv4)["bound_ratio"] <= 1.0


def test_set_feed_rejects_unknown_values(feed_setting):
    mm.set_feed(mm.FEED_AUTO)
    with pytest.raises(ValueError):
        mm.set_feed(3)
    assert mm.get_feed() == mm.FEED_AUTO
    assert mm.set_feed(mm.FEED_DIRECT) == mm.FEED_AUTO


ABORT_SCRIPT = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import torch
import maxmulsk_amd as mm
from oracle import oracle as orc
M, N, K = 256, 256, 4096
A = orc.uniform(1, (M, K)); B = orc.uniform(2, (K, N))
dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
def run(mode):
    dC = torch.full((M, N), 3.75, dtype=torch.float32, device="cuda")
    st = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=mode, plan=(64, 1))
    torch.cuda.synchronize()
    return st, dC.cpu().numpy()
out = {"feed_at_start": mm.get_feed()}
for mode in (1, 2):
    mm.set_feed(mm.FEED_DIRECT)
    st_d, Cd = run(mode)
    mm.set_feed(mm.FEED_DECOUPLED)
    variant = mm.feed_variant(M, K, N, mode, (64, 1))
    st_f, Cf = run(mode)
    fin = np.isfinite(Cf)
    out[str(mode)] = {"variant": variant, "status": [st_d, st_f], "nan": int(np.isnan(Cf).sum()),
                      "finite_equal_direct": bool(np.array_equal(Cf[fin], Cd[fin])),
                      "direct_ratio": orc.errors(Cd, orc.gemm_f64(A, B), orc.absgemm_f64(A, B), K)["bound_ratio"]}
print("RESULT " + json.dumps(out), flush=True)
"""


def test_exhausted_handshake_ends_with_nan_not_a_hang():
    """MAXMULSK_TEST_FEED_SPIN_LIMIT=0: every wait that is not already
    satisfied gives up at once. A consumer's first wait comes before any
    producer can have stored a block (the loads take microseconds), so the
    tiles abort: the kernel must end, the call return OK, and every tile hold
    NaN or, if it never had to wait, the direct tile's numbers. The direct
    tile still runs correctly afterwards in the same process."""
    torch_cuda()
    env = dict(os.environ, MAXMULSK_TEST_FEED_SPIN_LIMIT="0", MAXMULSK_FEED="decoupled")
    r = subprocess.run([sys.executable, "-c", ABORT_SCRIPT, ROOT], env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")][0]
    out = json.loads(line[len("RESULT "):])
    assert out["feed_at_start"] == mm.FEED_DECOUPLED
    for mode in ("1", "2"):
        o = out[mode]
        assert o["variant"] > 0 and o["status"] == [mm.OK, mm.OK], o
        assert o["nan"] > 0 and o["nan"] % 64 == 0, o  # whole tile rows, never a partial tile
        assert o["finite_equal_direct"], o
        assert o["direct_ratio"] <= 1.0, o
