"""The grouped persistent launch on the decoupled feed (split modes, the
default there): the direct grouped kernel's work items and tiles, fed by
producer waves, as the reference's Trio feeds its compute thread
(sme/trio/trio_sme.cpp:274-334) and run_batch runs a batch
(sme/v4/sme-1x4-v4c-onestream.cpp:683-694). Every element equals the direct
grouped launch's and the standalone call's of the same plan bit for bit (-0
may become +0), within the FP32 bound; nothing is written outside C; a
handshake that runs out of polls ends the launch with NaN instead of a
hang."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import maxmulsk_amd as mm
from oracle import oracle as orc
from tests.gpu_util import SENTINEL, Guarded, full_check, torch_cuda
from tests.test_grouped_gpu import SPECS, as_args, build, standalone_plan

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def feed_setting():
    torch_cuda()
    prev = mm.get_feed()
    yield
    mm.set_feed(prev)


def run_grouped(probs, mode, feed, plans=None):
    """Launches the problems' batch under a feed setting; returns whether the
    plan took the feed."""
    torch = torch_cuda()
    mm.set_feed(feed)
    gp = mm.GroupedPlan(as_args(probs), mode=mode, plans=plans)
    used = gp.info()["feed"]
    gp.launch()
    torch.cuda.synchronize()
    gp.close()
    return used


def check_against_direct(specs, mode, seed, plans=None, gen=orc.uniform):
    torch = torch_cuda()
    fed, host = build(specs, seed, gen)
    direct, _ = build(specs, seed, gen)
    assert run_grouped(fed, mode, mm.FEED_DECOUPLED, plans)
    assert not run_grouped(direct, mode, mm.FEED_DIRECT, plans)
    for (dA, dB, dC, M, K, N), (_, _, rC, _, _, _), (A, B) in zip(fed, direct, host):
        assert dA.guards_intact() and dB.guards_intact() and dC.guards_intact(), (M, K, N)
        if M == 0 or N == 0:
            assert np.all(dC.numpy((-1,)) == np.float32(SENTINEL))
            continue
        Cf = dC.numpy((M, N))
        if K == 0:
            assert np.all(Cf == 0.0)
            continue
        assert np.array_equal(Cf, rC.numpy((M, N))), (mode, M, K, N)
        assert full_check(A, B, Cf)["bound_ratio"] <= 1.0, (mode, M, K, N)
    torch.cuda.synchronize()
    return fed, host


@pytest.mark.parametrize("mode", [1, 2])
def test_grouped_feed_matches_direct_and_standalone(feed_setting, mode):
    """The mixed batch (empty, K = 0, M <= 64, ragged, long K) with its default
    plans: identical to the direct grouped launch and to standalone calls."""
    torch = torch_cuda()
    fed, host = check_against_direct(SPECS, mode, 500)
    mm.set_feed(mm.FEED_DIRECT)
    plans = mm.grouped_plans(as_args(fed), mode)
    for (dA, dB, dC, M, K, N), plan in zip(fed, plans):
        if M == 0 or N == 0 or K == 0:
            continue
        ref = Guarded(M * N, SENTINEL)
        assert mm.sgemm_ex(dA.ptr, dB.ptr, ref.ptr, M, K, N, mode=mode, plan=standalone_plan(M, plan)) == mm.OK
        torch.cuda.synchronize()
        assert np.array_equal(ref.numpy((M, N)), dC.numpy((M, N))), (mode, M, K, N, plan)


@pytest.mark.parametrize("mode", [1, 2])
def test_grouped_feed_every_plan_edge(feed_setting, mode):
    """Plans of every tile edge and several K slices, on ragged problems with
    partial last k-blocks and N below one tile, so every geometry of the
    kernel runs."""
    specs = [(129, 700, 256), (192, 1536, 130), (193, 300, 200), (255, 513, 260), (320, 2048, 512),
             (130, 4100, 384), (64, 3000, 900), (33, 4096, 77), (300, 777, 3), (1, 2500, 1000)]
    plans = [(256, 2), (128, 1), (128, 4), (256, 1), (256, 4), (128, 16), (64, 2), (256, 8), (64, 1), (128, 3)]
    check_against_direct(specs, mode, 900, plans)


@pytest.mark.parametrize("mode", [1, 2])
def test_grouped_feed_reduce_every_split_count(feed_setting, mode):
    """Thirty-two problems whose K spans 1..32 k-blocks of 32 under plan
    (64, 32) (so slices of 32, shorter than the feed's k-blocks): exact on
    small-integer inputs."""
    specs = [(40 + s, 32 * s, 70 + 3 * s) for s in range(1, 33)]
    gen = lambda seed, shape: orc.small_ints(seed, shape, 8)
    fed, host = check_against_direct(specs, mode, 800, [(64, 32)] * len(specs), gen)
    for (dA, dB, dC, M, K, N), (A, B) in zip(fed, host):
        exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
        assert np.array_equal(dC.numpy((M, N)), exact), (mode, K)


@pytest.mark.parametrize("mode", [1, 2])
def test_grouped_feed_config4_batch(feed_setting, mode):
    """The 256-problem batch bench.py times: three graph replays on the feed,
    every one identical to the direct grouped launch, every element within
    the FP32 bound (GPU FP64 oracle)."""
    torch = torch_cuda()
    import bench
    from maxmulsk_amd import inputs
    from oracle import gpu_oracle as go

    specs, pool, outs, probs = bench.config4_batch(torch, inputs)
    mm.set_feed(mm.FEED_DIRECT)
    gd = mm.GroupedPlan(probs, mode=mode)
    assert not gd.info()["feed"]
    gd.launch()
    torch.cuda.synchronize()
    ref = [c.clone() for c in outs]
    gd.close()
    mm.set_feed(mm.FEED_DECOUPLED)
    gf = mm.GroupedPlan(probs, mode=mode)
    assert gf.info()["feed"] and gf.info()["graph"]
    for replay in range(3):
        for c in outs:
            c.fill_(SENTINEL)
        gf.launch()
        torch.cuda.synchronize()
        same = sum(bool(torch.equal(c, r)) for c, r in zip(outs, ref))
        assert same == len(outs), f"replay {replay}: {len(outs) - same} problems differ"
    gf.close()
    worst = 0.0
    for (a, b, c, M, K, N) in probs:
        e = go.check(a, b, c, M, K, N)
        assert e["nonfinite"] == 0 and e["bound_ratio"] <= 1.0, (mode, M, K, N, e)
        worst = max(worst, e["bound_ratio"])
    print(f"mode {mode}: 256 problems on the feed, worst bound ratio {worst:.3g}")


def test_grouped_feed_is_for_split_modes(feed_setting):
    """Split-mode grouped launches take the feed under the automatic and the
    forced setting and the direct kernel under the direct one; plain grouped
    launches keep their own kernel."""
    probs, _ = build([(100, 512, 300), (37, 1000, 64)], 50)
    for mode in (1, 2):
        assert run_grouped(probs, mode, mm.FEED_AUTO)
        assert run_grouped(probs, mode, mm.FEED_DECOUPLED)
        assert not run_grouped(probs, mode, mm.FEED_DIRECT)
    for feed in (mm.FEED_AUTO, mm.FEED_DECOUPLED, mm.FEED_DIRECT):
        assert not run_grouped(probs, 0, feed)


ABORT_SCRIPT = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import torch
import maxmulsk_amd as mm
from oracle import oracle as orc
specs = [(256, 4096, 256), (40, 2048, 300), (130, 999, 129)]
host = [(orc.uniform(1 + 2 * i, (M, K)), orc.uniform(2 + 2 * i, (K, N))) for i, (M, K, N) in enumerate(specs)]
dev = [(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()) for A, B in host]
def run(mode, feed):
    mm.set_feed(feed)
    outs = [torch.full((M, N), 3.75, dtype=torch.float32, device="cuda") for M, K, N in specs]
    probs = [(a.data_ptr(), b.data_ptr(), c.data_ptr(), M, K, N) for (a, b), c, (M, K, N) in zip(dev, outs, specs)]
    gp = mm.GroupedPlan(probs, mode=mode)
    used = gp.info()["feed"]
    gp.launch()
    torch.cuda.synchronize()
    gp.close()
    return used, [c.cpu().numpy() for c in outs]
out = {}
for mode in (1, 2):
    used_f, Cf = run(mode, mm.FEED_DECOUPLED)
    used_d, Cd = run(mode, mm.FEED_DIRECT)
    out[str(mode)] = {"feed": [used_f, used_d], "nan": int(sum(np.isnan(c).sum() for c in Cf)),
                      "finite_equal_direct": all(bool(np.array_equal(f[np.isfinite(f)], d[np.isfinite(f)]))
                                                 for f, d in zip(Cf, Cd)),
                      "direct_ratio": max(orc.errors(d, orc.gemm_f64(A, B), orc.absgemm_f64(A, B), A.shape[1])["bound_ratio"]
                                          for d, (A, B) in zip(Cd, host))}
print("RESULT " + json.dumps(out), flush=True)
"""


def test_grouped_feed_exhausted_handshake_ends_with_nan_not_a_hang():
    """MAXMULSK_TEST_FEED_SPIN_LIMIT=0: every wait that is not already
    satisfied gives up at once, so workgroups abort their first item and every
    later one: the persistent kernel still drains its work list and ends, C
    holds NaN (or, where no wait was needed, the direct numbers), and the
    direct grouped launch in the same process is still correct."""
    torch_cuda()
    env = dict(os.environ, MAXMULSK_TEST_FEED_SPIN_LIMIT="0")
    print("grouped feed abort subprocess: spin limit 0, modes 1 and 2", flush=True)
    r = subprocess.run([sys.executable, "-c", ABORT_SCRIPT, ROOT], env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("RESULT ")][-1]
    out = json.loads(line[len("RESULT "):])
    for mode in ("1", "2"):
        o = out[mode]
        assert o["feed"] == [True, False], o
        assert o["nan"] > 0, o
        assert o["finite_equal_direct"], o
        assert o["direct_ratio"] <= 1.0, o
