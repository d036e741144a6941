"""Ragged and degenerate shapes on every mode: sizes around the tile and
k-block edges (1, 2, 3, 31..33, 63..65, 127..129, 255..257, 511..513) drawn
with a fixed seed, each run with the planner and with a random candidate plan
(every split-K count up to 32 among them), C prefilled with a sentinel and
guarded by poisoned bands, every element checked against FP64 on the GPU.
Plain is held to the FP32 bound K * 2^-24 * |A||B|. The split modes drop
terms that do not shrink with K: split3 up to 3 * 2^-16 * |a||b| per product,
held to (K * 2^-24 + 3 * 2^-16) * |A||B| (a ratio of 1 + 768 / K against the
FP32 bound); split6 under 4 * 2^-24 * |a||b| per product, held to
(K + 4) * 2^-24 * |A||B| (1 + 4 / K), which only matters for K of a few."""
import numpy as np
import pytest

import maxmulsk_amd as mm
from oracle import oracle as orc
from tests.gpu_util import SENTINEL, Guarded, report, torch_cuda

pytestmark = pytest.mark.gpu

EDGES = [1, 2, 3, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513]


def _shapes(n, seed):
    rng = np.random.default_rng(seed)
    out = set()
    while len(out) < n:
        out.add(tuple(int(x) for x in rng.choice(EDGES, 3)))
    return sorted(out)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_edge_shapes_planner_and_random_plans(mode):
    torch_cuda()
    import torch
    from oracle import gpu_oracle as go

    rng = np.random.default_rng(100 + mode)
    cands = mm.candidates()
    for i, (M, K, N) in enumerate(_shapes(40, 7 + mode)):
        A = orc.uniform(5000 + 97 * i + mode, (M, K))
        B = orc.uniform(6000 + 89 * i + mode, (K, N))
        dA, dB = Guarded(M * K, A), Guarded(K * N, B)
        limit = {0: 1.0, 1: 1.0 + 768.0 / K, 2: 1.0 + 4.0 / K}[mode]
        for plan in (None, cands[int(rng.integers(len(cands)))]):
            dC = Guarded(M * N, SENTINEL)
            torch.cuda.synchronize()
            st = mm.sgemm_ex(dA.ptr, dB.ptr, dC.ptr, M, K, N, mode=mode, plan=plan)
            assert st == mm.OK, (M, K, N, plan, mm.last_error())
            torch.cuda.synchronize()
            assert dC.guards_intact(), (M, K, N, plan)
            e = go.check(dA.ptr, dB.ptr, dC.ptr, M, K, N)
            assert e["nonfinite"] == 0 and e["bound_ratio"] <= limit, (mode, M, K, N, plan, e)
            report("edge_shapes", shape_mkn=[M, K, N], mode=mode, plan=list(plan) if plan else None, **e)
