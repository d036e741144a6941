"""The reference's 35 benchmark shapes through the planner in every mode.

The shape list is the one of the reference's final comparison
(bench/results/2026-09-13/final5.csv), as recorded row by row in
tests/golden/ref_plan_rows.npz: squares 256..4096, M = 64/128 rows against
wide or deep operands, and the LLM-style 4096..13824 x 256 shapes. Like the
reference's full-path check (check_nb.cpp:41-81: every shape, C prefilled with
a sentinel that must be overwritten, a repeat call with a new B), every call
here starts from a sentinel-filled C, and every element of the result is held
to |C - C64| <= K * 2^-24 * |A||B| by the FP64 GPU checker. Plain mode goes
through maxmulsk_sgemm_v4, the reference's own entry point; the split modes
through maxmulsk_sgemm_ex with the planner. Inputs are uniform[-1, 1) from
torch's seeded device generator (the checker needs no host copy)."""
import os

import numpy as np
import pytest

import maxmulsk_amd as mm
from tests.gpu_util import SENTINEL, report, torch_cuda

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = [tuple(int(x) for x in r[:3])
          for r in np.load(os.path.join(HERE, "golden", "ref_plan_rows.npz"))["final5"]]  # (M, K, N)


def _uniform(torch, gen, n):
    return torch.rand(n, generator=gen, device="cuda", dtype=torch.float32) * 2 - 1


def test_shape_list_is_the_reference_final_comparison():
    assert len(SHAPES) == 35
    assert (4096, 4096, 4096) in SHAPES and (64, 8192, 512) in SHAPES and (13824, 5120, 256) in SHAPES


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_reference_shapes_every_element_within_fp32_bound(mode):
    torch = torch_cuda()
    from oracle import gpu_oracle as go

    gen = torch.Generator(device="cuda")
    worst = (0.0, None)
    for i, (M, K, N) in enumerate(SHAPES):
        gen.manual_seed(1000 * mode + i)
        dA, dB = _uniform(torch, gen, M * K), _uniform(torch, gen, K * N)
        dC = torch.full((M * N,), SENTINEL, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        if mode == 0:
            st = mm.sgemm_v4(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N)
        else:
            st = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=mode)
        assert st == mm.OK, (M, K, N, mm.last_error())
        torch.cuda.synchronize()
        e = go.check(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N)
        assert e["nonfinite"] == 0 and e["bound_ratio"] <= 1.0, (mode, M, K, N, e)
        report("reference_shapes", shape_mkn=[M, K, N], mode=mode, **e)
        if e["bound_ratio"] > worst[0]:
            worst = (e["bound_ratio"], (M, K, N))
        del dA, dB, dC
    print(f"mode {mode}: 35 reference shapes, worst bound ratio {worst[0]:.3g} at {worst[1]}")


def test_reference_shapes_repeat_call_with_new_b():
    """Second call on the same buffers with a new B (the reference's stale-packing
    check, check_nb.cpp:61-66), on the shapes that take the packed path (256-wide
    tiles, where both operands are rewritten into the workspace on every call)."""
    torch = torch_cuda()
    from oracle import gpu_oracle as go

    gen = torch.Generator(device="cuda")
    for M, K, N in [(4096, 4096, 4096), (4096, 1536, 24576), (13824, 5120, 256)]:
        gen.manual_seed(M + K + N)
        dA, dB = _uniform(torch, gen, M * K), _uniform(torch, gen, K * N)
        dC = torch.full((M * N,), SENTINEL, dtype=torch.float32, device="cuda")
        for rep in range(2):
            if rep:
                dB.copy_(_uniform(torch, gen, K * N))
                dC.fill_(SENTINEL)
            torch.cuda.synchronize()
            st = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=1, plan=(256, 1))
            assert st == mm.OK, mm.last_error()
            torch.cuda.synchronize()
            e = go.check(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N)
            assert e["nonfinite"] == 0 and e["bound_ratio"] <= 1.0, (M, K, N, rep, e)
        del dA, dB, dC


# The reference's GEMM-vs-scalar unit-test shapes (test_sme.cpp:28-37,431-453), as (M, K, N).
UNIT_SHAPES = [(16, 16, 16), (32, 32, 32), (64, 64, 64), (67, 67, 67), (128, 128, 128), (20, 35, 41),
               (20, 41, 35), (256, 256, 256), (512, 512, 512)]


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_reference_unit_shapes_every_plan(mode):
    """Every candidate plan on the reference's unit-test shapes, every element of
    C within its mode's bound (the reference held these to 1e-3 absolute against
    its scalar loop, utils.hpp:24-25). Plain and split6: the FP32 bound
    K * 2^-24 * |A||B|. Split3 drops lo*lo and each operand's representation
    error, up to 3 * 2^-16 * |a||b| per product whatever K is, so its bound is
    (K * 2^-24 + 3 * 2^-16) * |A||B|: a bound ratio of 1 + 768 / K against the
    FP32 bound, which it exceeds on uniform data below K of about 64."""
    torch = torch_cuda()
    from oracle import gpu_oracle as go

    gen = torch.Generator(device="cuda")
    for M, K, N in UNIT_SHAPES:
        gen.manual_seed(M * 7 + K * 5 + N)
        dA, dB = _uniform(torch, gen, M * K), _uniform(torch, gen, K * N)
        dC = torch.empty((M * N,), dtype=torch.float32, device="cuda")
        for plan in mm.candidates():
            dC.fill_(SENTINEL)
            torch.cuda.synchronize()
            st = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=mode, plan=plan)
            assert st == mm.OK, (M, K, N, plan, mm.last_error())
            torch.cuda.synchronize()
            e = go.check(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N)
            limit = 1.0 + 768.0 / K if mode == 1 else 1.0
            assert e["nonfinite"] == 0 and e["bound_ratio"] <= limit, (mode, M, K, N, plan, e)
            assert e["max_abs"] <= 1e-3, (mode, M, K, N, plan, e)
            if plan == (64, 1):
                report("unit_shapes", shape_mkn=[M, K, N], mode=mode, **e)
