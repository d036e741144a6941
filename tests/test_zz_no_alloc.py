"""maxmulsk_sgemm_noalloc (workspace in the library's static device scratch)
against maxmulsk_sgemm_ws (caller workspace), through the C ABI: the same plan
runs the same kernels, so the results must be identical bit for bit, on two
calls in a row (the second smaller, reusing the scratch the first left
behind), with no allocation counted by the library and with sizes above the
scratch refused."""
import numpy as np
import pytest

import maxmulsk_amd as mm
from oracle import oracle as orc
from tests.gpu_util import SENTINEL, Guarded, full_check, torch_cuda

pytestmark = pytest.mark.gpu


def _pair(torch, M, K, N, mode, plan, seed):
    """(noalloc result, ws result, phase counters of the noalloc call)."""
    A = orc.uniform(seed, (M, K))
    B = orc.uniform(seed + 1, (K, N))
    dA, dB = Guarded(M * K, A), Guarded(K * N, B)
    c_na, c_ws = Guarded(M * N, SENTINEL), Guarded(M * N, SENTINEL)
    need = mm.workspace_bytes(M, K, N, mode, plan)
    ws = torch.empty(max(need // 4, 1) + 64, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    mm.reset_phase_counters()
    st = mm.sgemm_noalloc(dA.ptr, dB.ptr, c_na.ptr, M, K, N, mode, plan)
    counters = mm.phase_counters()
    assert st == mm.OK, mm.last_error()
    assert mm.sgemm_ws(dA.ptr, dB.ptr, c_ws.ptr, M, K, N, mode, plan, ws.data_ptr(), need) == mm.OK
    torch.cuda.synchronize()
    assert c_na.guards_intact() and c_ws.guards_intact()
    a, b = c_na.numpy((M, N)), c_ws.numpy((M, N))
    assert full_check(A, B, a)["bound_ratio"] <= 1.0
    return a, b, counters


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_noalloc_matches_caller_workspace_two_calls(mode):
    torch = torch_cuda()
    # first call: split-K partials plus (256-wide plan) packed operands fill
    # most of the scratch; second call: a smaller problem on the same scratch
    for (M, K, N), plan in [((640, 2048, 900), (256, 4)), ((67, 8191, 513), (64, 16))]:
        a, b, counters = _pair(torch, M, K, N, mode, plan, 500 + M)
        assert np.array_equal(a, b), (mode, M, K, N, plan, int((a != b).sum()))
        assert counters["allocation"] == 0, counters


def test_noalloc_planner_and_limits():
    torch = torch_cuda()
    M, K, N = 64, 8192, 512
    A = orc.uniform(610, (M, K))
    B = orc.uniform(611, (K, N))
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL)
    torch.cuda.synchronize()
    # no plan: override / cache / rule (no selector), like sgemm_ws with that plan
    mm.reset_phase_counters()
    assert mm.sgemm_noalloc(dA.ptr, dB.ptr, dC.ptr, M, K, N, 1, None) == mm.OK, mm.last_error()
    assert mm.phase_counters()["allocation"] == 0
    torch.cuda.synchronize()
    assert full_check(A, B, dC.numpy((M, N)))["bound_ratio"] <= 1.0
    # a plan whose workspace exceeds the scratch: refused, C untouched
    M2 = N2 = K2 = 3000
    assert mm.workspace_bytes(M2, K2, N2, 1, (64, 16)) > mm.NOALLOC_SCRATCH_BYTES
    a2 = torch.zeros(M2 * K2, dtype=torch.float32, device="cuda")
    b2 = torch.zeros(K2 * N2, dtype=torch.float32, device="cuda")
    c2 = torch.full((M2 * N2,), SENTINEL, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert mm.sgemm_noalloc(a2.data_ptr(), b2.data_ptr(), c2.data_ptr(), M2, K2, N2, 1, (64, 16)) == mm.ERR_ALLOC
    assert "static scratch" in mm.last_error()
    torch.cuda.synchronize()
    assert bool(torch.all(c2 == SENTINEL))
    # host pointers are refused (staging would allocate)
    Ch = np.full((M, N), SENTINEL, np.float32)
    assert mm.sgemm_noalloc(A, B, Ch, M, K, N, 0, (64, 1)) == mm.ERR_UNSUPPORTED
    assert np.all(Ch == SENTINEL)


@pytest.mark.parametrize("mode", [1, 2])
def test_noalloc_whole_tile_split_call_needs_no_scratch(mode):
    """A split-mode 256-wide call on whole 256 x 256 tiles splits its operands
    on the fly, so it needs no workspace at all: through the static scratch it
    runs as through a caller workspace, bit for bit, packing nothing."""
    torch = torch_cuda()
    M, K, N = 1024, 2048, 1024
    assert mm.workspace_bytes(M, K, N, mode, (256, 1)) == 0
    a, b, counters = _pair(torch, M, K, N, mode, (256, 1), 700 + mode)
    assert np.array_equal(a, b), (mode, int((a != b).sum()))
    assert counters["allocation"] == 0 and counters["pack"] == 0 and counters["compute"] == 1, counters
