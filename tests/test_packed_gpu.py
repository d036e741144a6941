"""Pack-once operands (maxmulsk_pack / maxmulsk_sgemm_packed) through the C ABI:
the packed product must equal, bit for bit, the per-call path of
maxmulsk_sgemm_ex with a 256-wide plan on more than 128 rows (the same
fragments and kernel, packed on every call; in the split modes on whole tiles
and k-blocks, the direct 256 x 256 tile that splits on the fly), and stay
within the FP32 bound of the FP64 oracle."""
import numpy as np
import pytest

import maxmulsk_amd as mm
from oracle import oracle as orc
from tests.gpu_util import SENTINEL, Guarded, full_check, gpu_gemm, torch_cuda

pytestmark = pytest.mark.gpu


def packed_gemm(A, B, mode, split_k=1, reuse=None):
    torch_cuda()
    import torch
    M, K = A.shape
    N = B.shape[1]
    dA, dB = Guarded(M * K, A), Guarded(K * N, B)
    dC = Guarded(M * N, SENTINEL)
    torch.cuda.synchronize()
    pa = mm.Packed(dA.ptr, M, K, 0, mode)
    pb = reuse if reuse is not None else mm.Packed(dB.ptr, N, K, 1, mode)
    mm.check(mm.sgemm_packed(pa, pb, dC.ptr, split_k=split_k))
    assert dA.guards_intact() and dB.guards_intact() and dC.guards_intact()
    out = dC.numpy((M, N))
    pa.close()
    return out, pb


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("mnk", [(1024, 1024, 1024), (1023, 1025, 999), (300, 257, 4097)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("split", [1, 4])
def test_packed_equals_per_call_path(mode, mnk, split):
    M, N, K = mnk
    A = orc.uniform(51, (M, K))
    B = orc.log_uniform(52, (K, N))
    Cp, pb = packed_gemm(A, B, mode, split)
    pb.close()
    Ce = gpu_gemm(A, B, mode=mode, plan=(256, split))
    assert np.array_equal(Cp, Ce)
    e = full_check(A, B, Cp)
    assert e["nonfinite"] == 0 and e["bound_ratio"] <= 1.0, e


def test_packed_B_reused_over_several_A():
    """One packed B (weights, freed after packing) against three A of different
    M; small integers make every product exact."""
    K, N = 777, 515
    B = orc.small_ints(70, (K, N), 8)
    pb = None
    for i, M in enumerate((256, 129, 600)):
        A = orc.small_ints(61 + i, (M, K), 8)
        Cp, pb = packed_gemm(A, B, 1, 1, reuse=pb)
        exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
        assert np.array_equal(Cp, exact), M
    pb.close()


def test_packed_rejects_mismatch():
    torch_cuda()
    import torch
    x = torch.zeros(256 * 64, device="cuda")
    y = torch.zeros(256 * 96, device="cuda")
    c = torch.zeros(256 * 256, device="cuda")
    pa = mm.Packed(x.data_ptr(), 256, 64, 0, 1)
    pb = mm.Packed(y.data_ptr(), 256, 96, 1, 1)   # K differs
    pb2 = mm.Packed(x.data_ptr(), 256, 64, 1, 2)  # mode differs
    assert mm.sgemm_packed(pa, pb, c.data_ptr()) == mm.ERR_UNSUPPORTED
    assert mm.sgemm_packed(pa, pb2, c.data_ptr()) == mm.ERR_UNSUPPORTED
    assert mm.sgemm_packed(pb, pa, c.data_ptr()) == mm.ERR_UNSUPPORTED  # operands swapped
    assert mm.sgemm_packed(pa, pa, c.data_ptr(), split_k=17) == mm.ERR_UNSUPPORTED
    h = mm.C.c_void_p()
    host = np.zeros(256 * 64, np.float32)
    assert mm.lib().maxmulsk_pack(host.ctypes.data, 256, 64, 0, 1, None, mm.C.byref(h)) == mm.ERR_UNSUPPORTED
    for p in (pa, pb, pb2):
        p.close()


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("mnk,split", [((512, 768, 1024), 1), ((768, 512, 4096), 4), ((256, 256, 96), 2),
                                       ((1024, 1024, 1024), 32), ((2048, 1024, 512), 1)],
                         ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else str(v))
def test_whole_tile_split_calls_skip_the_pack(mode, mnk, split):
    """Split-mode 256-wide plans on whole tiles and k-blocks run the direct
    256 x 256 tile, which splits the operands on the fly: no pack launch, one
    GEMM (plus the reduce with split-K), and C bit-identical to the packed
    kernel on pre-packed operands, within each mode's bound: the FP32 bound
    plus split3's K-independent 3 * 2^-16 * |a||b| per product (a ratio of
    1 + 768 / K) and split6's 2^-24-scale terms (1 + 4 / K)."""
    torch = torch_cuda()
    M, N, K = mnk
    A = orc.uniform(81, (M, K))
    B = orc.log_uniform(82, (K, N))
    dA, dB, dC = Guarded(M * K, A), Guarded(K * N, B), Guarded(M * N, SENTINEL)
    torch.cuda.synchronize()
    print(f"direct 256 x 256: mode {mode} {M}x{N}x{K} split {split}", flush=True)
    assert mm.sgemm_ex(dA.ptr, dB.ptr, dC.ptr, M, K, N, mode=mode, plan=(256, split)) == mm.OK  # warms the pool
    mm.reset_phase_counters()
    assert mm.sgemm_ex(dA.ptr, dB.ptr, dC.ptr, M, K, N, mode=mode, plan=(256, split)) == mm.OK
    torch.cuda.synchronize()
    c = mm.phase_counters()
    assert (c["pack"], c["compute"], c["reduce"]) == (0, 1, int(split > 1)), c
    assert dA.guards_intact() and dB.guards_intact() and dC.guards_intact()
    Cd = dC.numpy((M, N))
    Cp, pb = packed_gemm(A, B, mode, split)
    pb.close()
    assert np.array_equal(Cd, Cp)
    e = full_check(A, B, Cd)
    limit = {1: 1.0 + 768.0 / K, 2: 1.0 + 4.0 / K}[mode]
    assert e["nonfinite"] == 0 and e["bound_ratio"] <= limit, e
