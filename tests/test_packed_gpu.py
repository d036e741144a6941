"""Pack-once operands (maxmulsk_pack / maxmulsk_sgemm_packed) through the C ABI:
the packed product must equal, bit for bit, the per-call path of
maxmulsk_sgemm_ex with a 256-wide plan on more than 128 rows (the same
fragments and kernel, packed on every call), and stay within the FP32 bound of
the FP64 oracle."""
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
