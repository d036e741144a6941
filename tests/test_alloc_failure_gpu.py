"""Allocation failures (capi/maxmulsk.cpp:22-29 maps AllocationFailed to
MAXMULSK_ERR_ALLOC; sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:645-654): with
MAXMULSK_TEST_ALLOC_LIMIT set, every device allocation of the library above
the limit fails as hipMalloc would. Each entry point must then return
MAXMULSK_ERR_ALLOC and leave C untouched, and work again once memory is
available. Runs in a fresh process, so no pool has grown beforehand."""
import os
import subprocess
import sys

import pytest

from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import os, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import torch
import maxmulsk_amd as mm
from oracle import oracle as orc
S = 3.75
M, K, N = 64, 8192, 512                      # plan (64, 16): 16 x 64 x 512 partials = 2 MiB
A = orc.uniform(1, (M, K)); B = orc.uniform(2, (K, N))
dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
dC = torch.full((M, N), S, dtype=torch.float32, device="cuda")
torch.cuda.synchronize()
def untouched():
    torch.cuda.synchronize()
    return bool(torch.all(dC == S))
os.environ["MAXMULSK_TEST_ALLOC_LIMIT"] = str(1 << 20)
r = {}
r["ex_split"] = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=1, plan=(64, 16))
r["ex_split_untouched"] = untouched()
Ch = np.full((M, N), S, np.float32)
r["v4_host"] = mm.sgemm_v4(A, B, Ch, M, K, N)      # staging A, B, C: 2 + 16 + 0.1 MiB
r["v3_host"] = mm.sgemm_v3(A, B, Ch, M, K, N)
r["host_untouched"] = bool(np.all(Ch == S))
M2 = N2 = K2 = 600
A2 = torch.from_numpy(orc.uniform(3, (M2, K2))).cuda(); B2 = torch.from_numpy(orc.uniform(4, (K2, N2))).cuda()
C2 = torch.full((M2, N2), S, dtype=torch.float32, device="cuda")
r["packed_path"] = mm.sgemm_ex(A2.data_ptr(), B2.data_ptr(), C2.data_ptr(), M2, K2, N2, mode=1, plan=(256, 1))
torch.cuda.synchronize()
r["packed_untouched"] = bool(torch.all(C2 == S))
try:
    mm.Packed(A2.data_ptr(), M2, K2, 0, 1)
    r["pack"] = 0
except mm.MaxMulSKError as e:
    r["pack"] = str(e)
try:
    gp = mm.GroupedPlan([(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N)], mode=1, plans=[(64, 16)])
    r["grouped"] = 0
    gp.close()
except mm.MaxMulSKError as e:
    r["grouped"] = str(e)
r["grouped_untouched"] = untouched()
# no split-K and device pointers: nothing to allocate, the call succeeds
r["ex_nosplit"] = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=1, plan=(64, 1))
torch.cuda.synchronize()
del os.environ["MAXMULSK_TEST_ALLOC_LIMIT"]
dC.fill_(S)
r["ex_split_after"] = mm.sgemm_ex(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode=1, plan=(64, 16))
torch.cuda.synchronize()
C64, AB = orc.gemm_f64(A, B), orc.absgemm_f64(A, B)
r["after_ratio"] = orc.errors(dC.cpu().numpy(), C64, AB, K)["bound_ratio"]
print("RESULT", repr(r), flush=True)
"""


def test_allocation_failure_returns_err_alloc_and_leaves_c_untouched():
    torch_cuda()
    import maxmulsk_amd as mm
    out = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-3000:]
    line = [ln for ln in out.stdout.splitlines() if ln.startswith("RESULT ")][-1]
    r = eval(line[len("RESULT "):], {"__builtins__": {}}, {"True": True, "False": False})
    assert r["ex_split"] == mm.ERR_ALLOC and r["ex_split_untouched"], r
    assert r["v4_host"] == mm.ERR_ALLOC and r["v3_host"] == mm.ERR_ALLOC and r["host_untouched"], r
    assert r["packed_path"] == mm.ERR_ALLOC and r["packed_untouched"], r
    assert "MAXMULSK_ERR_ALLOC" in str(r["pack"]) and "MAXMULSK_ERR_ALLOC" in str(r["grouped"]), r
    assert r["grouped_untouched"], r
    assert r["ex_nosplit"] == mm.OK, r
    assert r["ex_split_after"] == mm.OK and r["after_ratio"] <= 1.0, r
