"""Pins the FP64 GPU oracle (oracle/gpu_oracle.hip) against the CPU oracle:
whole FP64 products on small ragged shapes, sampled rows and columns of large
ones, the reduced error figures against oracle.errors(), planted mistakes the
checker must catch, and bench.py's torch FP64 checker against both."""
import numpy as np
import pytest

from oracle import gpu_oracle as go
from oracle import oracle as orc
from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu

U53 = 2.0 ** -53


def _dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _gpu_f64(torch, dA, dB, M, K, N):
    c64 = torch.empty((M, N), dtype=torch.float64, device="cuda")
    ab = torch.empty((M, N), dtype=torch.float64, device="cuda")
    go.gemm_f64(dA.data_ptr(), dB.data_ptr(), c64.data_ptr(), ab.data_ptr(), M, K, N)
    return c64, ab


@pytest.mark.parametrize("mnk", [(1, 1, 1), (67, 91, 513), (130, 257, 1000), (256, 300, 64), (5, 1025, 4099)],
                         ids=lambda s: "x".join(map(str, s)))
def test_gpu_f64_product_matches_cpu_whole(mnk):
    torch = torch_cuda()
    M, N, K = mnk
    A = orc.uniform(301, (M, K))
    B = orc.log_uniform(302, (K, N), -10, 10)
    c64, ab = _gpu_f64(torch, _dev(torch, A), _dev(torch, B), M, K, N)
    C64, AB = orc.gemm_f64(A, B), orc.absgemm_f64(A, B)
    tol = 2 * K * U53 * AB
    assert np.all(np.abs(c64.cpu().numpy() - C64) <= tol)
    assert np.all(np.abs(ab.cpu().numpy() - AB) <= tol)


def test_gpu_f64_product_exact_on_integers():
    """Small integers: every FP64 partial sum is exact, so GPU and CPU agree
    bit for bit with the integer product."""
    torch = torch_cuda()
    M, N, K = 97, 131, 777
    A = orc.small_ints(303, (M, K), 100)
    B = orc.small_ints(304, (K, N), 100)
    c64, ab = _gpu_f64(torch, _dev(torch, A), _dev(torch, B), M, K, N)
    exact = A.astype(np.int64) @ B.astype(np.int64)
    assert np.array_equal(c64.cpu().numpy(), exact.astype(np.float64))
    assert np.array_equal(ab.cpu().numpy(), (np.abs(A.astype(np.int64)) @ np.abs(B.astype(np.int64))).astype(np.float64))


@pytest.mark.parametrize("mnk", [(4096, 4096, 4096), (64, 7168, 16384), (4096, 24576, 1536)],
                         ids=lambda s: "x".join(map(str, s)))
def test_gpu_f64_product_matches_cpu_sampled_rows_and_cols(mnk):
    torch = torch_cuda()
    M, N, K = mnk
    A = orc.uniform(305, (M, K))
    B = orc.uniform(306, (K, N))
    c64, ab = _gpu_f64(torch, _dev(torch, A), _dev(torch, B), M, K, N)
    rng = np.random.default_rng(7)
    rows = np.unique(np.concatenate([[0, M - 1], rng.integers(0, M, 16)]))
    cols = np.unique(np.concatenate([[0, N - 1], rng.integers(0, N, 16)]))
    g = c64[torch.from_numpy(rows).cuda()].cpu().numpy()
    ga = ab[torch.from_numpy(rows).cuda()].cpu().numpy()
    ref, refa = orc.gemm_f64(A[rows], B), orc.absgemm_f64(A[rows], B)
    assert np.all(np.abs(g - ref) <= 2 * K * U53 * refa)
    assert np.all(np.abs(ga - refa) <= 2 * K * U53 * refa)
    Bc = np.ascontiguousarray(B[:, cols])
    g = c64[:, torch.from_numpy(cols).cuda()].cpu().numpy()
    ref, refa = orc.gemm_f64(A, Bc), orc.absgemm_f64(A, Bc)
    assert np.all(np.abs(g - ref) <= 2 * K * U53 * refa)


def _close(a, b, rel=1e-9):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-300


def test_gpu_check_matches_cpu_error_figures():
    """The reduced figures equal oracle.errors() on the same candidate C (an
    FP32 rounding of the exact product with planted perturbations of known
    size), up to the two FP64 products' own rounding: relative 1e-5 on errors
    of 1e-7 |A||B| and more."""
    torch = torch_cuda()
    M, N, K = 150, 170, 600
    A = orc.uniform(307, (M, K))
    B = orc.log_uniform(308, (K, N), -6, 6)
    C64, AB = orc.gemm_f64(A, B), orc.absgemm_f64(A, B)
    rng = np.random.default_rng(9)
    Cg = (C64 + rng.standard_normal((M, N)) * 1e-7 * AB).astype(np.float32)
    Cg[3, 5] = np.float32(C64[3, 5] + 50 * K * orc.U32 * AB[3, 5])  # one element outside the bound
    dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, Cg)
    eg = go.check(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N)
    ec = orc.errors(Cg, C64, AB, K)
    for k in ("bound_ratio", "max_rel", "rel_frob", "max_abs"):
        assert _close(eg[k], ec[k], 1e-5), (k, eg[k], ec[k])
    assert eg["nonfinite"] == 0 and 40 < eg["bound_ratio"] < 60


def test_gpu_check_catches_planted_mistakes():
    torch = torch_cuda()
    M = N = K = 256
    A = orc.uniform(309, (M, K))
    B = orc.uniform(310, (K, N))
    dA, dB = _dev(torch, A), _dev(torch, B)
    good = (A.astype(np.float64) @ B.astype(np.float64)).astype(np.float32)
    e = go.check(dA.data_ptr(), dB.data_ptr(), _dev(torch, good).data_ptr(), M, K, N)
    assert e["bound_ratio"] <= 1.0 and e["nonfinite"] == 0
    # transposed operand
    bad = (A.astype(np.float64) @ B.T.astype(np.float64)).astype(np.float32)
    assert go.check(dA.data_ptr(), dB.data_ptr(), _dev(torch, bad).data_ptr(), M, K, N)["bound_ratio"] > 1
    # last k term dropped
    bad = (A[:, :-1].astype(np.float64) @ B[:-1].astype(np.float64)).astype(np.float32)
    assert go.check(dA.data_ptr(), dB.data_ptr(), _dev(torch, bad).data_ptr(), M, K, N)["bound_ratio"] > 1
    # sign flip of one element
    bad = good.copy()
    bad[10, 20] = -bad[10, 20]
    assert go.check(dA.data_ptr(), dB.data_ptr(), _dev(torch, bad).data_ptr(), M, K, N)["bound_ratio"] > 1
    # a NaN is counted and fails the bound
    bad = good.copy()
    bad[0, 0] = np.nan
    e = go.check(dA.data_ptr(), dB.data_ptr(), _dev(torch, bad).data_ptr(), M, K, N)
    assert e["nonfinite"] == 1 and not e["bound_ratio"] <= 1.0


def test_bench_torch_checker_matches_oracle():
    """bench.py reports accuracy with torch FP64 products on the GPU; its
    figures must equal the oracle's on the same candidate."""
    torch = torch_cuda()
    import bench

    M, N, K = 200, 333, 1111
    A = orc.uniform(311, (M, K))
    B = orc.log_uniform(312, (K, N), -8, 8)
    C64, AB = orc.gemm_f64(A, B), orc.absgemm_f64(A, B)
    rng = np.random.default_rng(11)
    Cg = (C64 + rng.standard_normal((M, N)) * 3e-7 * AB).astype(np.float32)
    dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, Cg)
    eb = bench.errors_torch(dA, dB, dC, K)
    ec = orc.errors(Cg, C64, AB, K)
    eg = go.check(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N)
    for k in ("bound_ratio", "max_rel", "rel_frob", "max_abs"):
        assert _close(eb[k], ec[k], 1e-5), (k, eb[k], ec[k])
        assert _close(eg[k], ec[k], 1e-5), (k, eg[k], ec[k])
