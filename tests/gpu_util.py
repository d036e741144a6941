"""Helpers shared by the GPU tests: device buffers (torch as plumbing), calls
through the C ABI, and checks against the oracle."""
from __future__ import annotations

import json
import os

import numpy as np

import maxmulsk_amd as mm
from oracle import gpu_oracle as go
from oracle import oracle as orc

SENTINEL = 3.75
GUARD = 4096  # floats of guard band on each side of every device buffer


def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    assert mm.device_count() > 0, "libmaxmulsk sees no gfx950 device: " + mm.last_error()
    return torch


class Guarded:
    """A device float buffer of n elements with poisoned guard bands around it."""

    def __init__(self, n: int, fill: float | np.ndarray):
        torch = torch_cuda()
        self.n = n
        self.buf = torch.full((n + 2 * GUARD,), -7.0e30, dtype=torch.float32, device="cuda")
        self.view = self.buf[GUARD:GUARD + n]
        if isinstance(fill, np.ndarray):
            self.view.copy_(torch.from_numpy(np.ascontiguousarray(fill, np.float32).reshape(-1)))
        else:
            self.view.fill_(fill)

    @property
    def ptr(self) -> int:
        return self.view.data_ptr()

    def guards_intact(self) -> bool:
        g = self.buf.cpu().numpy()
        return bool(np.all(g[:GUARD] == np.float32(-7.0e30)) and np.all(g[GUARD + self.n:] == np.float32(-7.0e30)))

    def numpy(self, shape) -> np.ndarray:
        return self.view.cpu().numpy().reshape(shape)


def gpu_gemm(A: np.ndarray, B: np.ndarray, mode: int = 0, plan=None, entry: str = "ex",
             check_guards: bool = True) -> np.ndarray:
    """C = A*B through the C ABI on device buffers; C prefilled with a sentinel."""
    torch = torch_cuda()
    M, K = A.shape
    N = B.shape[1]
    dA, dB = Guarded(M * K, A), Guarded(K * N, B)
    dC = Guarded(M * N, SENTINEL)
    torch.cuda.synchronize()
    if entry == "ex":
        st = mm.sgemm_ex(dA.ptr, dB.ptr, dC.ptr, M, K, N, mode=mode, plan=plan)
    elif entry == "v3":
        st = mm.sgemm_v3(dA.ptr, dB.ptr, dC.ptr, M, K, N)
    elif entry == "v4":
        st = mm.sgemm_v4(dA.ptr, dB.ptr, dC.ptr, M, K, N)
    else:
        raise ValueError(entry)
    assert st == mm.OK, f"status {st}: {mm.last_error()}"
    torch.cuda.synchronize()
    if check_guards:
        assert dA.guards_intact() and dB.guards_intact() and dC.guards_intact(), "guard band written"
    return dC.numpy((M, N))


def gpu_gemm_checked(A: np.ndarray, B: np.ndarray, mode: int = 0, plan=None, entry: str = "ex"):
    """gpu_gemm, then every element of C checked against FP64 A*B by the GPU
    oracle on the same device buffers. Returns (C, errors)."""
    torch = torch_cuda()
    M, K = A.shape
    N = B.shape[1]
    dA, dB = Guarded(M * K, A), Guarded(K * N, B)
    dC = Guarded(M * N, SENTINEL)
    torch.cuda.synchronize()
    if entry == "ex":
        st = mm.sgemm_ex(dA.ptr, dB.ptr, dC.ptr, M, K, N, mode=mode, plan=plan)
    elif entry == "v4":
        st = mm.sgemm_v4(dA.ptr, dB.ptr, dC.ptr, M, K, N)
    else:
        raise ValueError(entry)
    assert st == mm.OK, f"status {st}: {mm.last_error()}"
    torch.cuda.synchronize()
    assert dA.guards_intact() and dB.guards_intact() and dC.guards_intact(), "guard band written"
    e = go.check(dA.ptr, dB.ptr, dC.ptr, M, K, N)
    return dC.numpy((M, N)), e


def report(kind: str, **fields) -> None:
    """Appends one JSON line to $MAXMULSK_REPORT_DIR/<kind>.jsonl when that
    variable is set (accuracy tables collected from a test run)."""
    d = os.environ.get("MAXMULSK_REPORT_DIR")
    if not d:
        return
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, f"{kind}.jsonl"), "a") as f:
        f.write(json.dumps(fields) + "\n")


def full_check(A, B, Cg, mode: int = 0, K_override=None):
    K = A.shape[1] if K_override is None else K_override
    return orc.errors(Cg, orc.gemm_f64(A, B), orc.absgemm_f64(A, B), K)


def rows_check(A, B, Cg, nrows: int = 48, seed: int = 0):
    rng = np.random.default_rng(seed)
    M = A.shape[0]
    rows = np.unique(np.concatenate([[0, M - 1], rng.integers(0, M, nrows)]))
    return orc.sampled_rows_check(A, B, Cg, rows)


def cols_check(A, B, Cg, ncols: int = 48, seed: int = 1):
    rng = np.random.default_rng(seed)
    N = B.shape[1]
    cols = np.unique(np.concatenate([[0, N - 1], rng.integers(0, N, ncols)]))
    return orc.sampled_cols_check(A, B, Cg, cols)
