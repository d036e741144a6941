"""ORACLE — test infrastructure only: the FP64 GPU ground truth.

Only tests/ import this module, as the checker. It binds liboracle_gpu.so
(gpu_oracle.hip): an FP64 GEMM on the device that widens every FP32 input
exactly and accumulates with FP64 FMAs, and the error figures of
oracle.errors() reduced over every element of a candidate C on the device.
Pinned against the CPU oracle (oracle.gemm_f64 / absgemm_f64 / errors) in
tests/test_gpu_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_gpu.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        h = C.CDLL(LIB_PATH)
        P, S = C.c_void_p, C.c_size_t
        h.oracle_gpu_check.argtypes = [P, P, P, S, S, S, P]
        h.oracle_gpu_check.restype = C.c_int
        h.oracle_gpu_gemm_f64.argtypes = [P, P, P, P, S, S, S]
        h.oracle_gpu_gemm_f64.restype = C.c_int
        _lib = h
    return _lib


def check(dA: int, dB: int, dC: int, M: int, K: int, N: int) -> dict:
    """Errors of the device FP32 C (M x N) against FP64 A*B, every element.
    Same keys as oracle.errors()."""
    out = (C.c_double * 5)()
    rc = lib().oracle_gpu_check(dA, dB, dC, M, K, N, out)
    if rc != 0:
        raise RuntimeError(f"oracle_gpu_check failed ({rc})")
    return {"bound_ratio": out[0], "max_rel": out[1], "rel_frob": out[2], "max_abs": out[3],
            "nonfinite": int(out[4])}


def gemm_f64(dA: int, dB: int, dC64: int, dABS: int, M: int, K: int, N: int) -> None:
    """C64 = A*B and ABS = |A|*|B| into device FP64 buffers (0 = skip)."""
    rc = lib().oracle_gpu_gemm_f64(dA, dB, dC64 or None, dABS or None, M, K, N)
    if rc != 0:
        raise RuntimeError(f"oracle_gpu_gemm_f64 failed ({rc})")
