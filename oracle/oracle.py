"""ORACLE — test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module, and only as the checker; the product (maxmulsk_amd and
libmaxmulsk.so) never imports, links or calls it.

What it restates, and what pins each function (tests/test_oracle.py):

* ``gemm_f64`` / ``absgemm_f64`` — FP64 GEMM ground truth (the reference checks
  against cblas_dgemm, docs/BENCHMARKS.md:112-121). Pinned by: brute-force
  triple loop on tiny shapes, the C restatement (oracle.c) vs numpy's dgemm,
  identities A*I = A and I*B = B, exact small-integer products, and a planted
  transposition that the checks must catch.
* ``sgemm_portable`` — Utils::multiply_scalar (common/utils.hpp:50-61). Pinned
  by: exact results on small-integer inputs (every partial sum exact), its FP32
  error staying inside the K*u32*|A||B| bound against FP64, and equality with
  an FMA-chain emulation on inputs where FP64 is exact.
* ``bf16_rn`` / ``split`` / ``split_gemm_f64`` — BF16 rounding and the
  hi/lo(/mid) split (no reference counterpart). Pinned by: exhaustive round
  trip of all 65536 BF16 bit patterns, ties-to-even cases built by hand, the
  representation bounds |x - sum(terms)| <= 2^-16|x| (2 terms) and 2^-24|x|
  (3 terms) on every exponent, and exactness on inputs that are already BF16.
* ``v4c_choose`` / ``v3_select_from_table`` / ``nb_classify`` /
  ``nb_capacity`` / ``nb_counts`` — the reference's plan rules
  (sme/v4/sme-1x4-v4c-dispatch.cpp:17-22, sme/support/gemm_tuning.cpp:75-151,
  sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:410-466). Pinned row by row
  against the figures the reference's own result files print
  (tests/golden/ref_plan_rows.npz, extracted by gen_ref_plan_rows.py):
  a_packs, b_packs, mk_calls and both buffer sizes on all 7,862 rows of
  nb_sweep.csv + nb_sweep_ext.csv; the v4c pick and its buffers on the 35
  rows of v4c3.csv; the v3 pick on the 35 rows of final5.csv and the 66
  MaxMulSK rows of headline_run*.csv / tuned_run*.csv (ir_outer included);
  every field of nb_counts against a walk of the v4 loop nest itself
  (ncblock.cpp:483-595); and the bad-blocking cases of check_nb.cpp:69-79.
* ``uniform`` / ``log_uniform`` — the repo's own seeded generator
  (splitmix64 -> 24-bit or 53-bit fraction), pinned by committed golden values.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
U32 = 2.0 ** -24

_lib = None


def build() -> str:
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        P = C.c_void_p
        S = C.c_size_t
        sigs = {
            "oracle_gemm_f64": [P, P, P, S, S, S],
            "oracle_absgemm_f64": [P, P, P, S, S, S],
            "oracle_sgemm_portable": [P, P, P, S, S, S],
            "oracle_split": [P, S, C.c_int, P],
            "oracle_split_gemm_f64": [P, P, P, S, S, S, C.c_int],
            "oracle_v4c_choose": [S, S, S, P],
            "oracle_v3_select_from_table": [S, S, S, P],
            "oracle_nb_capacity": [S, S, S, S, S, S, P],
            "oracle_nb_counts": [S, S, S, S, S, S, P],
        }
        for name, args in sigs.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = None
        h.oracle_bf16_rn.argtypes = [C.c_float]
        h.oracle_bf16_rn.restype = C.c_float
        h.oracle_nb_classify.argtypes = [S] * 6
        h.oracle_nb_classify.restype = C.c_int
        _lib = h
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


# ---- seeded inputs -------------------------------------------------------------

_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _splitmix64(seed: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = np.uint64(seed) * _GOLD + np.arange(n, dtype=np.uint64)
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, shape: tuple[int, ...]) -> np.ndarray:
    """U[-1, 1) on the 2^-23 grid (exact in FP32): k * 2^-23 - 1, k = top 24 bits."""
    n = int(np.prod(shape))
    k = (_splitmix64(seed, n) >> np.uint64(40)).astype(np.float64)
    return (k * 2.0 ** -23 - 1.0).astype(np.float32).reshape(shape)


def log_uniform(seed: int, shape: tuple[int, ...], lo: float = -20.0, hi: float = 20.0) -> np.ndarray:
    """x = s * 2^e, e ~ U[lo, hi] continuous, s = +-1 fair (bit 0 of the draw)."""
    n = int(np.prod(shape))
    z = _splitmix64(seed, n)
    f = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    s = np.where((z & np.uint64(1)) == 1, -1.0, 1.0)
    return (s * np.exp2(lo + (hi - lo) * f)).astype(np.float32).reshape(shape)


def small_ints(seed: int, shape: tuple[int, ...], bound: int) -> np.ndarray:
    """Integers in [-bound, bound] as FP32 (exact)."""
    n = int(np.prod(shape))
    z = _splitmix64(seed, n) >> np.uint64(33)
    return ((z % np.uint64(2 * bound + 1)).astype(np.int64) - bound).astype(np.float32).reshape(shape)


# ---- GEMM ground truth -----------------------------------------------------------

def gemm_f64(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """C64 = A * B with FP32 inputs promoted to FP64 (numpy dgemm)."""
    return A.astype(np.float64) @ B.astype(np.float64)


def absgemm_f64(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    return np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))


def gemm_f64_c(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """The same product through the plain-C triple loop (oracle.c)."""
    M, K = A.shape
    N = B.shape[1]
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    Cm = np.empty((M, N), np.float64)
    lib().oracle_gemm_f64(_p(A), _p(B), _p(Cm), M, K, N)
    return Cm


def absgemm_f64_c(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    M, K = A.shape
    N = B.shape[1]
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    Cm = np.empty((M, N), np.float64)
    lib().oracle_absgemm_f64(_p(A), _p(B), _p(Cm), M, K, N)
    return Cm


def sgemm_portable(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """Utils::multiply_scalar restated (FP32 accumulator, ijk order)."""
    M, K = A.shape
    N = B.shape[1]
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    Cm = np.empty((M, N), np.float32)
    lib().oracle_sgemm_portable(_p(A), _p(B), _p(Cm), M, N, K)
    return Cm


# ---- BF16 split ----------------------------------------------------------------------

def bf16_rn(x: float) -> float:
    return lib().oracle_bf16_rn(float(x))


def split(x: np.ndarray, terms: int) -> np.ndarray:
    """terms x shape array of BF16-valued FP32: t0 = RN(x), t_i = RN(residual)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty((terms,) + x.shape, np.float32)
    lib().oracle_split(_p(x), x.size, terms, _p(out))
    return out


SPLIT_PRODUCTS = {1: [(1, 0), (0, 1), (0, 0)],
                  2: [(2, 0), (0, 2), (1, 1), (1, 0), (0, 1), (0, 0)]}


def split_gemm_f64(A: np.ndarray, B: np.ndarray, mode: int) -> np.ndarray:
    """Exact FP64 value of the mode's cross-product sum (mode 1: 3 products, 2: 6)."""
    terms = 2 if mode == 1 else 3
    At = split(A, terms).astype(np.float64)
    Bt = split(B, terms).astype(np.float64)
    out = np.zeros((A.shape[0], B.shape[1]), np.float64)
    for ta, tb in SPLIT_PRODUCTS[mode]:
        out += At[ta] @ Bt[tb]
    return out


def split_gemm_f64_c(A: np.ndarray, B: np.ndarray, mode: int) -> np.ndarray:
    terms = 2 if mode == 1 else 3
    At = split(A, terms)
    Bt = split(B, terms)
    M, K = A.shape
    N = B.shape[1]
    Cm = np.empty((M, N), np.float64)
    lib().oracle_split_gemm_f64(_p(At), _p(Bt), _p(Cm), M, K, N, mode)
    return Cm


# ---- error metrics ----------------------------------------------------------------

def errors(Cg: np.ndarray, C64: np.ndarray, absAB: np.ndarray, K: int) -> dict:
    """Error of an FP32 result against FP64.

    bound_ratio = max_ij |C - C64| / (K * u32 * (|A||B|)_ij); <= 1 is inside the
    FP32 bound |C - C64| <= K * u32 * |A||B| (gamma = 1).
    max_rel = max_ij |C - C64| / max(|C64|, u32 * (|A||B|)_ij).
    rel_frob = ||C - C64||_F / ||C64||_F (check_nb.cpp:13-17).
    """
    d = np.abs(Cg.astype(np.float64) - C64)
    scale = np.maximum(absAB, np.finfo(np.float64).tiny)
    bound_ratio = float(np.max(d / (K * U32 * scale))) if d.size else 0.0
    max_rel = float(np.max(d / np.maximum(np.abs(C64), U32 * scale))) if d.size else 0.0
    den = float(np.sqrt(np.sum(C64 * C64)))
    rel_frob = float(np.sqrt(np.sum(d * d)) / den) if den > 0 else 0.0
    return {"bound_ratio": bound_ratio, "max_rel": max_rel, "rel_frob": rel_frob,
            "max_abs": float(d.max()) if d.size else 0.0,
            "nonfinite": int(np.count_nonzero(~np.isfinite(Cg)))}


def sampled_rows_check(A: np.ndarray, B: np.ndarray, Cg: np.ndarray, rows: np.ndarray) -> dict:
    """errors() on a subset of rows, for shapes where the full FP64 product is slow."""
    Ar = A[rows]
    return errors(Cg[rows], gemm_f64(Ar, B), absgemm_f64(Ar, B), A.shape[1])


def sampled_cols_check(A: np.ndarray, B: np.ndarray, Cg: np.ndarray, cols: np.ndarray) -> dict:
    Bc = np.ascontiguousarray(B[:, cols])
    return errors(Cg[:, cols], gemm_f64(A, Bc), absgemm_f64(A, Bc), A.shape[1])


# ---- the reference's plan rules --------------------------------------------------------

def _sz(n: int):
    return (C.c_size_t * n)()


def v4c_choose(M: int, K: int, N: int) -> tuple[int, int, int]:
    o = _sz(3)
    lib().oracle_v4c_choose(M, K, N, o)
    return tuple(o)


def v3_select_from_table(M: int, K: int, N: int) -> tuple[int, int, int, int]:
    o = _sz(4)
    lib().oracle_v3_select_from_table(M, K, N, o)
    return tuple(o)


def nb_classify(M: int, K: int, N: int, Mc: int, Nc: int, Kc: int) -> int:
    return lib().oracle_nb_classify(M, K, N, Mc, Nc, Kc)


def nb_capacity(M: int, K: int, N: int, Mc: int, Nc: int, Kc: int) -> tuple[int, int]:
    o = _sz(2)
    lib().oracle_nb_capacity(M, K, N, Mc, Nc, Kc, o)
    return tuple(o)


def nb_counts(M: int, K: int, N: int, Mc: int, Nc: int, Kc: int) -> tuple[int, ...]:
    o = _sz(6)
    lib().oracle_nb_counts(M, K, N, Mc, Nc, Kc, o)
    return tuple(o)
