"""Python mirror of the MaxMulSK C ABI on MI355X (ctypes over libmaxmulsk.so).

The reference exposes two C entry points, ``maxmulsk_sgemm_v3`` and
``maxmulsk_sgemm_v4`` (capi/maxmulsk.h:32-37): row-major FP32 ``C = A*B`` with
dimensions in (M, K, N) order and C overwritten. ``sgemm_v3`` / ``sgemm_v4``
below take the same arguments (numpy arrays in host memory, or torch tensors /
raw device pointers on the GPU) and return the same status values. The
extensions (modes, plans, workspaces, grouped launches, autotuner, learned
selector) follow include/maxmulsk.h.

The library is loaded from this directory; importing this package never falls
back to any other implementation: when libmaxmulsk.so is missing, every call
raises ``MaxMulSKError``.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Any, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmaxmulsk.so")

OK = 0
ERR_UNSUPPORTED = 1
ERR_ALLOC = 2
ERR_DEVICE = 3
STATUS_NAMES = {OK: "MAXMULSK_OK", ERR_UNSUPPORTED: "MAXMULSK_ERR_UNSUPPORTED",
                ERR_ALLOC: "MAXMULSK_ERR_ALLOC", ERR_DEVICE: "MAXMULSK_ERR_DEVICE"}

MODE_PLAIN = 0
MODE_SPLIT3 = 1
MODE_SPLIT6 = 2
MODES = {"plain": MODE_PLAIN, "split3": MODE_SPLIT3, "split6": MODE_SPLIT6}

MLP_IN = 4
MLP_HIDDEN = 64


class MaxMulSKError(RuntimeError):
    pass


class Plan(C.Structure):
    _fields_ = [("tile", C.c_int), ("split_k", C.c_int)]

    def __repr__(self) -> str:
        return f"Plan(tile={self.tile}, split_k={self.split_k})"

    def as_tuple(self) -> tuple[int, int]:
        return (self.tile, self.split_k)


class Options(C.Structure):
    _fields_ = [("mode", C.c_int), ("plan", Plan), ("stream", C.c_void_p), ("async_", C.c_int),
                ("workspace", C.c_void_p), ("workspace_bytes", C.c_size_t)]


class Gemm(C.Structure):
    _fields_ = [("A", C.c_void_p), ("B", C.c_void_p), ("C", C.c_void_p),
                ("M", C.c_size_t), ("K", C.c_size_t), ("N", C.c_size_t)]


_SIZE = C.c_size_t
_P = C.c_void_p
_SIGS: dict[str, tuple[Any, list[Any]]] = {
    "maxmulsk_sgemm_v3": (C.c_int, [_P, _P, _P, _SIZE, _SIZE, _SIZE]),
    "maxmulsk_sgemm_v4": (C.c_int, [_P, _P, _P, _SIZE, _SIZE, _SIZE]),
    "maxmulsk_sgemm_ex": (C.c_int, [_P, _P, _P, _SIZE, _SIZE, _SIZE, C.POINTER(Options)]),
    "maxmulsk_workspace_bytes": (C.c_size_t, [_SIZE, _SIZE, _SIZE, C.c_int, C.POINTER(Plan)]),
    "maxmulsk_sgemm_ws": (C.c_int, [_P, _P, _P, _SIZE, _SIZE, _SIZE, C.c_int, C.POINTER(Plan), _P, _SIZE, _P]),
    "maxmulsk_sgemm_noalloc": (C.c_int, [_P, _P, _P, _SIZE, _SIZE, _SIZE, C.c_int, C.POINTER(Plan), _P]),
    "maxmulsk_probe_mfma": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "maxmulsk_probe_stream": (C.c_int, [_SIZE, _SIZE, _SIZE, C.POINTER(Plan), C.c_int, _P, C.POINTER(C.c_double)]),
    "maxmulsk_set_feed": (C.c_int, [C.c_int]),
    "maxmulsk_get_feed": (C.c_int, []),
    "maxmulsk_feed_variant": (C.c_int, [_SIZE, _SIZE, _SIZE, C.c_int, C.POINTER(Plan)]),
    "maxmulsk_pack": (C.c_int, [_P, _SIZE, _SIZE, C.c_int, C.c_int, _P, C.POINTER(_P)]),
    "maxmulsk_sgemm_packed": (C.c_int, [_P, _P, _P, C.POINTER(Options)]),
    "maxmulsk_packed_destroy": (None, [_P]),
    "maxmulsk_choose_plan": (C.c_int, [_SIZE, _SIZE, _SIZE, C.c_int, C.POINTER(Plan), C.POINTER(C.c_int)]),
    "maxmulsk_rule_plan": (Plan, [_SIZE, _SIZE, _SIZE, C.c_int]),
    "maxmulsk_sgemm_grouped": (C.c_int, [C.POINTER(Gemm), _SIZE, C.POINTER(Options)]),
    "maxmulsk_grouped_create": (C.c_int, [C.POINTER(Gemm), _SIZE, C.c_int, C.POINTER(Plan), _P, C.POINTER(_P)]),
    "maxmulsk_grouped_launch": (C.c_int, [_P, _P]),
    "maxmulsk_grouped_info": (C.c_int, [_P, C.POINTER(_SIZE), C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "maxmulsk_grouped_uses_feed": (C.c_int, [_P]),
    "maxmulsk_grouped_destroy": (None, [_P]),
    "maxmulsk_grouped_default_plan": (Plan, [_SIZE, _SIZE, _SIZE, C.c_int]),
    "maxmulsk_grouped_plans": (C.c_int, [C.POINTER(Gemm), _SIZE, C.c_int, C.POINTER(Plan)]),
    "maxmulsk_grouped_autotune": (C.c_int, [C.POINTER(Gemm), _SIZE, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(Plan), C.POINTER(C.c_float)]),
    "maxmulsk_num_candidates": (C.c_int, []),
    "maxmulsk_candidate": (Plan, [C.c_int]),
    "maxmulsk_autotune": (C.c_int, [_P, _P, _P, _SIZE, _SIZE, _SIZE, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_float), C.POINTER(Plan)]),
    "maxmulsk_cache_put": (C.c_int, [_SIZE, _SIZE, _SIZE, C.c_int, Plan]),
    "maxmulsk_cache_clear": (None, []),
    "maxmulsk_cache_size": (C.c_size_t, []),
    "maxmulsk_cache_save": (C.c_int, [C.c_char_p]),
    "maxmulsk_cache_load": (C.c_int, [C.c_char_p]),
    "maxmulsk_mlp_num_params": (C.c_size_t, []),
    "maxmulsk_mlp_forward": (C.c_int, [_P, _P, C.c_int, _P, _P, _P, _P]),
    "maxmulsk_mlp_backward": (C.c_int, [_P, _P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P]),
    "maxmulsk_mlp_adam": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, _P]),
    "maxmulsk_selector_set": (C.c_int, [_P]),
    "maxmulsk_selector_features": (None, [_SIZE, _SIZE, _SIZE, C.c_int, C.POINTER(C.c_float)]),
    "maxmulsk_selector_load": (C.c_int, [C.c_char_p]),
    "maxmulsk_selector_save": (C.c_int, [C.c_char_p, _P]),
    "maxmulsk_selector_active": (C.c_int, []),
    "maxmulsk_selector_predict": (C.c_int, [_SIZE, _SIZE, _SIZE, C.c_int, C.POINTER(C.c_float), C.POINTER(Plan)]),
    "maxmulsk_num_phase_counters": (C.c_int, []),
    "maxmulsk_phase_counter_name": (C.c_char_p, [C.c_int]),
    "maxmulsk_phase_counter": (C.c_ulonglong, [C.c_int]),
    "maxmulsk_phase_counters_reset": (None, []),
    "maxmulsk_device_count": (C.c_int, []),
    "maxmulsk_last_error": (C.c_char_p, []),
    "maxmulsk_lds_chunk_offset": (C.c_int, [C.c_int] * 7),
    "maxmulsk_lds_epilogue_write": (C.c_int, [C.c_int] * 4),
    "maxmulsk_lds_epilogue_read": (C.c_int, [C.c_int] * 3 + [C.POINTER(C.c_int)] * 2),
    "maxmulsk_num_tile_geometries": (C.c_int, []),
    "maxmulsk_tile_info": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int)]),
}
EXPORTED = tuple(_SIGS)

_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """The loaded libmaxmulsk.so; raises MaxMulSKError when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MaxMulSKError(f"{LIB_PATH} is not built (run __graft_entry__.build() "
                                "or python -m maxmulsk_amd.build)")
        # torch wheels bundle their own libamdhip64 (same soname, other file
        # name). Loading torch first makes libmaxmulsk bind to that runtime, so
        # the process holds one HIP runtime whichever of the two is used first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def phase_counters() -> dict[str, int]:
    """Host-runtime event counts since the last reset (maxmulsk_phase_counter)."""
    h = lib()
    return {h.maxmulsk_phase_counter_name(i).decode(): int(h.maxmulsk_phase_counter(i))
            for i in range(h.maxmulsk_num_phase_counters())}


def reset_phase_counters() -> None:
    lib().maxmulsk_phase_counters_reset()


def last_error() -> str:
    return lib().maxmulsk_last_error().decode()


def device_count() -> int:
    return lib().maxmulsk_device_count()


def _ptr(x: Any) -> int:
    """Address of a numpy array, torch tensor or raw integer pointer."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if x.dtype.__str__() not in ("torch.float32",) and not getattr(x, "_mmsk_any_dtype", False):
            raise MaxMulSKError(f"expected float32 tensor, got {x.dtype}")
        if not x.is_contiguous():
            raise MaxMulSKError("tensor must be contiguous")
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        import numpy as np
        if x.dtype != np.float32 or not x.flags["C_CONTIGUOUS"]:
            raise MaxMulSKError("array must be C-contiguous float32")
        return x.ctypes.data
    raise MaxMulSKError(f"cannot take the address of {type(x)}")


def _stream_ptr(stream: Any) -> int:
    if stream is None:
        return 0
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream  # torch.cuda.Stream


def sgemm_v3(A: Any, B: Any, Cm: Any, M: int, K: int, N: int) -> int:
    """maxmulsk_sgemm_v3: C = A*B, (M, K, N) order; returns the status."""
    return lib().maxmulsk_sgemm_v3(_ptr(A), _ptr(B), _ptr(Cm), M, K, N)


def sgemm_v4(A: Any, B: Any, Cm: Any, M: int, K: int, N: int) -> int:
    """maxmulsk_sgemm_v4: C = A*B, (M, K, N) order; returns the status."""
    return lib().maxmulsk_sgemm_v4(_ptr(A), _ptr(B), _ptr(Cm), M, K, N)


def sgemm_ex(A: Any, B: Any, Cm: Any, M: int, K: int, N: int, mode: int | str = MODE_PLAIN,
             plan: tuple[int, int] | None = None, stream: Any = None, async_: bool = False,
             workspace: Any = None, workspace_bytes: int = 0) -> int:
    if isinstance(mode, str):
        mode = MODES[mode]
    opt = Options(mode=mode, plan=Plan(*(plan or (0, 0))), stream=_stream_ptr(stream),
                  async_=1 if async_ else 0, workspace=_ptr(workspace) if workspace is not None else None,
                  workspace_bytes=workspace_bytes)
    return lib().maxmulsk_sgemm_ex(_ptr(A), _ptr(B), _ptr(Cm), M, K, N, C.byref(opt))


def workspace_bytes(M: int, K: int, N: int, mode: int = MODE_PLAIN,
                    plan: tuple[int, int] | None = None) -> int:
    p = C.byref(Plan(*plan)) if plan else None
    return lib().maxmulsk_workspace_bytes(M, K, N, mode, p)


def sgemm_ws(A: Any, B: Any, Cm: Any, M: int, K: int, N: int, mode: int, plan: tuple[int, int] | None,
             ws: Any, ws_bytes: int, stream: Any = None) -> int:
    p = C.byref(Plan(*plan)) if plan else None
    return lib().maxmulsk_sgemm_ws(_ptr(A), _ptr(B), _ptr(Cm), M, K, N, mode, p,
                                   _ptr(ws) if ws is not None else None, ws_bytes, _stream_ptr(stream))


def probe_mfma(mode: int, random_operands: bool, iterations: int = 20000) -> float:
    """Issued MFMA rate (TFLOP/s) of the register-only probe (maxmulsk_probe_mfma)."""
    out = C.c_double()
    check(lib().maxmulsk_probe_mfma(mode, int(random_operands), iterations, C.byref(out)))
    return out.value


def probe_stream(M: int, K: int, N: int, plan: tuple[int, int], calls: int = 400, stream: Any = None) -> dict:
    """maxmulsk_probe_stream: per-call us of an empty kernel on the plan's grid,
    a streaming kernel over each workgroup's operand bytes, and that kernel plus
    the split-K reduce, on `stream` (None: the legacy default stream)."""
    us = (C.c_double * 3)()
    check(lib().maxmulsk_probe_stream(M, K, N, C.byref(Plan(*plan)), calls, _stream_ptr(stream), us))
    return {"empty_us": us[0], "stream_us": us[1], "stream_reduce_us": us[2]}


FEED_DIRECT = 0     # MAXMULSK_FEED_DIRECT
FEED_DECOUPLED = 1  # MAXMULSK_FEED_DECOUPLED
FEED_AUTO = 2       # MAXMULSK_FEED_AUTO


def set_feed(feed: int) -> int:
    """maxmulsk_set_feed: operand feed of the 64 x 64 tile; returns the previous one."""
    prev = lib().maxmulsk_set_feed(int(feed))
    if prev < 0:
        raise ValueError(f"unknown feed {feed}")
    return prev


def get_feed() -> int:
    return lib().maxmulsk_get_feed()


def feed_variant(M: int, K: int, N: int, mode: int, plan: tuple[int, int]) -> int:
    """maxmulsk_feed_variant: 0 = direct tile, 1..8 = decoupled-feed variant, -1 = invalid plan."""
    return lib().maxmulsk_feed_variant(M, K, N, mode, C.byref(Plan(*plan)))


NOALLOC_SCRATCH_BYTES = 32 << 20  # MAXMULSK_NOALLOC_SCRATCH_BYTES


def sgemm_noalloc(A: Any, B: Any, Cm: Any, M: int, K: int, N: int, mode: int = MODE_PLAIN,
                  plan: tuple[int, int] | None = None, stream: Any = None) -> int:
    """maxmulsk_sgemm_noalloc: workspace from the library's static device scratch."""
    p = C.byref(Plan(*plan)) if plan else None
    return lib().maxmulsk_sgemm_noalloc(_ptr(A), _ptr(B), _ptr(Cm), M, K, N, mode, p, _stream_ptr(stream))


def choose_plan(M: int, K: int, N: int, mode: int = MODE_PLAIN) -> tuple[tuple[int, int], int]:
    out = Plan()
    src = C.c_int(-1)
    st = lib().maxmulsk_choose_plan(M, K, N, mode, C.byref(out), C.byref(src))
    if st != OK:
        raise MaxMulSKError(last_error())
    return out.as_tuple(), src.value


def rule_plan(M: int, K: int, N: int, mode: int = MODE_PLAIN) -> tuple[int, int]:
    return lib().maxmulsk_rule_plan(M, K, N, mode).as_tuple()


def candidates() -> list[tuple[int, int]]:
    n = lib().maxmulsk_num_candidates()
    return [lib().maxmulsk_candidate(i).as_tuple() for i in range(n)]


def lds_chunk_offset(mode: int, op: int, rows: int, bk: int, row: int, c: int, term: int) -> int:
    return lib().maxmulsk_lds_chunk_offset(mode, op, rows, bk, row, c, term)


def lds_epilogue_write(wtn: int, lane: int, tj: int, q: int) -> int:
    return lib().maxmulsk_lds_epilogue_write(wtn, lane, tj, q)


def lds_epilogue_read(wtn: int, lane: int, i: int) -> tuple:
    """(float offset, row, first column) of lane's read in step i; offset -1
    when out of range."""
    row, col = C.c_int(-1), C.c_int(-1)
    off = lib().maxmulsk_lds_epilogue_read(wtn, lane, i, C.byref(row), C.byref(col))
    return off, row.value, col.value


def tile_info(cfg: int, mode: int) -> dict:
    out = (C.c_int * 6)()
    if lib().maxmulsk_tile_info(cfg, mode, out) != 0:
        raise MaxMulSKError(f"no tile geometry {cfg} for mode {mode}")
    return dict(zip(("bm", "bn", "bk", "threads", "b_cols", "waves"), list(out)))


def tile_infos() -> list[dict]:
    n = lib().maxmulsk_num_tile_geometries()
    return [dict(tile_info(c, m), cfg=c, mode=m) for c in range(n) for m in (0, 1, 2)]


def check(status: int) -> None:
    if status != OK:
        raise MaxMulSKError(f"{STATUS_NAMES.get(status, status)}: {last_error()}")


def make_gemms(problems: Sequence[tuple[Any, Any, Any, int, int, int]]) -> Any:
    arr = (Gemm * len(problems))()
    for i, (a, b, c, m, k, n) in enumerate(problems):
        arr[i] = Gemm(_ptr(a), _ptr(b), _ptr(c), m, k, n)
    return arr


class GroupedPlan:
    """A prepared grouped launch (maxmulsk_grouped_create); launch() replays its graph."""

    def __init__(self, problems: Sequence[tuple[Any, Any, Any, int, int, int]], mode: int = MODE_PLAIN,
                 plans: Sequence[tuple[int, int]] | None = None):
        self._gemms = make_gemms(problems)
        self.n = len(problems)
        pl = None
        if plans is not None:
            arr = (Plan * self.n)(*[Plan(*p) for p in plans])
            pl = arr
        h = C.c_void_p()
        check(lib().maxmulsk_grouped_create(self._gemms, self.n, mode, pl, None, C.byref(h)))
        self.handle = h

    def launch(self, stream: Any = None) -> None:
        check(lib().maxmulsk_grouped_launch(self.handle, _stream_ptr(stream)))

    def info(self) -> dict:
        items = C.c_size_t()
        flops = C.c_double()
        graph = C.c_int()
        check(lib().maxmulsk_grouped_info(self.handle, C.byref(items), C.byref(flops), C.byref(graph)))
        return {"work_items": items.value, "flops": flops.value, "graph": bool(graph.value),
                "feed": bool(lib().maxmulsk_grouped_uses_feed(self.handle))}

    def close(self) -> None:
        if self.handle:
            lib().maxmulsk_grouped_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Packed:
    """An operand packed once (maxmulsk_pack): operand 0 = A (rows x K,
    row-major), 1 = B (K x rows, row-major). Multiply with sgemm_packed."""

    def __init__(self, X: Any, rows: int, K: int, operand: int, mode: int | str = MODE_PLAIN,
                 stream: Any = None):
        if isinstance(mode, str):
            mode = MODES[mode]
        h = C.c_void_p()
        check(lib().maxmulsk_pack(_ptr(X), rows, K, operand, mode, _stream_ptr(stream), C.byref(h)))
        self.handle = h
        self.rows, self.K, self.operand, self.mode = rows, K, operand, mode

    def close(self) -> None:
        if self.handle:
            lib().maxmulsk_packed_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sgemm_packed(A: Packed, B: Packed, Cm: Any, split_k: int = 1, stream: Any = None,
                 async_: bool = False, workspace: Any = None, workspace_bytes: int = 0) -> int:
    """C (A.rows x B.rows) = A*B from packed operands; returns the status."""
    opt = Options(mode=A.mode, plan=Plan(0, split_k), stream=_stream_ptr(stream), async_=1 if async_ else 0,
                  workspace=_ptr(workspace) if workspace is not None else None, workspace_bytes=workspace_bytes)
    return lib().maxmulsk_sgemm_packed(A.handle, B.handle, _ptr(Cm), C.byref(opt))


def sgemm_grouped(problems: Sequence[tuple[Any, Any, Any, int, int, int]], mode: int = MODE_PLAIN,
                  stream: Any = None) -> int:
    gemms = make_gemms(problems)
    opt = Options(mode=mode, plan=Plan(0, 0), stream=_stream_ptr(stream), async_=0, workspace=None,
                  workspace_bytes=0)
    return lib().maxmulsk_sgemm_grouped(gemms, len(problems), C.byref(opt))


def grouped_default_plan(M: int, K: int, N: int, mode: int = MODE_PLAIN) -> tuple[int, int]:
    return lib().maxmulsk_grouped_default_plan(M, K, N, mode).as_tuple()


def grouped_plans(problems, mode: int = MODE_PLAIN) -> list[tuple[int, int]]:
    """Plans a grouped launch of these problems takes by default, one per problem."""
    gemms = make_gemms(problems)
    out = (Plan * max(len(problems), 1))()
    check(lib().maxmulsk_grouped_plans(gemms, len(problems), mode, out))
    return [out[i].as_tuple() for i in range(len(problems))]


def grouped_autotune(problems, mode: int = MODE_PLAIN, max_classes: int = 0, reps: int = 3):
    gemms = make_gemms(problems)
    out = (Plan * len(problems))()
    best = C.c_float()
    check(lib().maxmulsk_grouped_autotune(gemms, len(problems), mode, max_classes, reps, out, C.byref(best)))
    return [p.as_tuple() for p in out], best.value


def selector_features(M: int, K: int, N: int, mode: int = MODE_PLAIN) -> list[float]:
    out = (C.c_float * 4)()
    lib().maxmulsk_selector_features(M, K, N, mode, out)
    return list(out)


def selector_predict(M: int, K: int, N: int, mode: int = MODE_PLAIN) -> tuple[list[float], tuple[int, int]]:
    n = lib().maxmulsk_num_candidates()
    pred = (C.c_float * n)()
    best = Plan()
    check(lib().maxmulsk_selector_predict(M, K, N, mode, pred, C.byref(best)))
    return list(pred), best.as_tuple()


def autotune(A: Any, B: Any, Cm: Any, M: int, K: int, N: int, mode: int = MODE_PLAIN,
             stage1_reps: int = 0, stage2_reps: int = 0) -> tuple[list[float], tuple[int, int]]:
    n = lib().maxmulsk_num_candidates()
    times = (C.c_float * n)()
    best = Plan()
    check(lib().maxmulsk_autotune(_ptr(A), _ptr(B), _ptr(Cm), M, K, N, mode, stage1_reps, stage2_reps,
                                  times, C.byref(best)))
    return list(times), best.as_tuple()
