"""Shape sweep with the reference's benchmark method, on one GPU.

    python -m maxmulsk_amd.sweep [--shapes M,N,K;...] [--runs 5] [--out sweep.csv]

Method of bench/results/2026-09-13/source/bench_final5.cpp:72-151:
  - one common warm-up before any shape;
  - per shape: seeded inputs, a correctness check of every entrant (relative
    Frobenius error and the FP32-bound ratio against an FP64 product, C
    prefilled with a sentinel), then two warm-up calls per entrant;
  - adaptive repetitions: reps = clamp(0.3 s / slowest single call, 3, 2000);
  - `runs` rounds, entrant order rotated per shape and round (Latin square);
    a round's sample is the median of its reps;
  - reported: median over rounds (ms, TFLOP/s = 2MNK / t) and spread % =
    (max - min) / median over rounds, plus the speed ratios.
Entrants: the three modes of this library through maxmulsk_sgemm_ex on device
arrays (plans from the autotuner cache, filled before timing), and the vendor
FP32 GEMM (torch.matmul with TF32 off) for context. Times are per call with
HIP events on the launch stream, inputs resident in device memory.
"""
from __future__ import annotations

import argparse
import csv
import ctypes
import sys
import time

import numpy as np

import maxmulsk_amd as mm
from maxmulsk_amd import inputs

# The measured workload shapes, M x N x K.
DEFAULT_SHAPES = [
    (256, 256, 256), (1024, 1024, 1024), (2048, 2048, 2048), (4096, 4096, 4096),
    (64, 512, 8192), (128, 512, 32768), (64, 24576, 1536), (64, 32768, 512),
    (1023, 1025, 999), (67, 513, 8191), (64, 7168, 16384), (4096, 24576, 1536),
]
ENTRANTS = ("plain", "split3", "split6", "vendor")
U32 = 2.0 ** -24


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--shapes", default="", help="M,N,K;M,N,K;... (default: the workload shapes)")
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--sample-seconds", type=float, default=0.3)
    ap.add_argument("--out", default="", help="CSV path (default: stdout)")
    args = ap.parse_args(argv)
    import torch

    shapes = ([tuple(int(x) for x in s.split(",")) for s in args.shapes.split(";") if s]
              if args.shapes else DEFAULT_SHAPES)
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    torch.backends.cuda.matmul.allow_tf32 = False

    def log(msg):
        print(f"[sweep] {msg}", file=sys.stderr, flush=True)

    def make_call(name, dA, dB, dC, M, K, N):
        if name == "vendor":
            return lambda: torch.matmul(dA, dB, out=dC)
        # arguments built once, as a C caller holds them
        opt = mm.Options(mode=mm.MODES[name], plan=mm.Plan(0, 0), stream=sptr, async_=1,
                         workspace=None, workspace_bytes=0)
        fn = mm.lib().maxmulsk_sgemm_ex
        args_ = (dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, ctypes.byref(opt))

        def call():
            st = fn(*args_)
            if st != mm.OK:
                raise RuntimeError(f"{name} status {st}: {mm.last_error()}")
        call.opt = opt
        return call

    def time_calls(fn, reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        ts = []
        for _ in range(reps):
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        return ts

    # common warm-up: clocks and first-use costs are paid before any shape
    w = torch.rand(512, 512, device=dev)
    wc = torch.empty_like(w)
    for _ in range(10):
        for name in ENTRANTS:
            make_call(name, w, w, wc, 512, 512, 512)()
    torch.cuda.synchronize()

    cols = (["tag", "M", "N", "K"] + [f"plan_{e}" for e in ENTRANTS[:3]] + ["reps"] +
            [f"{e}_{q}" for e in ENTRANTS for q in ("ms", "tf", "sp")] +
            ["split3_vs_plain", "split6_vs_plain", "split3_vs_vendor", "plain_vs_vendor"] +
            [f"{q}_{e}" for e in ENTRANTS for q in ("relfrob", "bound")])
    rows = []
    for si, (M, N, K) in enumerate(shapes):
        t0 = time.time()
        A = inputs.uniform(1 + si, (M, K))
        B = inputs.uniform(101 + si, (K, N))
        dA, dB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
        dC = torch.empty((M, N), dtype=torch.float32, device=dev)
        row = {"tag": f"s{si}", "M": M, "N": N, "K": K}
        for name in ENTRANTS[:3]:
            _, best = mm.autotune(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mm.MODES[name], 3, 6)
            row[f"plan_{name}"] = f"{best[0]}x{best[1]}"
        calls = {e: make_call(e, dA, dB, dC, M, K, N) for e in ENTRANTS}
        # correctness once per entrant, C prefilled with a sentinel
        A64, B64 = dA.double(), dB.double()
        C64 = A64 @ B64
        AB = A64.abs() @ B64.abs()
        for e in ENTRANTS:
            dC.fill_(3.75)
            calls[e]()
            torch.cuda.synchronize()
            d = (dC.double() - C64).abs()
            row[f"relfrob_{e}"] = float(d.norm() / C64.norm()) if float(C64.norm()) > 0 else 0.0
            row[f"bound_{e}"] = float((d / (K * U32 * AB.clamp_min(1e-300))).max())
        del A64, B64, C64, AB
        for e in ENTRANTS:
            for _ in range(2):
                calls[e]()
        torch.cuda.synchronize()
        slowest = max(min(time_calls(calls[e], 3)) for e in ENTRANTS) * 1e-3
        reps = int(min(2000, max(3, args.sample_seconds / max(slowest, 1e-9))))
        row["reps"] = reps
        samples = {e: [] for e in ENTRANTS}
        for r in range(args.runs):
            for slot in range(len(ENTRANTS)):
                e = ENTRANTS[(si + r + slot) % len(ENTRANTS)]  # Latin square over shapes and runs
                samples[e].append(float(np.median(time_calls(calls[e], reps))))
        flops = 2.0 * M * N * K
        for e in ENTRANTS:
            ms = float(np.median(samples[e]))
            row[f"{e}_ms"] = ms
            row[f"{e}_tf"] = flops / (ms * 1e-3) / 1e12
            row[f"{e}_sp"] = 100.0 * (max(samples[e]) - min(samples[e])) / ms
        row["split3_vs_plain"] = row["split3_tf"] / row["plain_tf"]
        row["split6_vs_plain"] = row["split6_tf"] / row["plain_tf"]
        row["split3_vs_vendor"] = row["split3_tf"] / row["vendor_tf"]
        row["plain_vs_vendor"] = row["plain_tf"] / row["vendor_tf"]
        rows.append(row)
        log(f"{M}x{N}x{K}: plain {row['plain_tf']:.1f} split3 {row['split3_tf']:.1f} split6 {row['split6_tf']:.1f} "
            f"vendor {row['vendor_tf']:.1f} TFLOP/s ({time.time() - t0:.1f} s)")
        del dA, dB, dC
        torch.cuda.empty_cache()
    mm.lib().maxmulsk_cache_clear()

    out = open(args.out, "w", newline="") if args.out else sys.stdout
    wr = csv.DictWriter(out, fieldnames=cols)
    wr.writeheader()
    for row in rows:
        wr.writerow({k: (f"{v:.4g}" if isinstance(v, float) else v) for k, v in row.items()})
    if args.out:
        out.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
