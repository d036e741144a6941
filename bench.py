"""Benchmark: TFLOP/s at FP32 accuracy, split-BF16 vs plain FP32 MFMA, on one MI355X.

A step is one C = A*B call through the C ABI (maxmulsk_sgemm_ex, device
pointers, inputs resident in HBM, the library's pooled workspace) at
M = N = K = 4096 in the split-BF16 3-product mode. The same run also times the
plain FP32-MFMA path and split6 at 4096^3, both paths on the narrow-M shapes
(M x N x K), and the 256-problem grouped batch with its plans picked by the
grouped autotuner, and reports each result's error against an FP64 product.

Timing follows the reference's benchmark method (bench_final5.cpp:72-151): a
warm-up, then rounds in rotated (Latin-square) order over the entrants of one
shape, each round timing one entrant over `steps` back-to-back calls between
two HIP events on the launch stream; a figure is the median over the rounds
and its spread is (max - min) / median. The reference treats differences
below about 5% as noise (docs/BENCHMARKS.md:2262-2280); the line carries that
floor next to every spread.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TFLOP/s at FP32 accuracy: split-BF16 vs plain FP32 MFMA (4096³, M=64 shapes)"
U32 = 2.0 ** -24
BF16_PEAK = 2516.6   # TFLOP/s, MI355X dense BF16 MFMA
FP32_PEAK = 157.3    # TFLOP/s, MI355X FP32 MFMA
HBM_PEAK = 8000.0    # GB/s
NARROW = [(64, 512, 8192), (64, 24576, 1536), (64, 32768, 512), (128, 512, 32768), (64, 7168, 16384)]
# Operand feed of the 64 x 64 and 128 x 128 tiles, decoupled (producer and
# consumer waves) against direct, on plans of one workgroup per CU: (M, N, K,
# mode, plan). The last three are plans the automatic feed leaves direct.
FEED = [(768, 768, 4096, 1, (64, 1)), (1024, 1024, 1024, 1, (64, 1)), (1000, 1000, 1000, 1, (64, 1)),
        (512, 512, 8192, 1, (64, 4)), (768, 768, 4096, 2, (64, 1)), (1024, 1024, 1024, 2, (64, 1)),
        (700, 900, 3000, 2, (64, 1)), (64, 512, 8192, 2, (64, 16)), (2048, 2048, 2048, 1, (128, 1)),
        (512, 4096, 4096, 1, (128, 1)), (1000, 1000, 4000, 1, (128, 1)), (1152, 1152, 8192, 2, (128, 1)),
        (1024, 1024, 1024, 0, (64, 1)), (64, 512, 8192, 1, (64, 32)), (1920, 1920, 2000, 1, (128, 1))]
NOISE_FLOOR_PCT = 5.0  # docs/BENCHMARKS.md:2262-2280
MIN_ROUND_MS = 5.0     # GPU time per timed round, at least


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spread_pct(xs: list[float]) -> float:
    return 100.0 * (max(xs) - min(xs)) / float(np.median(xs))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rounds", type=int, default=5, help="rotated rounds per entrant (median reported)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--only", default="", help="comma list of sections: main,narrow,feed,grouped (default all)")
    ap.add_argument("--modes", default="plain,split3,split6", help="modes timed at 4096^3")
    ap.add_argument("--grouped-classes", type=int, default=16,
                    help="shape classes the grouped autotuner searches (heaviest first); -1: no tuning, "
                         "the grouped rule's plans (for counter runs that should see only the timed launches)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 256 MiB device copy (reference for traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the GEMM path has no "
              "multi-GPU mode in this build.", file=sys.stderr)
        return 2

    import torch

    import maxmulsk_amd as mm
    from maxmulsk_amd import inputs

    assert torch.cuda.is_available(), "no GPU visible"
    if mm.device_count() < 1:
        raise SystemExit("libmaxmulsk sees no gfx950 device: " + mm.last_error())
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    only = set(args.only.split(",")) if args.only else {"main", "narrow", "feed", "grouped"}

    def upload(x):
        return torch.from_numpy(x).to(dev)

    def make_call(dA, dB, dC, M, K, N, mode, plan=None):
        """One maxmulsk_sgemm_ex call with its arguments built once, as a C
        caller holds them: the timed loop then measures the library and the
        GPU, not the Python wrapper (which costs more than a small GEMM)."""
        opt = mm.Options(mode=mode, plan=mm.Plan(*(plan or (0, 0))), stream=sptr, async_=1,
                         workspace=None, workspace_bytes=0)
        fn, args_ = mm.lib().maxmulsk_sgemm_ex, (dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N,
                                                  ctypes.byref(opt))

        def call():
            st = fn(*args_)
            if st != mm.OK:
                raise RuntimeError(f"sgemm_ex status {st}: {mm.last_error()}")
        call.opt = opt  # keeps the options alive with the closure
        return call

    def block(fn, steps):
        """ms per call over `steps` back-to-back calls, HIP events on the stream."""
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for _ in range(steps):
            fn()
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / steps

    def rounds(fns: dict, steps: int, offset: int = 0) -> dict:
        """Warm-up, then args.rounds rounds in rotated order; per entrant the
        list of per-round ms per call. A round runs at least `steps` calls
        and at least MIN_ROUND_MS of GPU time (the reference sizes its samples
        the same way, adaptive repetitions of at least 0.3 s per sample), so
        short calls are not timed over a handful of microseconds."""
        names = list(fns)
        reps = {}
        for n in names:
            for _ in range(args.warmup):
                fns[n]()
            est = block(fns[n], steps)
            reps[n] = max(steps, int(np.ceil(MIN_ROUND_MS / max(est, 1e-4))))
        res = {n: [] for n in names}
        for r in range(args.rounds):
            for j in range(len(names)):
                n = names[(j + r + offset) % len(names)]
                res[n].append(block(fns[n], reps[n]))
        return res

    def tune(dA, dB, dC, M, K, N, mode):
        """Fill the plan cache for this shape with the autotuner's pick, so the
        planner's cache tier serves the timed calls. A MAXMULSK_PLAN override
        (the planner's first tier) skips tuning, e.g. under a profiler."""
        if os.environ.get("MAXMULSK_PLAN"):
            plan, src = mm.choose_plan(M, K, N, mode)
            assert src == 0, "MAXMULSK_PLAN did not parse as 'tile,split'"
            return plan
        # 5 repetitions per candidate and 15 per leader: the settings the
        # selector's table was collected with
        _, best = mm.autotune(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode, 5, 15)
        plan, src = mm.choose_plan(M, K, N, mode)
        assert src == 1 and plan == best, (plan, src, best)
        return plan

    if args.copy_calibration:
        # A plain streaming kernel of known size (256 MiB read, 256 MiB
        # written; torch's elementwise multiply), run three times, for scaling
        # hardware traffic counters.
        src_buf = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst_buf = torch.empty_like(src_buf)
        for _ in range(3):
            torch.mul(src_buf, 2.0, out=dst_buf)
        torch.cuda.synchronize()
        del src_buf, dst_buf

    def packed_kernel_ms(dA, dB, dC, M, K, N, mode, split, steps):
        """Average duration of the packed GEMM launch alone: both operands
        packed once, then rounds of `steps` launches between two events on
        the stream the kernel runs on (median of the rounds)."""
        pa = mm.Packed(dA.data_ptr(), M, K, 0, mode, stream=sptr)
        pb = mm.Packed(dB.data_ptr(), N, K, 1, mode, stream=sptr)

        def run():
            st = mm.sgemm_packed(pa, pb, dC.data_ptr(), split_k=split, stream=sptr, async_=True)
            if st != mm.OK:
                raise RuntimeError(f"sgemm_packed status {st}: {mm.last_error()}")

        ms = rounds({"k": run}, steps)["k"]
        pa.close()
        pb.close()
        return float(np.median(ms)), spread_pct(ms)

    out: dict = {}
    shapes = []

    if "main" in only:
        M = N = K = 4096
        log("generating 4096^3 inputs")
        dA, dB = upload(inputs.uniform(1, (M, K))), upload(inputs.uniform(2, (K, N)))
        dC = torch.empty((M, N), dtype=torch.float32, device=dev)
        modes = [(n, m) for n, m in (("plain", 0), ("split3", 1), ("split6", 2)) if n in args.modes.split(",")]
        plans = {n: tune(dA, dB, dC, M, K, N, m) for n, m in modes}
        # a plan call: planner (cache tier) + launch on the library's pooled
        # workspace, as maxmulsk_sgemm_v4 runs it
        times = rounds({n: make_call(dA, dB, dC, M, K, N, m) for n, m in modes}, args.steps)
        res = {}
        for n, m in modes:
            make_call(dA, dB, dC, M, K, N, m)()
            e = errors_torch(dA, dB, dC, K)
            ms = float(np.median(times[n]))
            tf = 2.0 * M * N * K / (ms * 1e-3) / 1e12
            source = "MAXMULSK_PLAN" if os.environ.get("MAXMULSK_PLAN") else "autotune cache"
            r = {"tflops": tf, "ms": ms, "round_ms": times[n], "spread_pct": spread_pct(times[n]),
                 "plan": list(plans[n]), "plan_source": source, "errors": e}
            # Does the call pack its operands first (the 256 x 256 packed path)?
            # The library's launch counters say.
            mm.reset_phase_counters()
            make_call(dA, dB, dC, M, K, N, m)()
            torch.cuda.synchronize()
            counts = mm.phase_counters()
            r["packs_operands"] = counts["pack"] > 0
            if r["packs_operands"]:
                # time its GEMM kernel alone on pre-packed operands
                kms, ksp = packed_kernel_ms(dA, dB, dC, M, K, N, m, plans[n][1], args.steps)
                r["gemm_kernel_ms"] = kms
                r["gemm_kernel_spread_pct"] = ksp
                r["packed_operands_tflops"] = 2.0 * M * N * K / (kms * 1e-3) / 1e12
                r["pack_share_of_call"] = 1.0 - kms / ms
                r["gemm_kernel"] = f"gemm_packed_kernel<{m}>"
            else:
                # one GEMM launch per call (plus the reduce with split-K)
                r["gemm_kernel_ms"] = ms
                r["gemm_kernel"] = (f"gemm_kernel<{m}, 256, 256> (direct tile, operands split on the fly)"
                                    if plans[n][0] == 256 and M > 128 else "gemm_kernel (direct tile)")
            res[n] = r
            log(f"4096^3 {n}: {ms:.4f} ms (spread {r['spread_pct']:.1f}%) {tf:.1f} TFLOP/s plan {plans[n]} "
                f"kernel {r['gemm_kernel_ms']:.4f} ms err {e}")
        if "split3" in res:
            # The drop-in case: maxmulsk_sgemm_v4 on host (pageable) arrays, so
            # each call stages A and B to the device and C back over PCIe.
            A_h, B_h = dA.cpu().numpy(), dB.cpu().numpy()
            C_h = np.empty((M, N), np.float32)
            os.environ["MAXMULSK_MODE"] = "split3"
            try:
                mm.check(mm.sgemm_v4(A_h, B_h, C_h, M, K, N))
                ts = []
                for _ in range(5):
                    t = time.perf_counter()
                    mm.check(mm.sgemm_v4(A_h, B_h, C_h, M, K, N))
                    ts.append(time.perf_counter() - t)
            finally:
                del os.environ["MAXMULSK_MODE"]
            hms = float(np.median(ts)) * 1e3
            res["split3"]["host_pointer_ms"] = hms
            res["split3"]["host_pointer_spread_pct"] = spread_pct(ts)
            res["split3"]["host_pointer_tflops"] = 2.0 * M * N * K / (hms * 1e-3) / 1e12
            log(f"4096^3 split3 through maxmulsk_sgemm_v4 on host arrays (PCIe staging): {hms:.2f} ms "
                f"{res['split3']['host_pointer_tflops']:.1f} TFLOP/s")
            del A_h, B_h, C_h
        # Context: the vendor library (torch.matmul -> hipBLASLt) on the same
        # operands, FP32 (no reduced-precision math) and plain BF16.
        torch.backends.cuda.matmul.allow_tf32 = False
        vendor = {}
        for name, dt in (("fp32", torch.float32), ("bf16", torch.bfloat16)):
            a, b = dA.to(dt), dB.to(dt)
            c = torch.empty((M, N), dtype=dt, device=dev)
            vt = rounds({name: lambda: torch.matmul(a, b, out=c)}, args.steps)[name]
            vms = float(np.median(vt))
            vendor[name] = {"ms": vms, "tflops": 2.0 * M * N * K / (vms * 1e-3) / 1e12, "spread_pct": spread_pct(vt)}
            log(f"4096^3 torch.matmul {name}: {vms:.4f} ms {vendor[name]['tflops']:.1f} TFLOP/s")
            del a, b, c
        # What the matrix cores sustain with register operands only (the
        # library's probe): random operand bits against constant ones.
        probe = {}
        for name, mode, iters in (("bf16_32x32x16", 1, 40000), ("fp32_32x32x2", 0, 20000)):
            probe[name] = {"random_operands_tflops": mm.probe_mfma(mode, True, iters),
                           "constant_operands_tflops": mm.probe_mfma(mode, False, iters)}
            log(f"MFMA probe {name}: random {probe[name]['random_operands_tflops']:.0f} TFLOP/s, "
                f"constant {probe[name]['constant_operands_tflops']:.0f} TFLOP/s")
        out["mfma_probe"] = probe
        out["main"] = res
        # the headline: the fastest split mode that stays inside the FP32 bound
        ok = [m for m in ("split3", "split6") if m in res and res[m]["errors"]["within_fp32_bound"]]
        head = max(ok, key=lambda m: res[m]["tflops"]) if ok else next(iter(res))
        out["headline_mode"] = head
        out["vendor"] = vendor
        del dA, dB, dC
        torch.cuda.empty_cache()

    if "narrow" in only:
        for i, (M, N, K) in enumerate(NARROW):
            dA, dB = upload(inputs.uniform(3, (M, K))), upload(inputs.uniform(4, (K, N)))
            dC = torch.empty((M, N), dtype=torch.float32, device=dev)
            row = {"shape_mnk": [M, N, K]}
            modes = (("plain", 0), ("split3", 1))
            plans = {n: tune(dA, dB, dC, M, K, N, m) for n, m in modes}
            times = rounds({n: make_call(dA, dB, dC, M, K, N, m) for n, m in modes}, max(args.steps, 20), i)
            for n, m in modes:
                make_call(dA, dB, dC, M, K, N, m)()
                e = errors_torch(dA, dB, dC, K)
                ms = float(np.median(times[n]))
                flops = 2.0 * M * N * K
                alg_bytes = 4.0 * (M * K + K * N + M * N)
                row[n] = {"tflops": flops / (ms * 1e-3) / 1e12, "ms": ms, "spread_pct": spread_pct(times[n]),
                          "plan": list(plans[n]), "hbm_gbs": alg_bytes / (ms * 1e-3) / 1e9, "errors": e}
                # The floor of this plan's launch structure: an empty kernel and
                # a pure streaming kernel on the same grid, plus the same reduce.
                fl = mm.probe_stream(M, K, N, tuple(plans[n]), stream=stream)
                fl["call_over_floor"] = ms * 1e3 / fl["stream_reduce_us"]
                row[n]["floor"] = fl
                log(f"{M}x{N}x{K} {n}: {ms:.4f} ms (spread {row[n]['spread_pct']:.1f}%) "
                    f"{row[n]['tflops']:.1f} TFLOP/s {row[n]['hbm_gbs']:.0f} GB/s plan {plans[n]} "
                    f"ratio {e['bound_ratio']:.3g}; floor {fl['stream_reduce_us']:.1f} us (empty kernel "
                    f"{fl['empty_us']:.1f}, streaming {fl['stream_us']:.1f})")
            shapes.append(row)
            del dA, dB, dC
        torch.cuda.empty_cache()

    feed = []
    if "feed" in only:
        # The same plan under both feeds, in the same rotated rounds; the
        # feed setting is read at each launch, so each entrant sets it.
        prev = mm.get_feed()
        for i, (M, N, K, mode, plan) in enumerate(FEED):
            dA, dB = upload(inputs.uniform(5, (M, K))), upload(inputs.uniform(6, (K, N)))
            dC = torch.empty((M, N), dtype=torch.float32, device=dev)
            call = make_call(dA, dB, dC, M, K, N, mode, plan)

            def entrant(setting, call=call):
                def f():
                    mm.set_feed(setting)
                    call()
                return f
            names = {"direct": mm.FEED_DIRECT, "decoupled": mm.FEED_DECOUPLED}
            times = rounds({n: entrant(v) for n, v in names.items()}, max(args.steps, 20), i)
            outs = {}
            for n, v in names.items():
                mm.set_feed(v)
                call()
                torch.cuda.synchronize()
                outs[n] = dC.clone()
            e = errors_torch(dA, dB, dC, K)
            mm.set_feed(mm.FEED_AUTO)
            row = {"shape_mnk": [M, N, K], "mode": ["plain", "split3", "split6"][mode], "plan": list(plan),
                   "decoupled_variant": None, "auto_takes_decoupled": mm.feed_variant(M, K, N, mode, plan) > 0,
                   "identical": bool(torch.equal(outs["direct"], outs["decoupled"])), "errors": e}
            mm.set_feed(mm.FEED_DECOUPLED)
            row["decoupled_variant"] = mm.feed_variant(M, K, N, mode, plan)
            for n in names:
                ms = float(np.median(times[n]))
                row[n] = {"ms": ms, "spread_pct": spread_pct(times[n]), "tflops": 2.0 * M * N * K / (ms * 1e-3) / 1e12}
            row["speedup"] = row["direct"]["ms"] / row["decoupled"]["ms"]
            log(f"feed {M}x{N}x{K} {row['mode']} {plan}: direct {row['direct']['ms']*1e3:.1f} us, decoupled "
                f"{row['decoupled']['ms']*1e3:.1f} us ({row['speedup']:.3f}x, variant {row['decoupled_variant']}, "
                f"auto {'decoupled' if row['auto_takes_decoupled'] else 'direct'}), identical {row['identical']}")
            feed.append(row)
            del dA, dB, dC
        mm.set_feed(prev)
        torch.cuda.empty_cache()

    grouped = None
    if "grouped" in only:
        grouped = bench_grouped(torch, mm, inputs, args, log, rounds)

    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline(args.cpu_seconds)

    line = {
        "metric": METRIC,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "data": "synthetic: U[-1,1) FP32 from splitmix64 seeds (A seed 1, B seed 2; narrow shapes 3, 4)",
        "method": {"rounds": args.rounds, "steps_per_round": args.steps, "order": "rotated per round",
                   "figure": "median over rounds of ms per call (HIP events around each round's calls; "
                             "a round runs max(steps, 5 ms of calls))",
                   "spread": "(max - min) / median over rounds, percent",
                   "noise_floor_pct": NOISE_FLOOR_PCT,
                   "source": "bench_final5.cpp:72-151; docs/BENCHMARKS.md:2262-2280"},
    }
    if "main" in out and "plain" in out["main"] and len(out["main"]) > 1:
        head = out["headline_mode"]
        r = out["main"][head]
        line["value"] = r["tflops"]
        line["ms_per_step"] = r["ms"]
        line["spread_pct"] = r["spread_pct"]
        line["dtype"] = "fp32 in/out; bf16 " + ("3-term" if head == "split3" else "6-term") + " split MFMA, fp32 accumulate"
        line["config"] = {"workload": f"split-BF16 GEMM C=A*B, M=N=K=4096, mode {head}, FP32 inputs/outputs",
                          "M": 4096, "N": 4096, "K": 4096, "mode": head, "plan": r["plan"]}
        line["plain_fp32_mfma_tflops"] = out["main"]["plain"]["tflops"]
        line["speedup_vs_plain"] = r["tflops"] / out["main"]["plain"]["tflops"]
        line["max_rel_vs_plain"] = r["errors"]["max_rel"] / out["main"]["plain"]["errors"]["max_rel"]
        line["max_abs_4096"] = {m: out["main"][m]["errors"]["max_abs"] for m in out["main"]}
        line["reference_max_abs_4096"] = {"value": 2.09e-4, "hardware": "Apple M4, SME 4x1 kernel",
                                          "source": "docs/BENCHMARKS.md:112-121"}
        line["modes_4096"] = out["main"]
        line["vendor_torch_matmul_4096"] = out["vendor"]
        line["plan_source"] = r["plan_source"]
        # Dominant kernel: the split GEMM kernel. Its algorithmic work on the
        # BF16 MFMA units is products x 2MNK per launch (one launch covers the
        # whole 4096^3 product when the plan has no split-K).
        products = 3 if head == "split3" else 6
        issued = products * 2.0 * 4096 ** 3
        kms = r["gemm_kernel_ms"]
        achieved = issued / (kms * 1e-3) / 1e12
        how = ("kernel timed alone on pre-packed operands with HIP events on its stream" if r["packs_operands"]
               else "the call is this one launch, timed with HIP events on its stream")
        line["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK, "unit": "TFLOP/s",
                            "frac": achieved / BF16_PEAK,
                            "traffic": pmc_traffic("packed" if r["packs_operands"] else "direct256", head, r["plan"]),
                            "kernel": r["gemm_kernel"], "kernel_ms": kms,
                            "note": f"BF16 MFMA FLOPs = {products} x 2MNK per launch; {how}"}
        line["roofline"]["frac_of_random_operand_ceiling"] = (
            achieved / out["mfma_probe"]["bf16_32x32x16"]["random_operands_tflops"])
        line["mfma_probe"] = out["mfma_probe"]
        line["plain_roofline"] = {"bound": "mfma", "achieved": out["main"]["plain"]["tflops"],
                                  "peak": FP32_PEAK, "unit": "TFLOP/s",
                                  "frac": out["main"]["plain"]["tflops"] / FP32_PEAK}
    if shapes:
        line["narrow_m"] = shapes
    if feed:
        line["feed"] = feed
    if grouped:
        line["grouped"] = grouped
    if cpu:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)
    return 0


def errors_torch(dA, dB, dC, K: int) -> dict:
    """Error of the device FP32 product dC against FP64 dA*dB, every element,
    with torch FP64 matmuls on the GPU: the same figures as the FP64 oracle's
    (bound_ratio, max_rel, rel_frob, max_abs). Pinned against the oracle in
    tests/test_gpu_oracle.py."""
    import torch

    tiny = float(np.finfo(np.float64).tiny)
    A64, B64 = dA.double(), dB.double()
    C64 = A64 @ B64
    AB = A64.abs() @ B64.abs()
    d = (dC.double() - C64).abs()
    scale = AB.clamp_min(tiny)
    ratio = float((d / (K * U32 * scale)).max())
    max_rel = float((d / torch.maximum(C64.abs(), U32 * scale)).max())
    rel_frob = float(d.norm() / C64.norm())
    return {"bound_ratio": ratio, "max_rel": max_rel, "rel_frob": rel_frob, "max_abs": float(d.max()),
            "within_fp32_bound": ratio <= 1.0}


def pmc_traffic(kernel: str, mode: str, plan) -> int | None:
    """HBM bytes per launch of the 4096^3 GEMM kernel (kernel "packed" or
    "direct256"), from the committed rocprofv3 counter summary (FETCH_SIZE x
    its calibration factor + WRITE_SIZE). None when no summary matches this
    kernel, mode and plan."""
    path = os.path.join(ROOT, "profiles", f"gemm_{kernel}_{mode}_4096_pmc.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for ln in open(path):
        if ":" in ln and not ln.startswith(" "):
            k, _, v = ln.partition(":")
            vals[k.strip()] = v.strip()
    if vals.get("plan") != f"{plan[0]},{plan[1]}":
        return None
    try:
        return int(float(vals["traffic_bytes_per_launch"]))
    except (KeyError, ValueError):
        return None


def config4_batch(torch, inputs):
    """The 256-problem batch: M ~ U{1..512}, N in {512, 2112, 4096, 7168},
    K in {512, 1536, 7168, 16384}, drawn from one seed. A and B of every
    problem are slices of one random pool (read-only inputs may alias); every
    C is its own buffer. Returns (specs as (M, K, N), pool, outputs, problems
    as (A, B, C, M, K, N) device pointers)."""
    rng = np.random.default_rng(2024)
    specs = [(int(rng.integers(1, 513)), int(rng.choice([512, 1536, 7168, 16384])),
              int(rng.choice([512, 2112, 4096, 7168]))) for _ in range(256)]
    pool_n = 16384 * 7168 + 4096
    pool = torch.from_numpy(inputs.uniform(9, (pool_n,))).cuda()
    outs, probs = [], []
    for M, K, N in specs:
        oa = int(rng.integers(0, pool_n - M * K)) & ~3
        ob = int(rng.integers(0, pool_n - K * N)) & ~3
        c = torch.empty(M * N, dtype=torch.float32, device="cuda")
        outs.append(c)
        probs.append((pool.data_ptr() + 4 * oa, pool.data_ptr() + 4 * ob, c.data_ptr(), M, K, N))
    return specs, pool, outs, probs


def bench_grouped(torch, mm, inputs, args, log, rounds) -> dict:
    """The 256-problem batch (config4_batch) in one persistent launch replayed
    from a hipGraph, per mode: the grouped autotuner picks the plans of the
    heaviest shape classes (the others keep the grouped rule), the launch is
    timed, and every element of every problem's C is checked against FP64."""
    specs, pool, outs, probs = config4_batch(torch, inputs)
    flops = sum(2.0 * M * K * N for M, K, N in specs)
    alg_bytes = sum(4.0 * (M * K + K * N + M * N) for M, K, N in specs)
    res = {"problems": 256, "total_gflop": flops / 1e9, "total_alg_gb": alg_bytes / 1e9}
    poolv = pool
    plans_by_mode = {}
    gps = {}
    for mode_name, mode in (("plain", 0), ("split3", 1)):
        rule = mm.grouped_plans(probs, mode)  # fresh process: the grouped cache is empty
        t = time.perf_counter()
        if args.grouped_classes >= 0:
            tuned, tuned_ms = mm.grouped_autotune(probs, mode, max_classes=args.grouped_classes, reps=3)
        else:
            tuned, tuned_ms = rule, None
        tune_s = time.perf_counter() - t
        used = mm.grouped_plans(probs, mode)
        assert used == tuned, "the grouped cache does not serve the autotuned plans"
        plans_by_mode[mode_name] = {"autotune_seconds": tune_s, "autotune_launch_ms": tuned_ms,
                                    "classes_searched": args.grouped_classes,
                                    "problems_plan_changed_by_autotuner": int(sum(a != b for a, b in zip(rule, used))),
                                    "distinct_plans": sorted({tuple(p) for p in used})}
        gps[mode_name] = mm.GroupedPlan(probs, mode=mode)
        log(f"grouped {mode_name}: autotuned {args.grouped_classes} classes in {tune_s:.1f} s; "
            f"{plans_by_mode[mode_name]['problems_plan_changed_by_autotuner']} problems changed plan")
    # The split modes' launches run on the decoupled feed by default; the direct
    # grouped kernel on the same plans is timed beside them for context.
    directs = {}
    prev = mm.set_feed(mm.FEED_DIRECT)
    for mode_name, mode in (("plain", 0), ("split3", 1)):
        if gps[mode_name].info()["feed"]:
            directs[mode_name] = mm.GroupedPlan(probs, mode=mode)
    mm.set_feed(prev)
    times = rounds({**{n: gp.launch for n, gp in gps.items()},
                    **{n + "_direct": gp.launch for n, gp in directs.items()}}, 5)
    for gp in directs.values():
        gp.close()
    for mode_name, gp in gps.items():
        info = gp.info()
        # every problem checked against FP64 after a replay
        for c in outs:
            c.fill_(float("nan"))
        gp.launch()
        torch.cuda.synchronize()
        worst = {"bound_ratio": 0.0, "max_rel": 0.0, "nonfinite_problems": 0}
        for (a, b, c_ptr, M, K, N), c in zip(probs, outs):
            base = poolv.data_ptr()
            A = poolv[(a - base) // 4:(a - base) // 4 + M * K].view(M, K)
            B = poolv[(b - base) // 4:(b - base) // 4 + K * N].view(K, N)
            e = errors_torch(A, B, c.view(M, N), K)
            if not np.isfinite(e["bound_ratio"]):
                worst["nonfinite_problems"] += 1
            worst["bound_ratio"] = max(worst["bound_ratio"], e["bound_ratio"])
            worst["max_rel"] = max(worst["max_rel"], e["max_rel"])
        gp.close()
        ms = float(np.median(times[mode_name]))
        res[mode_name] = {"ms": ms, "spread_pct": spread_pct(times[mode_name]),
                          "tflops": flops / (ms * 1e-3) / 1e12, "work_items": info["work_items"],
                          "graph": info["graph"], "decoupled_feed": info["feed"],
                          "direct_kernel_ms": (float(np.median(times[mode_name + "_direct"]))
                                               if mode_name + "_direct" in times else None),
                          "plans": plans_by_mode[mode_name],
                          "errors_worst_of_256": worst,
                          "within_fp32_bound": worst["bound_ratio"] <= 1.0 and worst["nonfinite_problems"] == 0}
        dk = res[mode_name]["direct_kernel_ms"]
        log(f"grouped 256 problems {mode_name}: {ms:.3f} ms {res[mode_name]['tflops']:.1f} TFLOP/s "
            f"items {info['work_items']} worst bound ratio {worst['bound_ratio']:.3g}"
            + (f"; decoupled feed (direct grouped kernel {dk:.3f} ms)" if dk else ""))
    return res


def cpu_baseline(seconds: float) -> dict:
    """The oracle's restatement of the reference's portable FP32 loop
    (Utils::multiply_scalar), single thread, at 256^3, with its error against
    FP64; plus the host BLAS (numpy float32 matmul) at 256^3 and 4096^3 for
    context."""
    from oracle import oracle as orc

    n = 256
    A = orc.uniform(1, (n, n))
    B = orc.uniform(2, (n, n))
    C32 = orc.sgemm_portable(A, B)  # warm, and the result checked below
    err = orc.errors(C32, orc.gemm_f64(A, B), orc.absgemm_f64(A, B), n)
    reps, t0 = 0, time.perf_counter()
    samples = []
    while time.perf_counter() - t0 < seconds:
        t = time.perf_counter()
        orc.sgemm_portable(A, B)
        samples.append(time.perf_counter() - t)
        reps += 1
    med = float(np.median(samples))
    flops = 2.0 * n ** 3
    res = {"value": flops / med / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
           "sample": f"oracle restatement of Utils::multiply_scalar (common/utils.hpp:50-61), "
                     f"M=N=K=256, {reps} calls in {time.perf_counter() - t0:.1f} s, median",
           "gflops": flops / med / 1e9,
           "errors_vs_fp64": {k: err[k] for k in ("bound_ratio", "max_rel", "rel_frob", "max_abs")}}
    try:
        info = np.__config__.CONFIG["Build Dependencies"]["blas"]
        blas = f"{info.get('name')} {info.get('version')}"
    except Exception:  # pragma: no cover - informational only
        blas = "numpy BLAS"
    # spread of the portable samples: (p90 - p10) / median, in percent
    res["spread_pct_p10_p90"] = 100.0 * (float(np.quantile(samples, 0.9)) - float(np.quantile(samples, 0.1))) / med

    def blas_rate(m: int, seconds_each: float) -> float:
        a = orc.uniform(5, (m, m))
        b = orc.uniform(6, (m, m))
        a @ b
        ts = []
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < seconds_each or len(ts) < 3:
            t = time.perf_counter()
            a @ b
            ts.append(time.perf_counter() - t)
        return 2.0 * m ** 3 / float(np.median(ts)) / 1e12

    each = min(3.0, seconds)
    rows = {"all_threads": {f"{m}^3": blas_rate(m, each) for m in (256, 4096)}}
    try:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=1, user_api="blas"):
            rows["1_thread"] = {f"{m}^3": blas_rate(m, each) for m in (256, 1024)}
    except ImportError:  # pragma: no cover - informational only
        pass
    res["host_blas_sgemm_tflops"] = rows
    res["host_blas"] = f"{blas}; all_threads = {os.cpu_count()} logical CPUs visible"
    return res


if __name__ == "__main__":
    sys.exit(main())
