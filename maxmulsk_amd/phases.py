"""Run GEMM calls through the C ABI and print the library's phase counters.

    python -m maxmulsk_amd.phases --shape 512,1024,768 --mode split3 --plan 256,2 --repeat 3 [--host]

Shapes are M,K,N (the ABI's order). --host passes numpy arrays (staged through
the device); otherwise the operands live in device memory. With
MAXMULSK_ROCTX=1 the host phases also appear as roctx ranges under
rocprofv3 --marker-trace.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

import maxmulsk_amd as mm
from maxmulsk_amd import inputs


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--shape", default="512,1024,768", help="M,K,N")
    ap.add_argument("--mode", default="plain", choices=sorted(mm.MODES))
    ap.add_argument("--plan", default="", help="tile,split (default: the planner)")
    ap.add_argument("--repeat", type=int, default=3)
    ap.add_argument("--host", action="store_true", help="host arrays (reference-style call)")
    args = ap.parse_args(argv)
    M, K, N = (int(x) for x in args.shape.split(","))
    plan = tuple(int(x) for x in args.plan.split(",")) if args.plan else None
    A = inputs.uniform(1, (M, K))
    B = inputs.uniform(2, (K, N))
    if args.host:
        a, b, c = A, B, np.empty((M, N), np.float32)
        ptrs = (a, b, c)
    else:
        import torch
        a, b = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        c = torch.empty((M, N), dtype=torch.float32, device="cuda")
        ptrs = (a.data_ptr(), b.data_ptr(), c.data_ptr())
    mm.reset_phase_counters()
    for _ in range(args.repeat):
        mm.check(mm.sgemm_ex(*ptrs, M, K, N, mode=args.mode, plan=plan))
    print(json.dumps({"shape_mkn": [M, K, N], "mode": args.mode, "plan": plan, "repeat": args.repeat,
                      "host": args.host, "counters": mm.phase_counters()}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
