"""Learned tile selector: data collection with the autotuner, training with the
library's own MLP kernels (forward, backward, Adam on the GPU), and held-out
regret evaluation against the autotuner's best.

    python -m maxmulsk_amd.selector --train 5000 --eval 500 --out <dir>

writes <dir>/selector_dataset.npz (shapes, modes, per-candidate times),
<dir>/maxmulsk_selector.bin (weights) and <dir>/selector_report.json.
Installing the weights next to libmaxmulsk.so makes the planner use them.

Regret of a choice = t(chosen) / t(best) - 1, both read from the autotuner's
table for that shape (the method of the reference's blocking sweep,
bench/results/2026-09-13/source/bench_nb_sweep.cpp, and of the v5.1 fixed
plan's "0.978 of each shape's best" figure, sme/v5/sme-1x4-fusedpack-v51.hpp:57-62).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

import maxmulsk_amd as mm

NCAND = 18  # maxmulsk_num_candidates(): tile {64, 128, 256} x split-K {1, 2, 4, 8, 16, 32}
FIXED_PLAN = (128, 1)  # one plan for every shape: the "no selector" baseline


def draw_shapes(n: int, seed: int, lo: int = 1, hi: int = 8192) -> tuple[np.ndarray, np.ndarray]:
    """M, N, K = round(2^U[log2 lo, log2 hi]); mode uniform over plain/split3/split6."""
    rng = np.random.default_rng(seed)
    e = rng.uniform(np.log2(lo), np.log2(hi), (n, 3))
    mnk = np.clip(np.rint(np.exp2(e)), lo, hi).astype(np.int64)
    modes = rng.integers(0, 3, n).astype(np.int64)
    return mnk, modes


def collect(mnk: np.ndarray, modes: np.ndarray, stage1: int = 3, stage2: int = 6, log=print) -> np.ndarray:
    """Autotuner times (ms) per shape and candidate; -1 where a candidate failed."""
    import torch
    M_max, N_max, K_max = (int(mnk[:, i].max()) for i in range(3))
    g = torch.Generator(device="cuda").manual_seed(1234)
    A = torch.rand(M_max * K_max, device="cuda", generator=g) * 2 - 1
    B = torch.rand(K_max * N_max, device="cuda", generator=g) * 2 - 1
    Cb = torch.empty(M_max * N_max, device="cuda")
    torch.cuda.synchronize()
    times = np.full((len(mnk), NCAND), -1.0, np.float32)
    t0 = time.time()
    for i, ((M, N, K), mode) in enumerate(zip(mnk.tolist(), modes.tolist())):
        tm, _ = mm.autotune(A.data_ptr(), B.data_ptr(), Cb.data_ptr(), M, K, N, mode, stage1, stage2)
        times[i] = tm
        if (i + 1) % 250 == 0:
            log(f"[selector] collected {i + 1}/{len(mnk)} shapes in {time.time() - t0:.1f} s")
    mm.lib().maxmulsk_cache_clear()  # the tuning runs filled the plan cache
    return times


def features(mnk: np.ndarray, modes: np.ndarray) -> np.ndarray:
    X = np.empty((len(mnk), 4), np.float32)
    for i, ((M, N, K), mode) in enumerate(zip(mnk.tolist(), modes.tolist())):
        X[i] = mm.selector_features(M, K, N, mode)
    return X


def targets(times: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    valid = times > 0
    T = np.where(valid, np.log2(np.where(valid, times, 1.0)), 0.0).astype(np.float32)
    return T, valid.astype(np.float32)


def init_params(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    H, IN = mm.MLP_HIDDEN, mm.MLP_IN
    parts = [rng.normal(0, np.sqrt(2 / IN), H * IN), np.full(H, 0.01), rng.normal(0, np.sqrt(2 / H), H * H),
             np.full(H, 0.01), rng.normal(0, np.sqrt(1 / H), NCAND * H), np.zeros(NCAND)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == mm.lib().maxmulsk_mlp_num_params()
    return p


class Trainer:
    """Device buffers and one optimisation step through the C ABI."""

    def __init__(self, params: np.ndarray, batch: int):
        import torch
        self.torch = torch
        self.batch = batch
        self.p = torch.from_numpy(params.copy()).cuda()
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        H = mm.MLP_HIDDEN
        self.H1 = torch.empty(batch * H, device="cuda")
        self.H2 = torch.empty(batch * H, device="cuda")
        self.Y = torch.empty(batch * NCAND, device="cuda")
        self.loss = torch.zeros(1, device="cuda")
        self.t = 0

    def forward(self, X):
        n = X.shape[0]
        mm.check(mm.lib().maxmulsk_mlp_forward(self.p.data_ptr(), X.data_ptr(), n, self.H1.data_ptr(),
                                               self.H2.data_ptr(), self.Y.data_ptr(), None))
        return self.Y[: n * NCAND].view(n, NCAND)

    def step(self, X, T, Mk, lr: float) -> None:
        n = X.shape[0]
        self.forward(X)
        mm.check(mm.lib().maxmulsk_mlp_backward(self.p.data_ptr(), X.data_ptr(), n, self.H1.data_ptr(),
                                                self.H2.data_ptr(), self.Y.data_ptr(), T.data_ptr(),
                                                Mk.data_ptr(), self.g.data_ptr(), self.loss.data_ptr(), None))
        self.t += 1
        mm.check(mm.lib().maxmulsk_mlp_adam(self.p.data_ptr(), self.g.data_ptr(), self.m.data_ptr(),
                                            self.v.data_ptr(), self.t, lr, 0.9, 0.999, 1e-8, None))


def train(X: np.ndarray, T: np.ndarray, Mk: np.ndarray, steps: int = 40000, batch: int = 1024,
          lr: float = 3e-3, seed: int = 7, log=print) -> tuple[np.ndarray, list[float]]:
    import torch
    rng = np.random.default_rng(seed)
    tr = Trainer(init_params(seed), batch)
    Xd = torch.from_numpy(X).cuda()
    Td = torch.from_numpy(T).cuda()
    Md = torch.from_numpy(Mk).cuda()
    n = X.shape[0]
    hist = []
    for s in range(steps):
        idx = torch.from_numpy(rng.choice(n, min(batch, n), replace=False)).cuda()
        xb, tb, mb = Xd[idx].contiguous(), Td[idx].contiguous(), Md[idx].contiguous()
        cur_lr = lr * (0.1 ** (s / steps))  # exponential decay to lr/10
        tr.step(xb, tb, mb, cur_lr)
        if s % 500 == 0 or s == steps - 1:
            torch.cuda.synchronize()
            hist.append(float(tr.loss.item()))
            log(f"[selector] step {s} loss {hist[-1]:.4f}")
    torch.cuda.synchronize()
    return tr.p.cpu().numpy(), hist


def predict_batch(params: np.ndarray, X: np.ndarray) -> np.ndarray:
    import torch
    tr = Trainer(params, X.shape[0])
    Y = tr.forward(torch.from_numpy(X).cuda())
    torch.cuda.synchronize()
    return Y.cpu().numpy()


def regret_of(choice: np.ndarray, times: np.ndarray) -> np.ndarray:
    best = np.where(times > 0, times, np.inf).min(1)
    t = times[np.arange(len(times)), choice]
    t = np.where(t > 0, t, np.nan)
    return t / best - 1.0


def cand_index(plan) -> int:
    return mm.candidates().index(tuple(plan))


def evaluate(params: np.ndarray, mnk: np.ndarray, modes: np.ndarray, times: np.ndarray) -> dict:
    """Regret of the network, the shape rule and the fixed plan on a table."""
    X = features(mnk, modes)
    Y = predict_batch(params, X)
    mlp = Y.argmin(1)
    rule = np.array([cand_index(mm.rule_plan(int(M), int(K), int(N), int(md)))
                     for (M, N, K), md in zip(mnk.tolist(), modes.tolist())])
    fixed = np.full(len(mnk), cand_index(FIXED_PLAN))

    def stats(r):
        r = r[np.isfinite(r)]
        return {"mean": float(r.mean()), "median": float(np.median(r)), "p90": float(np.quantile(r, 0.9)),
                "within_5pct": float(np.mean(r <= 0.05)), "n": int(r.size)}

    return {"mlp": stats(regret_of(mlp, times)), "rule": stats(regret_of(rule, times)),
            "fixed_plan": stats(regret_of(fixed, times)), "mlp_choice": mlp.tolist()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=5000)
    ap.add_argument("--eval", type=int, default=500)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--out", default=os.path.dirname(os.path.abspath(__file__)))
    ap.add_argument("--dataset", default="", help="reuse a collected dataset instead of timing")
    ap.add_argument("--stage1", type=int, default=5, help="autotuner repetitions per candidate")
    ap.add_argument("--stage2", type=int, default=15, help="autotuner repetitions per leader")
    ap.add_argument("--seeds", default="7,11,13,17",
                    help="initialisation seeds; with several, the one with the lowest regret on a "
                         "validation split of the training shapes (the last 500) is trained on all of them")
    args = ap.parse_args()

    def log(s):
        print(s, file=sys.stderr, flush=True)

    if args.dataset:
        d = np.load(args.dataset)
        mnk, modes, times = d["mnk"], d["modes"], d["times"]
    else:
        mnk, modes = draw_shapes(args.train + args.eval, seed=2027)
        t0 = time.time()
        times = collect(mnk, modes, stage1=args.stage1, stage2=args.stage2, log=log)
        log(f"[selector] collection took {time.time() - t0:.1f} s")
        np.savez_compressed(os.path.join(args.out, "selector_dataset.npz"), mnk=mnk, modes=modes, times=times,
                            n_train=args.train)
    n_train = args.train
    seeds = [int(x) for x in args.seeds.split(",") if x]
    seed = seeds[0]
    val_regret = {}
    if len(seeds) > 1:
        # seed selection on a validation split of the training shapes; the
        # held-out shapes take no part in it
        n_fit = n_train - 500
        Xf = features(mnk[:n_fit], modes[:n_fit])
        Tf, Mf = targets(times[:n_fit])
        for sd in seeds:
            p_s, _ = train(Xf, Tf, Mf, steps=args.steps, seed=sd, log=lambda _s: None)
            val_regret[sd] = evaluate(p_s, mnk[n_fit:n_train], modes[n_fit:n_train], times[n_fit:n_train])["mlp"]["mean"]
            log(f"[selector] seed {sd}: validation regret {val_regret[sd]:.4f}")
        seed = min(val_regret, key=val_regret.get)
    X = features(mnk[:n_train], modes[:n_train])
    T, Mk = targets(times[:n_train])
    t0 = time.time()
    params, hist = train(X, T, Mk, steps=args.steps, seed=seed, log=log)
    train_s = time.time() - t0
    ev = evaluate(params, mnk[n_train:], modes[n_train:], times[n_train:])
    tr = evaluate(params, mnk[:n_train], modes[:n_train], times[:n_train])
    path = os.path.join(args.out, "maxmulsk_selector.bin")
    mm.check(mm.lib().maxmulsk_selector_save(path.encode(), params.ctypes.data))
    report = {"train_shapes": n_train, "eval_shapes": int(len(mnk) - n_train), "steps": args.steps,
              "seed": seed, "validation_regret_by_seed": val_regret,
              "train_seconds": train_s, "loss_history": hist,
              "heldout": {k: v for k, v in ev.items() if k != "mlp_choice"},
              "train": {k: v for k, v in tr.items() if k != "mlp_choice"}}
    with open(os.path.join(args.out, "selector_report.json"), "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(report["heldout"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
