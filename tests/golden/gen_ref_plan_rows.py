"""Writes tests/golden/ref_plan_rows.npz: the plan figures the reference's own
result files record, as plain numbers (data, not source).

Run from the repo root in a container that holds the reference tree:
    python tests/golden/gen_ref_plan_rows.py [/path/to/reference]

What it extracts (every row, nothing filtered except the header):

* ``nb``: bench/results/2026-09-13/nb_sweep.csv and nb_sweep_ext.csv, the
  Nc/Mc/Kc blocking sweep of the v4 kernel (bench_nb_sweep.cpp). Per row
  M, K, N, Mc, Nc, Kc, a_packs, b_packs, mk_calls, native(path == "native"),
  and ``nb_mib`` = packA_MiB, packB_MiB as printed (three decimals). The
  sweep prints NB::capacity and NB::counts of each blocking
  (sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:429-466).
* ``v4c``: bench/results/2026-09-13/v4c3.csv. Per row M, K, N,
  profile (0 = v4a, 1 = v4b), Mc, Nc, Kc, and ``v4c_mib`` = packA_MiB,
  packB_MiB: the v4c rule's pick (sme/v4/sme-1x4-v4c-dispatch.cpp:17-22) and
  its workspace.
* ``final5``: bench/results/2026-09-13/final5.csv. Per row M, K, N,
  v3_Mtile, v3_Ntile, v3_Kc (the v3 tuning table's pick,
  sme/support/gemm_tuning.cpp:81-100), and ``final5_mib`` = v4a_packA,
  v4a_packB, v4b_packA, v4b_packB (NB::capacity of both v4c blockings).
* ``headline``: bench/results/2026-09-09/headline_run{1,2,3}.csv and
  tuned_run{1,2,3}.csv, MaxMulSK rows. Per row M, K, N, M_tile, N_tile, Kc,
  ir_outer (the "blocking" column "Mt/Nt/Kc/swap|-").
"""
from __future__ import annotations

import csv
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "ref_plan_rows.npz")


def _rows(path: str) -> list[dict]:
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def main(ref: str) -> None:
    r13 = os.path.join(ref, "bench", "results", "2026-09-13")
    r09 = os.path.join(ref, "bench", "results", "2026-09-09")

    nb, nb_mib = [], []
    for name in ("nb_sweep.csv", "nb_sweep_ext.csv"):
        for r in _rows(os.path.join(r13, name)):
            nb.append([int(r[k]) for k in ("M", "K", "N", "Mc", "Nc", "Kc", "a_packs", "b_packs", "mk_calls")]
                      + [1 if r["path"] == "native" else 0])
            nb_mib.append([float(r["packA_MiB"]), float(r["packB_MiB"])])

    v4c, v4c_mib = [], []
    for r in _rows(os.path.join(r13, "v4c3.csv")):
        prof = {"v4a": 0, "v4b": 1}[r["v4c_profile"]]
        v4c.append([int(r["M"]), int(r["K"]), int(r["N"]), prof, int(r["Mc"]), int(r["Nc"]), int(r["Kc"])])
        v4c_mib.append([float(r["packA_MiB"]), float(r["packB_MiB"])])

    final5, final5_mib = [], []
    for r in _rows(os.path.join(r13, "final5.csv")):
        final5.append([int(r[k]) for k in ("M", "K", "N", "v3_Mtile", "v3_Ntile", "v3_Kc")])
        final5_mib.append([float(r[k]) for k in ("v4a_packA", "v4a_packB", "v4b_packA", "v4b_packB")])

    headline = []
    for stem in ("headline_run", "tuned_run"):
        for i in (1, 2, 3):
            for r in _rows(os.path.join(r09, f"{stem}{i}.csv")):
                if r["implementation"] != "MaxMulSK":
                    continue
                mt, nt, kc, sw = r["blocking"].split("/")
                headline.append([int(r["M"]), int(r["K"]), int(r["N"]), int(mt), int(nt), int(kc),
                                 1 if sw == "swap" else 0])

    np.savez_compressed(
        OUT,
        nb=np.array(nb, np.int64), nb_mib=np.array(nb_mib, np.float64),
        v4c=np.array(v4c, np.int64), v4c_mib=np.array(v4c_mib, np.float64),
        final5=np.array(final5, np.int64), final5_mib=np.array(final5_mib, np.float64),
        headline=np.array(headline, np.int64),
    )
    print(f"wrote {OUT}: nb {len(nb)}, v4c {len(v4c)}, final5 {len(final5)}, headline {len(headline)} rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
