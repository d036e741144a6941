"""The shape sweep (maxmulsk_amd/sweep.py, the reference's benchmark method of
bench_final5.cpp:72-151) end to end on two small shapes: one CSV row per shape,
every entrant checked against FP64 before timing, medians and spreads over the
rotated rounds."""
import csv

import pytest

import maxmulsk_amd as mm
from maxmulsk_amd import sweep
from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu


def test_sweep_two_shapes(tmp_path):
    torch_cuda()
    out = tmp_path / "sweep.csv"
    assert sweep.main(["--shapes", "64,512,8192;300,257,515", "--runs", "3", "--sample-seconds", "0.01",
                       "--out", str(out)]) == 0
    rows = list(csv.DictReader(open(out)))
    assert [(int(r["M"]), int(r["N"]), int(r["K"])) for r in rows] == [(64, 512, 8192), (300, 257, 515)]
    cands = {f"{t}x{s}" for t, s in mm.candidates()}
    for r in rows:
        for e in ("plain", "split3", "split6"):
            assert r[f"plan_{e}"] in cands
            assert float(r[f"bound_{e}"]) <= 1.0
            assert float(r[f"{e}_ms"]) > 0 and float(r[f"{e}_sp"]) >= 0
        assert float(r["bound_vendor"]) <= 1.0
        assert int(r["reps"]) >= 3
        assert abs(float(r["split3_vs_plain"]) - float(r["split3_tf"]) / float(r["plain_tf"])) < 1e-2
