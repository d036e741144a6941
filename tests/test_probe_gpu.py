"""The register-only MFMA probe (maxmulsk_probe_mfma): rates are positive,
below the dense peaks, and random operand bits never run faster than constant
ones by more than run-to-run noise; bad arguments are refused. A single short
run can land in a low clock state, so each rate is the best of three runs of
a few milliseconds."""
import pytest

import maxmulsk_amd as mm
from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu


def test_probe_rates_are_plausible():
    torch_cuda()
    for mode, peak, iters in ((1, 2516.6, 20000), (0, 157.3, 10000)):
        rnd = max(mm.probe_mfma(mode, True, iters) for _ in range(3))
        const = max(mm.probe_mfma(mode, False, iters) for _ in range(3))
        assert 0.2 * peak < rnd <= 1.02 * peak, (mode, rnd)
        assert 0.2 * peak < const <= 1.02 * peak, (mode, const)
        assert rnd <= const * 1.15, (mode, rnd, const)


def test_probe_rejects_bad_arguments():
    torch_cuda()
    with pytest.raises(mm.MaxMulSKError):
        mm.probe_mfma(2, True, 100)
    with pytest.raises(mm.MaxMulSKError):
        mm.probe_mfma(1, True, 0)


def test_stream_probe_orders_its_floors():
    """maxmulsk_probe_stream on the narrow-M plans: an empty kernel on the grid
    costs less than streaming the operands on it, which costs less than
    streaming plus the split-K reduce; without a split the two streaming
    figures agree; bad plans are rejected before anything runs."""
    torch_cuda()
    for M, N, K, plan in ((64, 512, 8192, (64, 32)), (128, 512, 32768, (64, 32)), (64, 24576, 1536, (64, 2))):
        f = mm.probe_stream(M, K, N, plan, calls=200)
        assert 0 < f["empty_us"] < f["stream_us"] < f["stream_reduce_us"] < 1000, f
    f1 = mm.probe_stream(256, 256, 256, (64, 1), calls=50)
    assert f1["stream_reduce_us"] == pytest.approx(f1["stream_us"], rel=0.5)
    for bad in ((256, 256, 256, (256, 1)), (64, 255, 512, (64, 1)), (64, 512, 255, (64, 1)), (64, 512, 512, (48, 1))):
        M, K, N, plan = bad
        with pytest.raises(mm.MaxMulSKError):
            mm.probe_stream(M, K, N, plan)
