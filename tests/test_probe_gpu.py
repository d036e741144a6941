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
