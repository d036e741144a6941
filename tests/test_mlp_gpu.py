"""The selector's GPU kernels against the FP64 MLP oracle: forward, loss,
gradients, Adam, the one-wave inference kernel, training progress, and the
planner tier that uses the network."""
import ctypes as C

import numpy as np
import pytest

import maxmulsk_amd as mm
from maxmulsk_amd import selector as sel
from oracle import mlp_oracle as mo
from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu
NOUT = sel.NCAND  # one output per candidate plan


def data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (n, 4)).astype(np.float32)
    T = rng.normal(-2, 2, (n, NOUT)).astype(np.float32)
    Mk = (rng.uniform(0, 1, (n, NOUT)) > 0.2).astype(np.float32)
    return X, T, Mk


def test_forward_loss_gradient_adam_match_oracle():
    torch = torch_cuda()
    p = mo.init_params(NOUT, 3)
    p[256:320] = 0.05
    X, T, Mk = data(300, 4)
    tr = sel.Trainer(p, 300)
    Xd, Td, Md = (torch.from_numpy(a).cuda() for a in (X, T, Mk))
    Y = tr.forward(Xd)
    torch.cuda.synchronize()
    _, _, Y64 = mo.forward(p, X, NOUT)
    assert np.allclose(Y.cpu().numpy(), Y64, rtol=1e-4, atol=1e-4)
    # one step: gradient and loss of the step, then the Adam update
    mm.check(mm.lib().maxmulsk_mlp_backward(tr.p.data_ptr(), Xd.data_ptr(), 300, tr.H1.data_ptr(),
                                            tr.H2.data_ptr(), tr.Y.data_ptr(), Td.data_ptr(), Md.data_ptr(),
                                            tr.g.data_ptr(), tr.loss.data_ptr(), None))
    torch.cuda.synchronize()
    L64, g64 = mo.backward(p, X, T, Mk, NOUT)
    assert abs(float(tr.loss.item()) - L64) <= 1e-5 * max(1.0, L64)
    g = tr.g.cpu().numpy()
    assert np.max(np.abs(g - g64)) <= 1e-4 * np.max(np.abs(g64)) + 1e-7
    # the same parameters receive no gradient (dead ReLU units) on both sides
    assert np.mean((g != 0) == (g64 != 0)) > 0.99
    mm.check(mm.lib().maxmulsk_mlp_adam(tr.p.data_ptr(), tr.g.data_ptr(), tr.m.data_ptr(), tr.v.data_ptr(), 1,
                                        1e-3, 0.9, 0.999, 1e-8, None))
    torch.cuda.synchronize()
    p64, _, _ = mo.adam(p.astype(np.float64), g.astype(np.float64), np.zeros(p.size), np.zeros(p.size), 1, 1e-3)
    assert np.allclose(tr.p.cpu().numpy(), p64, rtol=0, atol=1e-6)


def test_inference_kernel_and_planner_tier():
    torch_cuda()
    p = mo.init_params(NOUT, 5)
    lib = mm.lib()
    try:
        mm.check(lib.maxmulsk_selector_set(p.ctypes.data))
        assert lib.maxmulsk_selector_active() == 1
        for (M, K, N, mode) in [(64, 8192, 512, 0), (4096, 4096, 4096, 1), (1, 1, 1, 2), (300, 17, 9000, 0)]:
            pred, best = mm.selector_predict(M, K, N, mode)
            x = np.array([mm.selector_features(M, K, N, mode)])
            _, _, y64 = mo.forward(p, x, NOUT)
            assert np.allclose(pred, y64[0], rtol=1e-4, atol=1e-4)
            assert best == mm.candidates()[int(np.argmin(pred))]
            lib.maxmulsk_cache_clear()
            plan, src = mm.choose_plan(M, K, N, mode)
            assert src == 2 and plan == best
            # the autotuner cache outranks the network
            assert lib.maxmulsk_cache_put(M, K, N, mode, mm.Plan(64, 1)) == mm.OK
            assert mm.choose_plan(M, K, N, mode) == ((64, 1), 1)
        lib.maxmulsk_cache_clear()
    finally:
        lib.maxmulsk_selector_set(None)
    assert lib.maxmulsk_selector_active() == 0
    assert mm.choose_plan(64, 8192, 512)[1] == 3


def test_selector_file_round_trip(tmp_path):
    torch_cuda()
    p = mo.init_params(NOUT, 6)
    lib = mm.lib()
    path = str(tmp_path / "sel.bin").encode()
    mm.check(lib.maxmulsk_selector_save(path, p.ctypes.data))
    try:
        mm.check(lib.maxmulsk_selector_load(path))
        pred, _ = mm.selector_predict(100, 200, 300, 1)
        _, _, y64 = mo.forward(p, np.array([mm.selector_features(100, 200, 300, 1)]), NOUT)
        assert np.allclose(pred, y64[0], rtol=1e-4, atol=1e-4)
        raw = bytearray(open(path, "rb").read())
        raw[12] = 5  # wrong geometry
        bad = tmp_path / "bad.bin"
        bad.write_bytes(bytes(raw))
        assert lib.maxmulsk_selector_load(str(bad).encode()) == mm.ERR_UNSUPPORTED
    finally:
        lib.maxmulsk_selector_set(None)


def test_training_reduces_loss():
    """A few hundred steps on a smooth synthetic cost table cut the loss 10x."""
    torch_cuda()
    rng = np.random.default_rng(8)
    X = rng.uniform(0, 1, (2000, 4)).astype(np.float32)
    W = rng.normal(0, 1, (4, NOUT))
    T = (np.sin(3 * X @ W) + X.sum(1, keepdims=True)).astype(np.float32)
    Mk = np.ones_like(T)
    params, hist = sel.train(X, T, Mk, steps=600, batch=512, lr=3e-3, log=lambda s: None)
    L0 = mo.loss(sel.init_params(7), X, T, Mk, NOUT)
    L1 = mo.loss(params, X, T, Mk, NOUT)
    assert L1 < 0.1 * L0, (L0, L1, hist)
