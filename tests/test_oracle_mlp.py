"""Pins the MLP oracle (CPU): finite differences, a hand-computed tiny case, and
the Adam closed form on a constant gradient."""
import numpy as np

from oracle import mlp_oracle as mo


def test_layout_size():
    assert mo.num_params(15) == 64 * 4 + 64 + 64 * 64 + 64 + 15 * 64 + 15 == 5455


def test_gradient_matches_finite_differences():
    nout = 7
    rng = np.random.default_rng(1)
    p = mo.init_params(nout, 2).astype(np.float64)
    p[64 * 4:64 * 4 + 64] += 0.1  # biases away from the ReLU kink
    X = rng.uniform(0, 1, (9, 4))
    T = rng.normal(0, 1, (9, nout))
    Mk = (rng.uniform(0, 1, (9, nout)) > 0.3).astype(np.float64)
    L, g = mo.backward(p, X, T, Mk, nout)
    assert abs(L - mo.loss(p, X, T, Mk, nout)) < 1e-12
    idx = rng.choice(p.size, 200, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mo.loss(p + e, X, T, Mk, nout) - mo.loss(p - e, X, T, Mk, nout)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_tiny_network_by_hand():
    # One active hidden path: x -> h1[0] = relu(2x) -> h2[0] = relu(3 h1) -> y = 5 h2 + 1
    nout = 1
    p = np.zeros(mo.num_params(nout))
    W1, b1, W2, b2, W3, b3 = mo.unpack(p, nout)  # views into p
    W1[0, 0] = 2.0
    W2[0, 0] = 3.0
    W3[0, 0] = 5.0
    b3[0] = 1.0
    X = np.array([[0.5, 0, 0, 0]])
    _, _, Y = mo.forward(p, X, nout)
    assert Y[0, 0] == 5 * 3 * 2 * 0.5 + 1 == 16.0
    L, g = mo.backward(p, X, np.array([[10.0]]), np.ones((1, 1)), nout)
    assert L == 36.0
    gW1, gb1, gW2, gb2, gW3, gb3 = mo.unpack(g, nout)
    assert gb3[0] == 12.0          # dL/dy = 2 (16 - 10)
    assert gW3[0, 0] == 12.0 * 3.0  # * h2 = 3
    assert gW2[0, 0] == 12.0 * 5.0 * 1.0  # * W3 * h1
    assert gW1[0, 0] == 12.0 * 5.0 * 3.0 * 0.5
    # masked-out target contributes nothing
    L0, g0 = mo.backward(p, X, np.array([[10.0]]), np.zeros((1, 1)), nout)
    assert L0 == 0.0 and not np.any(g0)


def test_adam_constant_gradient_closed_form():
    # with a constant gradient g, bias-corrected m = g and v = g^2 at every step,
    # so each step moves by lr * g / (|g| + eps)
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -4.0])
    m = np.zeros(2)
    v = np.zeros(2)
    for step in range(1, 6):
        p_new, m, v = mo.adam(p, g, m, v, step, lr=0.01, eps=0.0)
        assert np.allclose(p_new - p, -0.01 * np.sign(g), rtol=0, atol=1e-12)
        p = p_new
