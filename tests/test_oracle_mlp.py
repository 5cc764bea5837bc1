"""CPU: pins for the O2 MLP oracle (oracle/mlp.py).

* its backward pass equals central finite differences of its forward pass
  (FP64, operands unrounded) for every weight and bias entry sampled;
* GELU and its derivative match known values and a numerical derivative;
* BF16 rounding is round-to-nearest-even (ties, carries, specials);
* one Adam step from zero moments moves each weight by lr * sign(g) (to
  eps), and the MSE gradient is the derivative of the MSE.
"""

import numpy as np
import pytest

from oracle import mlp as O2


def _net(sizes, seed):
    rng = np.random.default_rng(seed)
    return [(rng.normal(0, 0.7, (o, i)), rng.normal(0, 0.3, o)) for i, o in zip(sizes, sizes[1:])]


def test_backward_matches_finite_differences():
    params = _net([3, 7, 5, 4], 1)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(6, 3))
    T = rng.normal(size=(6, 4))
    w = rng.uniform(0.5, 2.0, 4)

    def loss(ps):
        Y, _ = O2.forward(ps, X, bf16_ops=False)
        return O2.mse(Y, T, w)[0]

    Y, cache = O2.forward(params, X, bf16_ops=False)
    _, dY = O2.mse(Y, T, w)
    grads = O2.backward(params, cache, dY, bf16_ops=False)
    h = 1e-6
    for li, (W, b) in enumerate(params):
        for (r, c) in [(0, 0), (W.shape[0] - 1, W.shape[1] - 1), (1, min(2, W.shape[1] - 1))]:
            Wp = [(w_.copy(), b_.copy()) for w_, b_ in params]
            Wm = [(w_.copy(), b_.copy()) for w_, b_ in params]
            Wp[li][0][r, c] += h
            Wm[li][0][r, c] -= h
            fd = (loss(Wp) - loss(Wm)) / (2 * h)
            assert abs(fd - grads[li][0][r, c]) <= 1e-7 + 1e-5 * abs(fd), (li, r, c)
        bp = [(w_.copy(), b_.copy()) for w_, b_ in params]
        bm = [(w_.copy(), b_.copy()) for w_, b_ in params]
        bp[li][1][0] += h
        bm[li][1][0] -= h
        fd = (loss(bp) - loss(bm)) / (2 * h)
        assert abs(fd - grads[li][1][0]) <= 1e-7 + 1e-5 * abs(fd), li


def test_gelu_values_and_derivative():
    assert O2.gelu(1.0) == pytest.approx(0.8413447460685429, rel=1e-15)
    assert O2.gelu(-1.0) == pytest.approx(-0.15865525393145707, rel=1e-14)
    assert O2.gelu(0.0) == 0.0
    z = np.linspace(-5, 5, 41)
    h = 1e-6
    num = (O2.gelu(z + h) - O2.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(O2.dgelu(z), num, atol=1e-8)


def test_bf16_rounding():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)          # bf16 spacing at 1
    assert O2.bf16(one) == 1.0
    assert O2.bf16(one + ulp / 2) == 1.0                 # tie -> even (1.0)
    assert O2.bf16(one + 3 * ulp / 2) == 1.0 + 2 * ulp   # tie -> even (1 + 2 ulp)
    assert O2.bf16(one + ulp * 0.51) == 1.0 + ulp
    assert O2.bf16(np.float32(np.inf)) == np.inf
    assert np.isnan(O2.bf16(np.float32(np.nan)))
    x = np.random.default_rng(3).normal(size=1000).astype(np.float32)
    r = O2.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


def test_adam_first_step_and_mse_gradient():
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = O2.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 0.01, 0.9, 0.999, 1e-12, 0.0, 1)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-9)
    Y = np.array([[1.0, 2.0], [0.5, -1.0]])
    T = np.zeros_like(Y)
    l, d = O2.mse(Y, T)
    assert l == pytest.approx(np.mean(Y ** 2))
    np.testing.assert_allclose(d, 2 * Y / Y.size)
