"""The MLP oracle pinned without the GPU: bfloat16 rounding known answers, backward against central
finite differences of the loss (float64, no rounding), Adam against its closed form."""
import numpy as np

from oracle import mlp_ref as R


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 3.0e-39, 65504.0],
                 dtype=np.float32)
    want = np.array([1.0, 1.0, 1.015625, 1.0078125, -2.5, np.float32(3.0e-39), 65536.0], dtype=np.float32)
    got = R.bf16(x)
    # 3e-39 is subnormal: the rounding keeps the top 16 bits of its pattern
    want[5] = np.array([np.float32(3.0e-39).view(np.uint32) + 0x7FFF + ((np.float32(3.0e-39).view(np.uint32) >> 16) & 1)
                        & 0xFFFF0000], dtype=np.uint32).view(np.float32)[0]
    assert np.array_equal(got, want)
    assert (R.bf16(x).view(np.uint32) & 0xFFFF == 0).all()


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    i, h, o, N = 5, 7, 4, 6
    theta = rng.standard_normal(i * h + h + h * h + h + h * o + o) * 0.5
    X = rng.standard_normal((N, i))
    T = rng.standard_normal((N, o))
    loss, g = R.backward(theta, X, T, i, h, o, rounded=False)
    eps = 1e-6
    for k in rng.choice(theta.size, 40, replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (R.backward(tp, X, T, i, h, o, rounded=False)[0] - R.backward(tm, X, T, i, h, o, rounded=False)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_adam_closed_form():
    theta, g = np.array([1.0, -2.0]), np.array([0.5, -0.25])
    t1, m, v = R.adam(theta, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    # first step: m_hat = g, v_hat = g^2, so the step is lr * sign(g) (up to eps)
    assert np.allclose(t1, theta - 0.1 * np.sign(g), atol=1e-7)
