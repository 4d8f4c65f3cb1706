"""The learned warm start's MLP kernels (bf16 matrix cores, float32 accumulation) against the CPU
oracle (oracle/mlp_ref.py, pinned by finite differences): the GEMM tile, forward, loss, every
gradient, the Adam step, and training on the GPU reducing the loss."""
import numpy as np
import pytest

from oracle import mlp_ref as R
from solver import _lib
from solver.learned import MLP

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K", [(16, 16, 16), (1, 1, 1), (33, 301, 77), (301, 512, 256), (64, 17, 512)])
def test_gemm_tile_matches_bf16_products(M, N, K):
    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C = np.empty((M, N), dtype=np.float32)
    ctx = _lib.context(0)
    _lib.check(ctx.lib.rumgpu_mlp_gemm_test(ctx.h, M, N, K, _lib.ptr(A), _lib.ptr(B), _lib.ptr(C)))
    want = R.gemm(A, B)
    scale = np.abs(R.bf16(A)).astype(np.float64) @ np.abs(R.bf16(B)).astype(np.float64)
    assert (np.abs(C - want) <= 1e-6 * scale + 1e-30).all(), np.abs(C - want).max()


def _net(i=301, h=512, o=301, seed=3):
    net = MLP(i, h, o, max_batch=128, seed=seed)
    theta = net.get_params()
    # non-zero biases so every term of the backward is exercised
    theta[i * h:i * h + h] = 0.05
    k = i * h + h + h * h
    theta[k:k + h] = -0.02
    theta[-o:] = 0.1
    net.set_params(theta)
    return net, theta


def test_forward_backward_adam_match_oracle():
    i, h, o, N = 301, 512, 301, 96
    net, theta = _net(i, h, o)
    rng = np.random.default_rng(11)
    X = rng.standard_normal((N, i)).astype(np.float32)
    T = rng.standard_normal((N, o)).astype(np.float32)
    Y = net.forward(X)
    _h1, _h2, Yo = R.forward(theta, X, i, h, o)
    assert np.abs(Y - Yo).max() <= 1e-4 * max(1.0, np.abs(Yo).max())
    loss, g = net.backward(X, T, want_grad=True)
    lo, go = R.backward(theta, X, T, i, h, o)
    assert abs(loss - lo) <= 1e-5 * lo
    # a ReLU whose pre-activation sits within float32 accumulation error of 0 may switch; compare
    # per parameter block against its scale
    bounds = np.cumsum([0, i * h, h, h * h, h, h * o, o])
    for a, b in zip(bounds[:-1], bounds[1:]):
        sc = np.abs(go[a:b]).max()
        assert np.abs(g[a:b] - go[a:b]).max() <= 2e-3 * sc, (a, b, np.abs(g[a:b] - go[a:b]).max(), sc)
    net.adam(1e-3)
    t1 = net.get_params()
    to, _m, _v = R.adam(theta.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    assert np.abs(t1 - to).max() <= 1e-6
    net.close()


def test_training_on_the_gpu_reduces_the_loss():
    i, h, o, N = 301, 512, 301, 256
    rng = np.random.default_rng(5)
    A = (rng.standard_normal((i, o)) / np.sqrt(i)).astype(np.float32)
    X = rng.standard_normal((N, i)).astype(np.float32)
    T = np.tanh(X @ A).astype(np.float32)
    net = MLP(i, h, o, max_batch=128, seed=1)
    hist = net.fit(X, T, epochs=150, batch=64, lr=1e-3)
    print("loss %.4f -> %.4f" % (hist[0], hist[-1]))
    assert hist[-1] < 0.1 * hist[0]
    net.close()
