"""CPU oracle of the learned warm start's MLP (test infrastructure only: the tests import it as the
checker of the GPU kernels in mlp.hip; nothing in the product path calls it).

The network: Y = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3 per row (row-major X @ W), loss
(1/N) sum_i 1/2 ||Y_i - T_i||^2, Adam on the float32 parameters. GEMM operands are rounded to
bfloat16 (round to nearest, ties to even) as the matrix cores take them; products are exact and
accumulated here in float64 (the GPU accumulates in float32, so the two agree to float32
accumulation error). No reference counterpart exists; the formulas are pinned by finite
differences (float64, no rounding) and by bfloat16 rounding known answers in tests/test_mlp_oracle.py.
"""
import numpy as np


def bf16(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32"""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def gemm(A, B, rounded=True):
    """bf16 operands, float64 sums, float32 result (rounded); plain float64 product otherwise"""
    if not rounded:
        return np.asarray(A, dtype=np.float64) @ np.asarray(B, dtype=np.float64)
    return (bf16(A).astype(np.float64) @ bf16(B).astype(np.float64)).astype(np.float32)


def split(theta, i, h, o):
    k = 0
    out = []
    for shape in ((i, h), (h,), (h, h), (h,), (h, o), (o,)):
        size = int(np.prod(shape))
        out.append(theta[k:k + size].reshape(shape))
        k += size
    return out


def forward(theta, X, i, h, o, rounded=True):
    W1, b1, W2, b2, W3, b3 = split(theta, i, h, o)
    H1 = np.maximum(gemm(X, W1, rounded) + b1, 0).astype(X.dtype)
    H2 = np.maximum(gemm(H1, W2, rounded) + b2, 0).astype(X.dtype)
    Y = (gemm(H2, W3, rounded) + b3).astype(X.dtype)
    return H1, H2, Y


def backward(theta, X, T, i, h, o, rounded=True):
    """(loss, grad) with the same factorisation as the GPU kernels"""
    W1, b1, W2, b2, W3, b3 = split(theta, i, h, o)
    H1, H2, Y = forward(theta, X, i, h, o, rounded)
    N = X.shape[0]
    D = (Y - T)
    loss = 0.5 * float((D.astype(np.float64) ** 2).sum()) / N
    dY = (D / N).astype(X.dtype)
    gW3 = gemm(H2.T, dY, rounded)
    gb3 = dY.sum(axis=0)
    dA2 = np.where(H2 > 0, gemm(dY, W3.T, rounded), 0).astype(X.dtype)
    gW2 = gemm(H1.T, dA2, rounded)
    gb2 = dA2.sum(axis=0)
    dA1 = np.where(H1 > 0, gemm(dA2, W2.T, rounded), 0).astype(X.dtype)
    gW1 = gemm(X.T, dA1, rounded)
    gb1 = dA1.sum(axis=0)
    grad = np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, grad


def adam(theta, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """one bias-corrected Adam step (Kingma and Ba, Algorithm 1); returns new (theta, m, v)"""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return theta - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
