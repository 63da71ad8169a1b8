"""CPU restatement of the learned planner's cost model (include/vsgm_mlp.h).

TEST INFRASTRUCTURE ONLY: the product's model runs in the HIP kernels of
vsgm_amd/csrc/hip/mlp.hip; this numpy version is the checker its forward
pass, gradients and Adam step are compared with. There is no reference
counterpart (the reference planner is heuristic only), so this file is pinned
by finite differences (tests/test_oracle_mlp.py) and by the textbook
definition of round-to-nearest-even BF16.

Model: y = w3q . relu(W2q relu(W1q x + b1) + b2) + b3 with Wq = bf16(W)
(the gradient flows through the rounding unchanged); loss = mean (y-t)^2/2.
Parameters: W1 [64][D] | b1 [64] | W2 [64][64] | b2 [64] | w3 [64] | b3 [1].

mfma=True restates the matrix-core kernels: the operands of every product
are BF16, so x and h1 = relu(z1) are rounded before the forward products,
and the deltas dh2 = dy w3 [z2 > 0] and dh1 = (bf16(dh2) W2) [h1 > 0]
before the backward ones (bias and w3 gradients, h2 and y stay FP32).
"""
import numpy as np

H = 64


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def n_params(d: int) -> int:
    return H * d + H + H * H + H + H + 1


def unpack(p: np.ndarray, d: int):
    o = 0
    W1 = p[o:o + H * d].reshape(H, d); o += H * d
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p: np.ndarray, x: np.ndarray, quantize: bool = True, dtype=np.float32, mfma: bool = False):
    d = x.shape[1]
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, dtype=np.float32 if quantize else dtype), d)
    q = bf16 if quantize else (lambda a: a)
    W1, W2, w3 = (q(W1).astype(dtype), q(W2).astype(dtype), q(w3).astype(dtype))
    b1, b2, b3 = b1.astype(dtype), b2.astype(dtype), dtype(b3)
    x = bf16(x) if mfma else x.astype(dtype)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0)
    if mfma:
        h1 = bf16(h1)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    return y, (x, z1, h1, z2, h2, W1, W2, w3)


def loss_and_grad(p: np.ndarray, x: np.ndarray, t: np.ndarray, quantize: bool = True, dtype=np.float32, mfma: bool = False):
    y, (xx, z1, h1, z2, h2, W1, W2, w3) = forward(p, x, quantize, dtype, mfma)
    n = len(x)
    err = y - t.astype(dtype)
    loss = 0.5 * np.sum(err * err) / n
    dy = err / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dh2 = np.outer(dy, w3) * (z2 > 0)
    gb2 = dh2.sum(0)
    if mfma:
        dh2 = bf16(dh2)
    gW2 = dh2.T @ h1
    dh1 = (dh2 @ W2) * (z1 > 0)
    gb1 = dh1.sum(0)
    gW1 = (bf16(dh1) if mfma else dh1).T @ xx
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]]).astype(dtype)
    return float(loss), g


def adam(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns new (p, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
