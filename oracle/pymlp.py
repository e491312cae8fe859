"""FP64 oracle of the warm-start network (TEST INFRASTRUCTURE ONLY: tests/ use it as the checker
of the device kernels, never as part of the product path).

Forward and backward of the three-layer perceptron
    z1 = u W1^T + b1, z2 = relu(z1) W2^T + b2, y = relu(z2) W3^T + b3,
    L = (1 / (B d_out)) sum (y - t)^2,
in float64.  With bf16=True every operand of a product (inputs, weights and the activations fed
to the next layer) is first rounded to bfloat16 (round to nearest even), the rounding the
device's MFMA products apply; the products themselves stay exact, so the remaining difference
to the device is its FP32 accumulation.  The parameter layout is the library's: W1 (h1 x d_in),
b1, W2 (h2 x h1), b2, W3 (d_out x h2), b3, concatenated.
"""
import numpy as np


def bf16(a):
    """float32 values rounded to the nearest bfloat16 (ties to even), returned as float64."""
    x = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((x + 0x7FFF + ((x >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def split(p, dims):
    d_in, h1, h2, d_out = dims
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = []
    for shape in ((h1, d_in), (h1,), (h2, h1), (h2,), (d_out, h2), (d_out,)):
        k = int(np.prod(shape))
        out.append(p[o:o + k].reshape(shape))
        o += k
    assert o == p.size
    return out


def forward(p, dims, u, bf=True):
    W1, b1, W2, b2, W3, b3 = split(p, dims)
    q = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    u = np.asarray(u, np.float64)
    z1 = q(u) @ q(W1).T + b1
    z2 = q(np.maximum(z1, 0)) @ q(W2).T + b2
    y = q(np.maximum(z2, 0)) @ q(W3).T + b3
    return z1, z2, y


def gradients(p, dims, u, t, bf=True):
    """(gradient in the parameter layout, loss).  Activations are rounded to float32 where the
    device stores them before they enter the next product."""
    W1, b1, W2, b2, W3, b3 = split(p, dims)
    q = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    f32 = (lambda a: np.asarray(a, np.float32)) if bf else (lambda a: a)
    u = np.asarray(u, np.float64)
    z1, z2, y = forward(p, dims, u, bf)
    if bf:
        z1, z2, y = (np.asarray(np.asarray(a, np.float32), np.float64) for a in (z1, z2, y))
    B, d_out = y.shape
    d = y - np.asarray(t, np.float64)
    loss = float((d * d).sum() / (B * d_out))
    g = f32(2.0 * d / (B * d_out))
    a2 = f32(np.maximum(z2, 0))
    a1 = f32(np.maximum(z1, 0))
    dW3 = q(g).T @ q(a2)
    db3 = np.asarray(g, np.float64).sum(0)
    g2 = f32((q(g) @ q(W3)) * (z2 > 0))
    dW2 = q(g2).T @ q(a1)
    db2 = np.asarray(g2, np.float64).sum(0)
    g1 = f32((q(g2) @ q(W2)) * (z1 > 0))
    dW1 = q(g1).T @ q(u)
    db1 = np.asarray(g1, np.float64).sum(0)
    grad = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return grad, loss
