"""CPU FP64 oracle of the learned part (an extension; the reference has no
learned component).  TEST INFRASTRUCTURE ONLY: imported by tests/ as the
checker of the library's HIP MLP kernels, never by the product.

The network (per vertex):
    xh = (x - mu) / sigma                 x: F = 8 features
    h1 = silu(W1 xh + b1)                 W1: 64 x 8
    h2 = silu(W2 h1 + b2)                 W2: 64 x 64
    y  = sigmoid(w3 . h2 + b3)            w3: 64
    loss = mean((y - t)^2)
with W1, W2, w3 rounded to BF16 before use (the library stores BF16 weights
and accumulates in FP32; here the same rounded values in FP64).
silu(z) = z sigmoid(z).

Parameter vector layout (shared with the library): W1 (64x8, row-major),
b1 (64), W2 (64x64), b2 (64), w3 (64), b3 (1): 4801 values.
Pinned by central finite differences in FP64 (tests/test_mlp.py)."""
import numpy as np

F, H = 8, 64
NPARAM = H * F + H + H * H + H + H + 1


def bf16(a):
    """Round FP32 values to BF16 (round to nearest even), returned as FP64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, x, mu, sd, round_weights=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    if round_weights:
        W1, W2, w3 = bf16(W1), bf16(W2), bf16(w3)
    xh = (np.asarray(x, np.float64) - mu) / sd
    z1 = xh @ W1.T + b1
    h1 = z1 * sig(z1)
    z2 = h1 @ W2.T + b2
    h2 = z2 * sig(z2)
    z3 = h2 @ w3 + b3
    return sig(z3), (xh, z1, h1, z2, h2, z3)


def loss_and_grad(p, x, t, mu, sd, round_weights=True):
    """Loss and its gradient with respect to the (rounded) weights."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    if round_weights:
        W1, W2, w3 = bf16(W1), bf16(W2), bf16(w3)
    y, (xh, z1, h1, z2, h2, z3) = forward(p, x, mu, sd, round_weights)
    n = len(t)
    loss = np.mean((y - t) ** 2)
    dz3 = 2.0 * (y - t) / n * y * (1 - y)
    dw3 = h2.T @ dz3
    db3 = dz3.sum()
    dh2 = np.outer(dz3, w3)
    s2 = sig(z2)
    dz2 = dh2 * s2 * (1 + z2 * (1 - s2))
    dW2 = dz2.T @ h1
    db2 = dz2.sum(0)
    dh1 = dz2 @ W2
    s1 = sig(z1)
    dz1 = dh1 * s1 * (1 + z1 * (1 - s1))
    dW1 = dz1.T @ xh
    db1 = dz1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])
    return loss, g


def init_params(seed):
    """He-style initial weights (the library's tr_mlp_create uses its own RNG;
    tests pass explicit parameters to both)."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((H, F)) * np.sqrt(1.0 / F)
    W2 = rng.standard_normal((H, H)) * np.sqrt(1.0 / H)
    w3 = rng.standard_normal(H) * np.sqrt(1.0 / H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, [0.0]])
