"""CPU reference of the proposal network (test infrastructure).

The network is 12 -> 64 -> 64 -> 16 with ReLU, parameters packed as in include/gstar4d.h
(W1[32][64] with rows 12..31 zero, b1[64], W2[64][64], b2[64], W3[64][16], b3[16]).  Loss per
sample: binary cross-entropy with logits over the slots whose candidate exists (label bit
8 + j) with target bit j, plus the softmax cross-entropy of outputs 8..15 against the bin in
bits 16..18.

`forward_backward(..., bf16=True)` mirrors the GPU's arithmetic: the matrix-product operands
(features, weights, hidden activations and back-propagated output gradients) are rounded to
BF16 (round to nearest even) and every product is accumulated in FP32; with bf16=False the
same computation runs in float64 (for finite differences).
"""
import numpy as np

IN, H, OUT, F, SLOTS, BINS = 32, 64, 16, 12, 8, 8
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * OUT
NPARAM = OFF_B3 + OUT


def unpack(p):
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(IN, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(H, OUT), p[OFF_B3:NPARAM])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3]).astype(np.float32)


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    W1 = np.zeros((IN, H), np.float32)
    W1[:F] = rng.normal(0, np.sqrt(2.0 / F), size=(F, H))
    W2 = rng.normal(0, np.sqrt(2.0 / H), size=(H, H)).astype(np.float32)
    W3 = rng.normal(0, np.sqrt(1.0 / H), size=(H, OUT)).astype(np.float32)
    b1 = rng.normal(0, 0.1, H).astype(np.float32)
    b2 = rng.normal(0, 0.1, H).astype(np.float32)
    b3 = rng.normal(0, 0.1, OUT).astype(np.float32)
    return pack(W1, b1, W2, b2, W3, b3)


def random_samples(rng, n):
    x = rng.normal(0, 1.5, size=(n, F)).astype(np.float32)
    valid = rng.random((n, SLOTS)) < 0.85
    lab = (rng.random((n, SLOTS)) < 0.6) & valid
    bins = rng.integers(0, BINS, n)
    y = np.zeros(n, np.uint32)
    for j in range(SLOTS):
        y |= lab[:, j].astype(np.uint32) << j
        y |= valid[:, j].astype(np.uint32) << (8 + j)
    y |= bins.astype(np.uint32) << 16
    return x, y


def _loss_grad(z, y):
    """Summed loss and dL/dz for logits z (n, 16) and label words y."""
    n = z.shape[0]
    j = np.arange(SLOTS)
    valid = ((y[:, None] >> (8 + j)) & 1).astype(z.dtype)
    t = ((y[:, None] >> j) & 1).astype(z.dtype)
    s = z[:, :SLOTS]
    bce = np.maximum(s, 0) - s * t + np.log1p(np.exp(-np.abs(s)))
    loss = (bce * valid).sum()
    dz = np.zeros_like(z)
    dz[:, :SLOTS] = (1.0 / (1.0 + np.exp(-s)) - t) * valid
    c = z[:, SLOTS:]
    mx = c.max(1, keepdims=True)
    e = np.exp(c - mx)
    se = e.sum(1, keepdims=True)
    bins = ((y >> 16) & 7).astype(np.int64)
    loss += (mx[:, 0] + np.log(se[:, 0]) - c[np.arange(n), bins]).sum()
    sm = e / se
    sm[np.arange(n), bins] -= 1.0
    dz[:, SLOTS:] = sm
    return loss, dz


def forward(p, x, bf16_ops=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    dt = np.float32 if bf16_ops else np.float64
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    X = np.zeros((x.shape[0], IN), dt)
    X[:, :F] = x
    Xb, W1b, W2b, W3b = r(X), r(W1), r(W2), r(W3)
    h1 = np.maximum(Xb @ W1b + b1.astype(dt), 0)
    h1b = r(h1)
    h2 = np.maximum(h1b @ W2b + b2.astype(dt), 0)
    h2b = r(h2)
    z = h2b @ W3b + b3.astype(dt)
    return z, (Xb, h1, h1b, h2, h2b, W1b, W2b, W3b)


def forward_backward(p, x, y, bf16_ops=True):
    """Returns (summed loss, gradient (NPARAM,), logits)."""
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    z, (Xb, h1, h1b, h2, h2b, W1b, W2b, W3b) = forward(p, x, bf16_ops)
    loss, dz3 = _loss_grad(z, y)
    dz3b = r(dz3)
    gW3 = h2b.T @ dz3b
    gb3 = dz3b.sum(0)
    dz2 = (dz3b @ W3b.T) * (h2 > 0)
    dz2b = r(dz2)
    gW2 = h1b.T @ dz2b
    gb2 = dz2b.sum(0)
    dz1 = (dz2b @ W2b.T) * (h1 > 0)
    dz1b = r(dz1)
    gW1 = Xb.T @ dz1b
    gW1[F:] = 0
    gb1 = dz1b.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return float(loss), g, z
