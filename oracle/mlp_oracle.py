"""FP64 restatement of the warm-start network's forward and backward pass.

TEST INFRASTRUCTURE ONLY (the checker of the GPU network). It rounds operands
to bf16 at the same points as the GPU path (GEMM inputs: X, W, H1, H2, dY,
dH2, dH1), then computes everything else in float64, so the GPU gradients can
be compared with a tolerance set by FP32 accumulation only.
"""
import numpy as np


def bf16(x, enabled=True):
    """Round float32 values to bf16 (round to nearest even), returned as float64.
    With enabled=False the values pass through in float64 (for finite-difference
    checks of the backward pass itself)."""
    if not enabled:
        return np.asarray(x, np.float64)
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, n_in, h, n_out):
    p = np.asarray(params, np.float64)
    sizes = [n_in * h, h, h * h, h, h * n_out, n_out]
    out, o = [], 0
    for s in sizes:
        out.append(p[o:o + s])
        o += s
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(n_in, h), b1, W2.reshape(h, h), b2, W3.reshape(h, n_out), b3


def loss_and_grads(params, x, t, n_in, h, n_out, rounding=True):
    W1, b1, W2, b2, W3, b3 = unpack(params, n_in, h, n_out)

    def rnd(v):  # the rounding points of the GPU path
        return bf16(v, rounding)

    W1r, W2r, W3r = rnd(W1), rnd(W2), rnd(W3)
    n = x.shape[0]
    xr = rnd(x)
    h1 = np.maximum(xr @ W1r + b1, 0.0)
    h1r = rnd(h1)
    h2 = np.maximum(h1r @ W2r + b2, 0.0)
    h2r = rnd(h2)
    y = h2r @ W3r + b3
    e = y - np.asarray(t, np.float64)
    loss = np.mean(e * e)
    dy = 2.0 * e / (n * n_out)
    dyr = rnd(dy)
    gW3 = h2r.T @ dyr
    gb3 = dy.sum(0)
    dh2 = (dyr @ W3r.T) * (h2 > 0)
    dh2r = rnd(dh2)
    gW2 = h1r.T @ dh2r
    gb2 = dh2.sum(0)
    dh1 = (dh2r @ W2r.T) * (h1 > 0)
    dh1r = rnd(dh1)
    gW1 = xr.T @ dh1r
    gb1 = dh1.sum(0)
    grads = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grads, y
