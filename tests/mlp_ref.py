"""CPU FP64 reference of the surrogate MLP (test infrastructure): the same
architecture and loss as the HIP kernels, forward and backward in numpy."""

import numpy as np


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(W, b, x):
    """W[l]: [N][K], b[l]: [N]; hidden layers SiLU, last linear."""
    acts, zs = [x], []
    a = x
    for l, (w, bb) in enumerate(zip(W, b)):
        z = a @ w.T + bb
        zs.append(z)
        a = silu(z) if l < len(W) - 1 else z
        acts.append(a)
    return acts, zs


def loss_and_grads(W, b, x, y, n, nout):
    """Mean squared error over the first n rows and nout outputs; gradients
    of every W and b."""
    acts, zs = forward(W, b, x)
    out = acts[-1]
    cnt = n * nout
    d = np.zeros_like(out)
    d[:n, :nout] = out[:n, :nout] - y[:n, :nout]
    loss = (d ** 2).sum() / cnt
    dz = 2.0 * d / cnt
    gW, gb = [None] * len(W), [None] * len(W)
    for l in range(len(W) - 1, -1, -1):
        gW[l] = dz.T @ acts[l]
        gb[l] = dz.sum(axis=0)
        if l > 0:
            dz = (dz @ W[l]) * dsilu(zs[l - 1])
    return loss, gW, gb
