"""Test infrastructure only: FP64 numpy restatement of the MisesMat
surrogate network (13 -> 3 x 128 SiLU -> 27) and of its mean-squared-error
gradient, used to check the HIP training kernels.  The network has no
counterpart in the reference; its parameters use the flat layout of
hfem_mlp (W1 128x13, b1, W2 128x128, b2, W3 128x128, b3, W4 27x128, b4,
row = output unit).  Pinned by central finite differences in
tests/test_oracle.py.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 13, 128, 27
SHAPES = [(NH, NIN), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NOUT, NH), (NOUT,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)


def unpack(p):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def forward(p, X, keep=False):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ W1.T + b1
    h1 = silu(z1)
    z2 = h1 @ W2.T + b2
    h2 = silu(z2)
    z3 = h2 @ W3.T + b3
    h3 = silu(z3)
    y = h3 @ W4.T + b4
    if keep:
        return y, (X, z1, h1, z2, h2, z3, h3)
    return y


def loss_and_grad(p, X, Y):
    """Mean over all rows and outputs of (y - Y)^2 and its gradient with
    respect to the flat parameters."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    y, (X, z1, h1, z2, h2, z3, h3) = forward(p, X, keep=True)
    d = y - np.asarray(Y, dtype=np.float64)
    n = d.size
    loss = float((d * d).sum() / n)
    dy = 2.0 * d / n

    def dsilu(z):
        s = 1.0 / (1.0 + np.exp(-z))
        return s * (1.0 + z * (1.0 - s))

    gW4 = dy.T @ h3
    gb4 = dy.sum(0)
    dz3 = (dy @ W4) * dsilu(z3)
    gW3 = dz3.T @ h2
    gb3 = dz3.sum(0)
    dz2 = (dz3 @ W3) * dsilu(z2)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * dsilu(z1)
    gW1 = dz1.T @ X
    gb1 = dz1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, g
