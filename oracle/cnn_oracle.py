"""CPU oracle of the warm-start CNN (forward, mean-squared-error loss and backward), in numpy FP64.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP forward and backward kernels.  The network has
no reference counterpart; its definition is the one stated in include/pg_amd.h (pgamd_cnn_*):

    x0 [n,18,18,18,8] -> y1 = relu(conv(x0)+b1)                       (valid 3^3 conv, 18 -> 16)
    y_l = crop(y_{l-1}) + relu(conv(y_{l-1}) + b_l),  l = 2..5        (16 -> 8)
    d = conv_same(y5) + b6                                            (8^3, 1 channel)

Parameters are the flat vector of pgamd_cnn_get_params (W1 b1 W2 b2 W3 b3 W4 b4 W5 b5 w6 b6) with
W_l[co][tap*CIN + ci], tap = dz*9 + dy*3 + dx.  With bf16=True the activations, the BF16 weight copies
and the ReLU gradients are rounded to BF16 at the same points as the GPU kernels (to compare closely);
with bf16=False everything is FP64 (used for the finite-difference pin of the backward pass).
"""
from __future__ import annotations

import numpy as np
from numpy.lib.stride_tricks import sliding_window_view

NPARAMS = 31617
OW = {1: 0, 2: 3472, 3: 10400, 4: 17328, 5: 24256}
OB = {1: 3456, 2: 10384, 3: 17312, 4: 24240, 5: 31168}
OW6, OB6 = 31184, 31616
SIN = {1: 18, 2: 16, 3: 14, 4: 12, 5: 10}


def bf16_round(a):
    """Round FP32 values to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits_to_float(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    W = {1: p[OW[1]:OW[1] + 16 * 216].reshape(16, 216)}
    for l in range(2, 6):
        W[l] = p[OW[l]:OW[l] + 16 * 432].reshape(16, 432)
    b = {l: p[OB[l]:OB[l] + 16] for l in range(1, 6)}
    return W, b, p[OW6:OW6 + 432], p[OB6]


def _windows(x):
    """[n,S,S,S,c] -> [n,So,So,So,27*c] with K ordered tap-major (tap*c + ci)."""
    n, S, _, _, c = x.shape
    w = sliding_window_view(x, (3, 3, 3), axis=(1, 2, 3))          # [n,So,So,So,c,3,3,3]
    w = np.moveaxis(w, 4, 7)                                         # [n,So,So,So,3,3,3,c]
    return w.reshape(n, S - 2, S - 2, S - 2, 27 * c)


def _scatter_windows(g, S, c):
    """Adjoint of _windows: [n,So,So,So,27*c] -> [n,S,S,S,c]."""
    n, So = g.shape[0], S - 2
    g = g.reshape(n, So, So, So, 3, 3, 3, c)
    out = np.zeros((n, S, S, S, c))
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                out[:, dz:dz + So, dy:dy + So, dx:dx + So, :] += g[:, :, :, :, dz, dy, dx, :]
    return out


def forward(p, x0, bf16=True):
    """Returns (d [n,512], cache)."""
    W, b, w6, b6 = unpack(p)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    Wq = {l: r(W[l]) for l in W}
    w6q = r(w6)
    x = np.asarray(x0, np.float64)
    cache = {"x": {1: x}, "h": {}, "win": {}}
    for l in range(1, 6):
        win = _windows(x)
        h = win @ Wq[l].T + b[l]
        y = np.maximum(h, 0.0)
        if l > 1:
            y = y + x[:, 1:-1, 1:-1, 1:-1, :]
        if bf16:
            h = r(h)
            y = r(y)
        cache["win"][l], cache["h"][l] = win, h
        x = y
        cache["x"][l + 1] = x
    xp = np.zeros((x.shape[0], 10, 10, 10, 16))
    xp[:, 1:9, 1:9, 1:9, :] = x
    win6 = _windows(xp)                                               # [n,8,8,8,432]
    d = win6 @ w6q + b6
    cache["win6"] = win6
    cache["Wq"], cache["w6q"] = Wq, w6q
    return d.reshape(-1, 512), cache


def loss_grad(p, x0, target, bf16=True):
    """Mean-squared-error loss and its gradient with respect to the parameter vector."""
    d, c = forward(p, x0, bf16)
    n = d.shape[0]
    t = np.asarray(target, np.float64).reshape(n, 512)
    e = d - t
    loss = float(np.mean(e * e))
    dd = (2.0 * e / e.size).reshape(n, 8, 8, 8)
    g = np.zeros(NPARAMS)
    g[OW6:OW6 + 432] = np.einsum("nzyx,nzyxk->k", dd, c["win6"])
    g[OB6] = dd.sum()
    dwin6 = dd[..., None] * c["w6q"]                                  # [n,8,8,8,432]
    dy = _scatter_windows(dwin6, 10, 16)[:, 1:9, 1:9, 1:9, :]
    r = bf16_round if bf16 else (lambda a: a)
    for l in range(5, 0, -1):
        dh = np.where(c["h"][l] > 0, dy, 0.0)
        if bf16:
            dh = r(dh)
        g[OB[l]:OB[l] + 16] = dh.sum(axis=(0, 1, 2, 3))
        cin = 8 if l == 1 else 16
        g[OW[l]:OW[l] + 16 * 27 * cin] = np.einsum("nzyxc,nzyxk->ck", dh, c["win"][l]).ravel()
        if l > 1:
            dx = _scatter_windows(dh @ c["Wq"][l], SIN[l], 16)
            dx[:, 1:-1, 1:-1, 1:-1, :] += dy
            dy = dx
    return loss, g
