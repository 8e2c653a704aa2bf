"""CPU oracle of the warm-start CNN (input packing, forward, loss and backward), numpy.  The loss is the mean
squared error of the correction plus an optional weight times the mean squared residual that the error leaves
in the guess (`loss_grad`, `res_weight`).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP kernels.  The network has no reference
counterpart; its definition is the one stated in include/pg_amd.h (pgamd_cnn_*):

    fields on the 18^3 frame (the solve cube of ghost size 5; smaller cubes centred, zeros around):
      q0 = interp / s,  q1 = 64 * r / s, both on the solve interior only (zero elsewhere), s = max |interp|
      over the solve cube, r the SOR residual of the interpolated cube (the right-hand side of the equation
      the correction satisfies; the boundary enters through r)
    x0 [n,10,10,10,16]: space-to-depth of the 20^3 zero-extended frame, fine index f = 2c - 1 + d
      (c = 0..9, d = 0..1), channel ((dz*2+dy)*2+dx)*2 + q
    y1 = relu(conv(x0) + b1)                                  (valid 3^3 conv, 10^3 -> 8^3)
    y_l = y_{l-1} + relu(conv(pad(y_{l-1})) + b_l), l = 2..5   ('same', zero padding, 8^3)
    D = conv(pad(y5)) + b6, 8 channels                         (linear)
    the correction on the 16^3 interior frame (fine f = 1..16): corr[2o+p] = D[o][p],
      p = (pz*2+py)*2+px, o = (oz*8+oy)*8+ox;  guess = interp + s*corr

Parameters (pgamd_cnn_get_params order): for l = 1..6, W_l [16][432] then b_l [16]; W_l[co][tap*16+ci],
tap = dz*9 + dy*3 + dx; rows 8..15 of W6 and b6[8..15] are unused (zero).  With bf16=True, weights,
stored activations, pre-activations and the gradients fed to the matrix cores are rounded to BF16 at the
same points as the kernels; with bf16=False everything is FP64 (used to pin the backward pass against
finite differences).
"""
from __future__ import annotations

import numpy as np
from numpy.lib.stride_tricks import sliding_window_view

LAYER = 16 * 432 + 16
NPARAMS = 6 * LAYER
OW = {l: (l - 1) * LAYER for l in range(1, 7)}
OB = {l: OW[l] + 16 * 432 for l in range(1, 7)}
NOUT = 8


def bf16_round(a):
    """Round FP32 values to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a):
    """BF16 bit patterns (uint16) of the rounding of a."""
    return (bf16_round(a).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def unpack(p):
    p = np.asarray(p, np.float64)
    W = {l: p[OW[l]:OW[l] + 16 * 432].reshape(16, 432) for l in range(1, 7)}
    b = {l: p[OB[l]:OB[l] + 16] for l in range(1, 7)}
    return W, b


# ---- input packing and output unpacking --------------------------------------------------------------

RES_SCALE = np.float32(64.0)


def frame18(cube, rho, gs, cst, scale):
    """(q0, q1) on the 18^3 frame for one patch: cube [N^3] interpolated potential, rho [R^3].

    On the solve interior (zero elsewhere): q0 = cube / s; q1 = 64 * r / s, r the SOR residual of the
    interpolated cube in the reference operand order ((((((kp+km)+jp)+jm)+ip)+im) - 6 phi) - Const*rho, all in
    FP32."""
    N, R, off = 8 + 2 * gs, 6 + 2 * gs, 5 - gs
    c = np.asarray(cube, np.float32)
    q0 = np.zeros((18, 18, 18), np.float32)
    q1 = np.zeros((18, 18, 18), np.float32)
    inv = np.float32(1.0) / np.float32(scale)
    i = slice(1, N - 1)
    q0[off + 1:off + 1 + R, off + 1:off + 1 + R, off + 1:off + 1 + R] = c[i, i, i] * inv
    r = (((((c[2:, i, i] + c[:-2, i, i]) + c[i, 2:, i]) + c[i, :-2, i]) + c[i, i, 2:]) + c[i, i, :-2])
    r = (r - np.float32(6.0) * c[i, i, i]) - np.float32(cst) * np.asarray(rho, np.float32)
    q1[off + 1:off + 1 + R, off + 1:off + 1 + R, off + 1:off + 1 + R] = (r * inv) * RES_SCALE
    return q0, q1


def space_to_depth(q0, q1):
    """[n,18,18,18] x 2 -> [n,10,10,10,16] (fine index f = 2c - 1 + d)."""
    n = q0.shape[0]
    ext = np.zeros((n, 20, 20, 20, 2), np.float64)
    ext[:, 1:19, 1:19, 1:19, 0] = q0
    ext[:, 1:19, 1:19, 1:19, 1] = q1
    x = ext.reshape(n, 10, 2, 10, 2, 10, 2, 2).transpose(0, 1, 3, 5, 2, 4, 6, 7)
    return x.reshape(n, 10, 10, 10, 16)


def depth_to_space(D):
    """[n,512,8] -> correction on the 16^3 interior frame [n,16,16,16]."""
    n = D.shape[0]
    d = np.asarray(D).reshape(n, 8, 8, 8, 2, 2, 2).transpose(0, 1, 4, 2, 5, 3, 6)
    return d.reshape(n, 16, 16, 16)


def interior_to_depth(corr):
    """Adjoint layout of depth_to_space: [n,16,16,16] -> [n,512,8]."""
    n = corr.shape[0]
    d = np.asarray(corr).reshape(n, 8, 2, 8, 2, 8, 2).transpose(0, 1, 3, 5, 2, 4, 6)
    return d.reshape(n, 512, 8)


# ---- network ------------------------------------------------------------------------------------------

def _windows(x):
    """[n,S,S,S,c] -> [n,S-2,S-2,S-2,27*c] with K ordered tap-major (tap*c + ci)."""
    n, S, _, _, c = x.shape
    w = sliding_window_view(x, (3, 3, 3), axis=(1, 2, 3))          # [n,So,So,So,c,3,3,3]
    w = np.moveaxis(w, 4, 7)                                         # [n,So,So,So,3,3,3,c]
    return w.reshape(n, S - 2, S - 2, S - 2, 27 * c)


def _scatter_windows(g, S, c):
    """Adjoint of _windows: [n,S-2,S-2,S-2,27*c] -> [n,S,S,S,c]."""
    n, So = g.shape[0], S - 2
    g = g.reshape(n, So, So, So, 3, 3, 3, c)
    out = np.zeros((n, S, S, S, c))
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                out[:, dz:dz + So, dy:dy + So, dx:dx + So, :] += g[:, :, :, :, dz, dy, dx, :]
    return out


def _pad(y):
    n = y.shape[0]
    out = np.zeros((n, 10, 10, 10, 16))
    out[:, 1:9, 1:9, 1:9, :] = y
    return out


def forward(p, x0, bf16=True):
    """x0 [n,10,10,10,16] -> (D [n,512,8], cache)."""
    W, b = unpack(p)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    Wq = {l: r(W[l]) for l in W}
    x = np.asarray(x0, np.float64)
    cache = {"in": {}, "h": {}, "win": {}, "Wq": Wq}
    y = None
    for l in range(1, 7):
        xin = x if l == 1 else _pad(y)
        win = _windows(xin)                                          # [n,8,8,8,432]
        h = win @ Wq[l].T + b[l]
        cache["win"][l] = win
        if l == 6:
            D = h[..., :NOUT]
            break
        yn = np.maximum(h, 0.0)
        if l > 1:
            yn = yn + y
        if bf16:
            h, yn = r(h), r(yn)
        cache["h"][l] = h
        y = yn
    return D.reshape(-1, 512, NOUT), cache


def stencil16(e):
    """7-point operator (sum of the six neighbours - 6 e) on [n,16,16,16] with zeros outside: the SOR residual
    operator applied to a correction that vanishes on the Dirichlet layer."""
    f = np.pad(np.asarray(e, np.float64), ((0, 0), (1, 1), (1, 1), (1, 1)))
    return (f[:, 2:, 1:-1, 1:-1] + f[:, :-2, 1:-1, 1:-1] + f[:, 1:-1, 2:, 1:-1] + f[:, 1:-1, :-2, 1:-1]
            + f[:, 1:-1, 1:-1, 2:] + f[:, 1:-1, 1:-1, :-2]) - 6.0 * f[:, 1:-1, 1:-1, 1:-1]


def loss_grad(p, x0, target, bf16=True, res_weight=0.0):
    """Loss over all n*512*8 outputs and its gradient (parameter vector order):
    L = mean(e^2) + res_weight * sum(r^2) / (n*512*8), e = D - target, r = stencil16(e on the 16^3 frame), the
    residual the error leaves in the guess (the stencil is symmetric, so dL/de gets 2 res_weight stencil16(r))."""
    D, c = forward(p, x0, bf16)
    n = D.shape[0]
    t = np.asarray(target, np.float64).reshape(n, 512, NOUT)
    e = D - t
    loss = float(np.mean(e * e))
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    dD = 2.0 * e / e.size
    if res_weight:
        res = stencil16(depth_to_space(e))
        loss += float(res_weight * np.sum(res * res) / e.size)
        dD = dD + interior_to_depth(2.0 * res_weight * stencil16(res) / e.size)
    dh = np.zeros((n, 8, 8, 8, 16))
    dh[..., :NOUT] = r(dD).reshape(n, 8, 8, 8, NOUT)
    g = np.zeros(NPARAMS)
    dy = None
    for l in range(6, 0, -1):
        if l < 6:
            dh = np.where(c["h"][l] > 0, dy, 0.0)
            if bf16:
                dh = r(dh)
        g[OB[l]:OB[l] + 16] = dh.sum(axis=(0, 1, 2, 3))
        g[OW[l]:OW[l] + 16 * 432] = np.einsum("nzyxc,nzyxk->ck", dh, c["win"][l]).ravel()
        if l > 1:
            dx = _scatter_windows(dh @ c["Wq"][l], 10, 16)[:, 1:9, 1:9, 1:9, :]
            if l < 6:
                dx = dx + dy
            dy = dx
    return loss, g
