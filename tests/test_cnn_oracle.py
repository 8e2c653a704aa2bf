"""Pins on the CNN oracle (CPU): the analytic backward pass against central finite differences in FP64,
the input/output packing against direct index arithmetic, and BF16 rounding against hand-worked values."""
import numpy as np
import pytest

from oracle import cnn_oracle as K


def _random_problem(seed, n=1):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(K.NPARAMS) * 0.05
    for l in range(1, 7):
        p[K.OB[l]:K.OB[l] + 16] = rng.uniform(0.0, 0.1, 16)   # keep most ReLUs open
    W6 = p[K.OW[6]:K.OW[6] + 16 * 432].reshape(16, 432)
    W6[K.NOUT:] = 0.0
    p[K.OB[6] + K.NOUT:K.OB[6] + 16] = 0.0
    x0 = rng.standard_normal((n, 10, 10, 10, 16))
    t = rng.standard_normal((n, 512, K.NOUT)) * 0.1
    return p, x0, t


def test_parameter_layout():
    W, b = K.unpack(np.arange(K.NPARAMS, dtype=np.float64))
    assert all(W[l].shape == (16, 432) for l in range(1, 7))
    assert b[1][0] == 16 * 432 and W[2][0, 0] == K.LAYER and b[6][-1] == K.NPARAMS - 1


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_finite_differences(seed):
    p, x0, t = _random_problem(seed)
    _, g = K.loss_grad(p, x0, t, bf16=False)
    rng = np.random.default_rng(100 + seed)
    probes = [K.OW[1] + 5 * 432 + 3 * 16 + 1, K.OB[1] + 2, K.OW[3] + 7 * 432 + 200, K.OB[4] + 9,
              K.OW[5] + int(rng.integers(0, 16 * 432)), K.OW[6] + 3 * 432 + 17, K.OB[6] + 5, K.OW[2] + 431,
              K.OW[4] + 15 * 432 + 430]
    eps = 1e-6
    for i in probes:
        e = np.zeros(K.NPARAMS)
        e[i] = eps
        lp, _ = K.loss_grad(p + e, x0, t, bf16=False)
        lm, _ = K.loss_grad(p - e, x0, t, bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])
    # the unused output channels carry no gradient
    assert np.all(g[K.OW[6] + K.NOUT * 432:K.OB[6]] == 0) and np.all(g[K.OB[6] + K.NOUT:K.OB[6] + 16] == 0)


@pytest.mark.parametrize("seed", [5, 6])
def test_residual_loss_backward_matches_finite_differences(seed):
    p, x0, t = _random_problem(seed)
    w = 0.05
    _, g = K.loss_grad(p, x0, t, bf16=False, res_weight=w)
    _, g0 = K.loss_grad(p, x0, t, bf16=False)
    assert np.abs(g - g0).max() > 1e-3 * np.abs(g0).max()    # the residual term changes the gradient
    probes = [K.OW[6] + 2 * 432 + 40, K.OB[6] + 7, K.OW[1] + 9 * 432 + 100, K.OB[3] + 4, K.OW[4] + 3 * 432 + 5]
    eps = 1e-6
    for i in probes:
        e = np.zeros(K.NPARAMS)
        e[i] = eps
        lp, _ = K.loss_grad(p + e, x0, t, bf16=False, res_weight=w)
        lm, _ = K.loss_grad(p - e, x0, t, bf16=False, res_weight=w)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_residual_stencil():
    # a unit spike: -6 at the cell, 1 on each neighbour inside the frame (a face cell loses the outside one)
    e = np.zeros((1, 16, 16, 16))
    e[0, 5, 6, 7] = 1.0
    e[0, 0, 9, 9] = 2.0
    r = K.stencil16(e)
    assert r[0, 5, 6, 7] == -6 and r[0, 4, 6, 7] == 1 and r[0, 6, 6, 7] == 1 and r[0, 5, 5, 7] == 1
    assert r[0, 5, 7, 7] == 1 and r[0, 5, 6, 6] == 1 and r[0, 5, 6, 8] == 1
    assert r[0, 0, 9, 9] == -12 and r[0, 1, 9, 9] == 2 and np.sum(r[0, :, 9, 9] != 0) == 2
    # quadratic x^2 + y^2 + z^2: 6 wherever no neighbour falls outside
    ax = np.arange(16.0)
    q = (ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2)[None]
    assert np.all(K.stencil16(q)[0, 1:-1, 1:-1, 1:-1] == 6.0)
    # symmetric (self-adjoint with the zero exterior), which the gradient 2 w stencil16(r) relies on
    rng = np.random.default_rng(7)
    a, b = rng.standard_normal((2, 1, 16, 16, 16))
    assert abs(np.sum(K.stencil16(a) * b) - np.sum(a * K.stencil16(b))) < 1e-9 * np.sum(np.abs(a))


def test_zero_output_layer_gives_zero_correction():
    p, x0, _ = _random_problem(3, n=2)
    p[K.OW[6]:] = 0.0
    d, _ = K.forward(p, x0, bf16=True)
    assert d.shape == (2, 512, K.NOUT) and np.all(d == 0)


def test_packing_index_arithmetic():
    rng = np.random.default_rng(4)
    q0, q1 = rng.standard_normal((2, 2, 18, 18, 18))
    x = K.space_to_depth(q0, q1)
    # fine index f = 2c - 1 + d; channel ((dz*2+dy)*2+dx)*2 + q
    for _ in range(50):
        c = rng.integers(0, 10, 3)
        d = rng.integers(0, 2, 3)
        q = int(rng.integers(0, 2))
        f = 2 * c - 1 + d
        ch = ((d[0] * 2 + d[1]) * 2 + d[2]) * 2 + q
        want = 0.0 if np.any(f < 0) or np.any(f > 17) else (q0, q1)[q][1, f[0], f[1], f[2]]
        assert x[1, c[0], c[1], c[2], ch] == want
    corr = rng.standard_normal((2, 16, 16, 16))
    D = K.interior_to_depth(corr)
    assert np.array_equal(K.depth_to_space(D), corr)
    # interior frame index i = 2o + p, the cell that x0 holds at coarse o+1, sub-cell p
    o, pp = np.array([3, 0, 7]), np.array([1, 0, 1])
    i = 2 * o + pp
    assert D[0, (o[0] * 8 + o[1]) * 8 + o[2], (pp[0] * 2 + pp[1]) * 2 + pp[2]] == corr[0, i[0], i[1], i[2]]
    f = i + 1
    assert np.array_equal((f + 1) // 2, o + 1) and np.array_equal((f + 1) % 2, pp)


def test_frame_centres_smaller_cubes():
    # q1 is 64 r / s with r the 7-point residual of the cube: on phi = x^2 + y^2 + z^2 (cell units) the
    # discrete Laplacian is exactly 6, so r = 6 - Const*rho at every interior cell
    ax = np.arange(10, dtype=np.float32)
    cube = (ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2).astype(np.float32)
    rho = np.ones((8, 8, 8), np.float32)
    q0, q1 = K.frame18(cube, rho, gs=1, cst=2.0, scale=4.0)
    # gs 1: the 10^3 cube sits at 4..13, its 8^3 interior at 5..12; both fields live on the interior only
    assert q0[5, 5, 5] == cube[1, 1, 1] / 4 and q0[6, 5, 5] == cube[2, 1, 1] / 4
    assert np.array_equal(q0[5:13, 5:13, 5:13], cube[1:9, 1:9, 1:9] / np.float32(4))
    assert np.all(q1[5:13, 5:13, 5:13] == 64.0 * (6.0 - 2.0) / 4.0)
    for q in (q0, q1):   # the Dirichlet layer (4 and 13) and outside are zero
        assert q[4].max() == 0 and q[13].max() == 0 and q[:, 4].max() == 0 and q[:, :, 13].max() == 0
        assert q[:4].max() == 0 and q[14:].max() == 0


def test_bf16_round_ties_to_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], np.float32)
    r = K.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert K.bf16_bits(np.float32(1.0)) == 0x3F80 and K.bf16_bits(np.float32(-2.0)) == 0xC000
