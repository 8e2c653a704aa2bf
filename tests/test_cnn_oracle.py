"""Pins on the CNN oracle (CPU): the analytic backward pass against central finite differences in FP64."""
import numpy as np
import pytest

from oracle import cnn_oracle as K


def _random_problem(seed, n=1):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(K.NPARAMS) * 0.05
    x0 = np.zeros((n, 18, 18, 18, 8))
    x0[..., :2] = rng.standard_normal((n, 18, 18, 18, 2))
    t = rng.standard_normal((n, 512)) * 0.1
    return p, x0, t


def test_parameter_layout():
    W, b, w6, b6 = K.unpack(np.arange(K.NPARAMS, dtype=np.float64))
    assert W[1].shape == (16, 216) and W[5].shape == (16, 432) and w6.shape == (432,)
    assert b[5][-1] + 1 == K.OW6 and b6 == K.NPARAMS - 1


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_finite_differences(seed):
    p, x0, t = _random_problem(seed)
    _, g = K.loss_grad(p, x0, t, bf16=False)
    rng = np.random.default_rng(100 + seed)
    # probe weights and biases of every layer (the padded input channels 2..7 of W1 have zero gradient)
    probes = [K.OW[1] + 5 * 216 + 3 * 8 + 1, K.OB[1] + 2, K.OW[3] + 7 * 432 + 200, K.OB[4] + 9,
              K.OW[5] + rng.integers(0, 16 * 432), K.OW6 + 17, K.OB6, K.OW[2] + 431]
    eps = 1e-6
    for i in probes:
        e = np.zeros(K.NPARAMS)
        e[i] = eps
        lp, _ = K.loss_grad(p + e, x0, t, bf16=False)
        lm, _ = K.loss_grad(p - e, x0, t, bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])
    assert np.all(g[K.OW[1]:K.OW[1] + 16 * 216].reshape(16, 27, 8)[:, :, 2:] == 0)


def test_forward_shapes_and_identity_at_zero_output_layer():
    p, x0, _ = _random_problem(3, n=2)
    p[K.OW6:K.OB6 + 1] = 0.0
    d, _ = K.forward(p, x0, bf16=True)
    assert d.shape == (2, 512) and np.all(d == 0)


def test_bf16_round_ties_to_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], np.float32)
    r = K.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
