"""pgamd_cnn_forward_noalloc (the network forward with its activations in the library's fixed device array)
against pgamd_cnn_forward, bit for bit, on two calls in a row, the second with a smaller batch (needs an
MI355X)."""
import numpy as np
import pytest

import gpu_testbed_poisson_gravity_amd as pg
from oracle import cnn_oracle as CO

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(CO.NPARAMS)
    for l in range(1, 7):
        p[CO.OW[l]:CO.OW[l] + 16 * 432] = rng.normal(0, np.sqrt(2 / 432), 16 * 432)
        p[CO.OB[l]:CO.OB[l] + 16] = rng.normal(0, 0.1, 16)
    p[CO.OW[6] + CO.NOUT * 432:CO.OB[6]] = 0.0
    p[CO.OB[6] + CO.NOUT:CO.OB[6] + 16] = 0.0
    return p.astype(np.float32)


def _x0(n, seed):
    """Packed inputs with zero borders (the layout pgamd_cnn_prepare writes), as BF16 bits."""
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 10, 10, 10, 16), np.float32)
    x[:, 1:9, 1:9, 1:9] = rng.uniform(-1, 1, (n, 8, 8, 8, 16))
    return CO.bf16_bits(x).view(np.int16).reshape(n, -1)


def test_forward_noalloc_matches_forward_on_two_calls():
    import torch

    with pg.Context(0) as ctx:
        net = pg.CNN(ctx, seed=1)
        try:
            net.set_params(_params(3))
            for n, seed in ((64, 11), (24, 12)):          # the second call is smaller
                d_x0 = torch.from_numpy(_x0(n, seed)).cuda()
                ref = torch.empty((n, 512, 8), device="cuda")
                got = torch.full((n, 512, 8), float("nan"), device="cuda")
                print(f"[case] forward_noalloc n={n}", flush=True)
                net.forward(d_x0, n, ref)
                net.forward_noalloc(d_x0, n, got)
                torch.cuda.synchronize()
                assert np.array_equal(got.cpu().numpy(), ref.cpu().numpy())
            with pytest.raises(pg.PgError) as e:            # beyond the fixed array: refused before any launch
                net.forward_noalloc(d_x0, 65, got)
            assert e.value.code == pg.ERR_ARG
        finally:
            net.close()
