"""Warm-start CNN on the GPU (needs an MI355X): MFMA convolutions against the numpy oracle, input
preparation and guess assembly against the oracle's interpolation, the guess check of the SOR kernel,
and the layout autotuner.

Tolerances: the oracle rounds activations, weights and ReLU gradients to BF16 at the same points as the
kernels and accumulates in FP64, the kernels in FP32; the normwise relative gap is then ~1e-6 plus the
rare BF16 rounding flip, so 2e-3 separates that from a layout or indexing error (which gives O(1)).
"""
import json
import os

import numpy as np
import pytest

import gpu_testbed_poisson_gravity_amd as pg
import oracle as O
from gpu_testbed_poisson_gravity_amd import inputs as I
from oracle import cnn_oracle as CO

pytestmark = pytest.mark.gpu
POI = 4.0 * np.pi * 1e-5
TOL_NET = 2e-3
# Gradients of the early layers sum many terms that cancel: one BF16 rounding flip of an activation or
# ReLU gradient (the FP32-vs-FP64 accumulation gap lands a value on the other side of a rounding boundary
# now and then) moves the W1 gradient by ~5e-3 normwise, measured on the oracle by perturbing the weights
# by 1e-7.  1e-2 admits a few such flips and still rejects any indexing or layout error.
TOL_GRAD = 1e-2


def _torch():
    import torch

    return torch


def _dev(a):
    return _torch().as_tensor(np.ascontiguousarray(a)).to("cuda")


def _bf16_bits(x):
    """float array -> (uint16 bit patterns of its BF16 rounding, the rounded values as float64)."""
    r = CO.bf16_round(x)
    bits = (r.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    return bits, r


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(CO.NPARAMS)
    for l in range(1, 7):
        p[CO.OW[l]:CO.OW[l] + 16 * 432] = rng.normal(0, np.sqrt(2 / 432) * (1.0 if l in (1, 6) else 0.5), 16 * 432)
        p[CO.OB[l]:CO.OB[l] + 16] = rng.normal(0, 0.1, 16)
    p[CO.OW[6] + CO.NOUT * 432:CO.OB[6]] = 0.0          # unused output rows
    p[CO.OB[6] + CO.NOUT:CO.OB[6] + 16] = 0.0
    return p.astype(np.float32)


def _random_x0(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 10, 10, 10, 16)).astype(np.float32)
    return CO.bf16_bits(x), CO.bf16_round(x)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def ctx():
    c = pg.Context(0)
    yield c
    c.close()


def test_num_params_and_param_roundtrip(ctx):
    assert pg.load().pgamd_cnn_num_params() == CO.NPARAMS
    net = pg.CNN(ctx, seed=3)
    try:
        p0 = net.get_params()
        assert p0.shape == (CO.NPARAMS,)
        W, b = CO.unpack(p0)
        assert np.all(W[6] == 0) and all(np.all(b[l] == 0) for l in b)   # untrained: d = 0
        assert np.all(W[1] != 0)
        p1 = _random_params(1)
        net.set_params(p1)
        assert np.array_equal(net.get_params(), p1)
        with pytest.raises(pg.PgError):
            pg._check(pg.load().pgamd_cnn_set_params(net.h, p1.ctypes.data_as(pg._FP), 100))
    finally:
        net.close()


@pytest.mark.parametrize("n", [1, 3])
def test_forward_matches_oracle(ctx, n):
    torch = _torch()
    p = _random_params(10 + n)
    bits, x0 = _random_x0(n, seed=20 + n)
    net = pg.CNN(ctx)
    try:
        net.set_params(p)
        d_x0 = _dev(bits.view(np.int16))
        d_out = torch.empty((n, 512, 8), device="cuda")
        print(f"[case] forward n={n}", flush=True)
        net.forward(d_x0, n, d_out)
        torch.cuda.synchronize()
        got = d_out.cpu().numpy()
    finally:
        net.close()
    want, _ = CO.forward(p, x0, bf16=True)
    assert np.all(np.isfinite(got))
    assert _rel(got, want) < TOL_NET, _rel(got, want)
    # each sample is computed independently of the others in the batch
    for k in range(n):
        assert _rel(got[k], want[k]) < TOL_NET


@pytest.mark.parametrize("res_weight", [0.0, 0.05])
def test_loss_and_gradient_match_oracle(ctx, res_weight):
    torch = _torch()
    n = 2
    p = _random_params(7)
    bits, x0 = _random_x0(n, seed=8)
    tgt = np.random.default_rng(9).normal(0, 0.5, (n, 512, 8)).astype(np.float32)
    net = pg.CNN(ctx)
    try:
        net.set_params(p)
        net.set_loss(res_weight)
        d_g = torch.empty(CO.NPARAMS, device="cuda")
        print(f"[case] loss_grad n=2 res_weight={res_weight}", flush=True)
        loss = net.loss_grad(_dev(bits.view(np.int16)), _dev(tgt), n, d_g)
        torch.cuda.synchronize()
        g = d_g.cpu().numpy()
        with pytest.raises(pg.PgError):
            net.set_loss(-1.0)
    finally:
        net.close()
    loss_o, g_o = CO.loss_grad(p, x0, tgt, bf16=True, res_weight=res_weight)
    assert abs(loss - loss_o) <= 1e-4 * loss_o
    blocks = [(f"W{l}", CO.OW[l], 16 * 432) for l in range(1, 7)] + [(f"b{l}", CO.OB[l], 16) for l in range(1, 7)]
    errs = {}
    for name, o, ln in blocks:
        errs[name] = _rel(g[o:o + ln], g_o[o:o + ln])
        print(f"  {name}: rel {errs[name]:.2e}", flush=True)
    assert np.all(np.any(g[CO.OW[1]:CO.OW[1] + 16 * 432].reshape(16, 27, 16) != 0, axis=(0, 1)))  # live inputs
    assert np.all(g[CO.OW[6] + CO.NOUT * 432:CO.OB[6]] == 0)                                  # unused outputs
    for name, e in errs.items():
        assert e < TOL_GRAD, (name, e)


def test_gradient_is_bitwise_reproducible(ctx):
    torch = _torch()
    n = 4
    p = _random_params(3)
    bits, _ = _random_x0(n, seed=4)
    tgt = np.random.default_rng(5).normal(0, 0.5, (n, 512, 8)).astype(np.float32)
    net = pg.CNN(ctx)
    try:
        net.set_params(p)
        gs = []
        for _ in range(2):
            d_g = torch.empty(CO.NPARAMS, device="cuda")
            net.loss_grad(_dev(bits.view(np.int16)), _dev(tgt), n, d_g)
            torch.cuda.synchronize()
            gs.append(d_g.cpu().numpy())
    finally:
        net.close()
    assert np.array_equal(gs[0], gs[1])


def test_training_reduces_loss(ctx):
    torch = _torch()
    n = 8
    bits, x0 = _random_x0(n, seed=11)
    # a learnable target: a fixed linear function of the two fields at the same fine cell
    core = x0[:, 1:9, 1:9, 1:9, :].reshape(n, 512, 8, 2)
    tgt = (0.3 * core[..., 0] - 0.2 * core[..., 1]).astype(np.float32)
    net = pg.CNN(ctx, seed=1)
    try:
        d_x0, d_t = _dev(bits.view(np.int16)), _dev(tgt)
        l0 = net.loss_grad(d_x0, d_t, n)
        losses = [net.train_step(d_x0, d_t, n, 3e-3, want_loss=True) for _ in range(150)]
        l1 = net.loss_grad(d_x0, d_t, n)
        p = net.get_params()
    finally:
        net.close()
    print(f"[train] loss {l0:.4e} -> {l1:.4e}", flush=True)
    assert np.isfinite(losses).all() and np.isfinite(p).all()
    assert l1 < 0.7 * l0


def _scale_and_interp(cpot, gs, ints):
    cube = O.interp(cpot, gs, ints, dtype=np.float32)
    s = np.abs(cube).reshape(cube.shape[0], -1).max(axis=1)
    s = np.where(s > 0, s, 1.0).astype(np.float32)
    return cube, s


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("ints", [pg.INT_CQUAD, pg.INT_QUAD])
def test_prepare_matches_oracle_interp(ctx, gs, ints):
    torch = _torch()
    n = 5
    rho, cpot, _, _ = I.plummer(n, gs, seed=gs)
    args = pg.solver_args(dh=0.5, poi_coeff=POI, int_scheme=ints)
    d_x0 = torch.empty((n, pg.CNN.INPUT), dtype=torch.int16, device="cuda")
    d_s = torch.empty(n, device="cuda")
    print(f"[case] prepare gs={gs} ints={ints}", flush=True)
    net = pg.CNN(ctx)
    try:
        net.prepare(_dev(rho), _dev(cpot), n, gs, args, d_x0, d_s)
        torch.cuda.synchronize()
    finally:
        net.close()
    x = d_x0.cpu().numpy().view(np.uint16).reshape(n, 10, 10, 10, 16)
    s_got = d_s.cpu().numpy()
    cube, s = _scale_and_interp(cpot, gs, ints)
    assert np.array_equal(s_got, s)
    cst = np.float32(np.float32(POI) * np.float32(0.5)) * np.float32(0.5)
    q = [CO.frame18(cube[k], rho[k], gs, cst, s[k]) for k in range(n)]
    want = CO.space_to_depth(np.stack([a for a, _ in q]), np.stack([b for _, b in q]))
    assert np.array_equal(x, CO.bf16_bits(want))


def _solve_dev(ctx, rho, cpot, flu, args, guess=None):
    torch = _torch()
    n, G = rho.shape[0], pg.geometry(ctx.opts.pot_ghost_size, ctx.opts.gra_ghost_size)["G"]
    d_out = torch.zeros((n, G, G, G), device="cuda")
    d_flu = _dev(flu)
    d_it = torch.zeros(n, dtype=torch.int32, device="cuda")
    d_st = torch.zeros(n, dtype=torch.int32, device="cuda")
    ctx.solve_device(_dev(rho), _dev(cpot), d_out, d_flu, n, args, guess=guess, iters=d_it, status=d_st)
    torch.cuda.synchronize()
    return d_out.cpu().numpy(), d_flu.cpu().numpy(), d_it.cpu().numpy(), d_st.cpu().numpy()


@pytest.mark.parametrize("gs", [1, 3, 5])
def test_warm_start_guess_assembly(gs):
    """guess = interp on the solve interior, + s*d on the patch-proper core (FP32, bit for bit); an
    untrained network gives the interpolated guess, so a guess solve equals the interpolation solve."""
    torch = _torch()
    n = 6
    rho, cpot, flu, _ = I.plummer(n, gs, seed=40 + gs)
    args = pg.solver_args(dh=1.0, poi_coeff=POI, sor_min_iter=20, sor_max_iter=20, sor_omega=1.6)
    R = pg.geometry(gs)["R"]
    with pg.Context(0) as c:
        c.set_options(pot_ghost_size=gs)
        net = pg.CNN(c, seed=2)
        try:
            d_rho, d_cpot = _dev(rho), _dev(cpot)
            d_guess = torch.empty((n, R, R, R), device="cuda")
            print(f"[case] warm start untrained gs={gs}", flush=True)
            net.warm_start(d_rho, d_cpot, n, args, d_guess)
            torch.cuda.synchronize()
            cube, s = _scale_and_interp(cpot, gs, pg.INT_CQUAD)
            interior = cube[:, 1:-1, 1:-1, 1:-1]
            assert np.array_equal(d_guess.cpu().numpy(), interior)
            c.set_options(init_mode=pg.INIT_INTERP)
            a = _solve_dev(c, rho, cpot, flu, args)
            c.set_options(init_mode=pg.INIT_GUESS)
            b = _solve_dev(c, rho, cpot, flu, args, guess=d_guess)
            for u, v in zip(a, b):
                assert np.array_equal(u, v)
            # a trained-looking network: the core moves by s*d
            net.set_params(_random_params(5))
            print(f"[case] warm start random gs={gs}", flush=True)
            net.warm_start(d_rho, d_cpot, n, args, d_guess)
            d_x0 = torch.empty((n, pg.CNN.INPUT), dtype=torch.int16, device="cuda")
            d_s = torch.empty(n, device="cuda")
            d_d = torch.empty((n, 512, 8), device="cuda")
            net.prepare(d_rho, d_cpot, n, gs, args, d_x0, d_s)
            net.forward(d_x0, n, d_d)
            torch.cuda.synchronize()
        finally:
            net.close()
    off = 5 - gs
    corr = CO.depth_to_space(d_d.cpu().numpy())[:, off:off + R, off:off + R, off:off + R].astype(np.float32)
    want = interior + s[:, None, None, None] * corr
    assert np.array_equal(d_guess.cpu().numpy(), want)


@pytest.mark.parametrize("gs", [2, 5])
@pytest.mark.parametrize("scheme", [pg.SOR, pg.MG])
def test_whole_cube_guess_matches_interior_guess(gs, scheme):
    """pgamd_cnn_warm_start_cube: the interior is bitwise the interior guess and the Dirichlet layer bitwise the
    interpolated coarse potential; a solve from it (init GUESS_CUBE, SOR reading its Dirichlet layer from the
    guess) is bitwise the solve from the interior guess, with and without the guess check."""
    torch = _torch()
    n = 12
    rho, cpot, flu, _ = I.plummer(n, gs, seed=60 + gs)
    args = pg.solver_args(dh=1.0, poi_coeff=POI, sor_min_iter=25, sor_max_iter=25, sor_omega=1.7, mg_max_iter=3,
                          mg_tol=0.0, mg_npre=2, mg_npost=2)
    g = pg.geometry(gs)
    R, N = g["R"], g["N"]
    with pg.Context(0) as c:
        c.set_options(pot_ghost_size=gs, pot_scheme=scheme)
        net = pg.CNN(c, seed=2)
        net.set_params(_random_params(7))
        try:
            d_rho, d_cpot = _dev(rho), _dev(cpot)
            d_int = torch.empty((n, R, R, R), device="cuda")
            d_cube = torch.full((n, N, N, N), float("nan"), device="cuda")
            print(f"[case] whole-cube guess gs={gs} scheme={scheme}", flush=True)
            net.warm_start(d_rho, d_cpot, n, args, d_int)
            net.warm_start_cube(d_rho, d_cpot, n, args, d_cube)
            torch.cuda.synchronize()
            cube_g, int_g = d_cube.cpu().numpy(), d_int.cpu().numpy()
            assert np.array_equal(cube_g[:, 1:-1, 1:-1, 1:-1], int_g)
            interp = O.interp(cpot, gs, pg.INT_CQUAD, dtype=np.float32)
            face = np.ones((N, N, N), bool)
            face[1:-1, 1:-1, 1:-1] = False
            assert np.array_equal(cube_g[:, face], interp[:, face])
            for tol in (0.0, 1e-2):
                c.set_options(init_mode=pg.INIT_GUESS, guess_fallback_tol=tol)
                a = _solve_dev(c, rho, cpot, flu, args, guess=d_int)
                c.set_options(init_mode=pg.INIT_GUESS_CUBE)
                b = _solve_dev(c, rho, cpot, flu, args, guess=d_cube)
                for u, v in zip(a, b):
                    assert np.array_equal(u, v)
        finally:
            net.close()


def test_guess_fallback_to_zero_guess():
    """With guess_fallback_tol > 0 the SOR kernel rejects a guess whose relative residual exceeds the
    tolerance, restarts from the zero guess and sets status bit 8; an accepted guess is used as given."""
    torch = _torch()
    gs, n = 5, 8
    rho, cpot, flu, _ = I.plummer(n, gs, seed=3)
    R = pg.geometry(gs)["R"]
    args = pg.solver_args(dh=1.0, poi_coeff=POI, sor_min_iter=30, sor_max_iter=30, sor_omega=1.69)
    interior = O.interp(cpot, gs, pg.INT_CQUAD, dtype=np.float32)[:, 1:-1, 1:-1, 1:-1]
    bad = (interior * 1.5 + np.float32(1e-3)).astype(np.float32)
    bad[::2] = interior[::2]                                    # even patches keep the interpolated guess
    with pg.Context(0) as c:
        c.set_options(pot_ghost_size=gs, init_mode=pg.INIT_ZERO)
        zero = _solve_dev(c, rho, cpot, flu, args)
        c.set_options(init_mode=pg.INIT_GUESS, guess_fallback_tol=0.0)
        plain = _solve_dev(c, rho, cpot, flu, args, guess=_dev(bad))
        c.set_options(guess_fallback_tol=0.05)
        print("[case] guess fallback", flush=True)
        chk = _solve_dev(c, rho, cpot, flu, args, guess=_dev(bad))
        with pytest.raises(pg.PgError):
            c.set_options(guess_fallback_tol=-1.0)
    # the oracle's relative residual (FP64 sum|R| / sum|phi|) puts the two guesses on either side of 0.05
    for k in (0, 1):
        cube = O.interp(cpot[k:k + 1], gs, pg.INT_CQUAD, dtype=np.float32)[0]
        cube[1:-1, 1:-1, 1:-1] = bad[k]
        rr = O.relres(cube, rho[k], gs, POI, 1.0)
        assert (rr > 0.05) == (k == 1), rr
    fell = (chk[3] & pg.STATUS_FALLBACK) != 0
    assert np.array_equal(fell, np.arange(n) % 2 == 1)
    for k in range(n):
        ref = zero if fell[k] else plain
        assert np.array_equal(chk[0][k], ref[0][k]) and np.array_equal(chk[1][k], ref[1][k])
        assert chk[2][k] == ref[2][k]
        assert (chk[3][k] & ~pg.STATUS_FALLBACK) == ref[3][k]


def test_mg_guess_fallback_to_zero_guess():
    """The MG twin of the SOR guess check: with guess_fallback_tol > 0 a guess whose relative residual
    (sum|R| / sum|phi|, the SOR residual) exceeds the tolerance is replaced by the zero guess before the first
    V-cycle and status bit 8 is set; an accepted guess is used as given."""
    _torch()
    gs, n = 4, 8
    rho, cpot, flu, _ = I.plummer(n, gs, seed=5)
    args = pg.solver_args(dh=1.0, poi_coeff=POI, mg_max_iter=3, mg_npre=2, mg_npost=2, mg_tol=0.0)
    interior = O.interp(cpot, gs, pg.INT_CQUAD, dtype=np.float32)[:, 1:-1, 1:-1, 1:-1]
    bad = (interior * 1.5 + np.float32(1e-3)).astype(np.float32)
    bad[::2] = interior[::2]                                    # even patches keep the interpolated guess
    with pg.Context(0) as c:
        c.set_options(pot_ghost_size=gs, pot_scheme=pg.MG, init_mode=pg.INIT_ZERO)
        zero = _solve_dev(c, rho, cpot, flu, args)
        c.set_options(init_mode=pg.INIT_GUESS, guess_fallback_tol=0.0)
        plain = _solve_dev(c, rho, cpot, flu, args, guess=_dev(bad))
        c.set_options(guess_fallback_tol=0.05)
        print("[case] MG guess fallback", flush=True)
        chk = _solve_dev(c, rho, cpot, flu, args, guess=_dev(bad))
    for k in (0, 1):
        cube = O.interp(cpot[k:k + 1], gs, pg.INT_CQUAD, dtype=np.float32)[0]
        cube[1:-1, 1:-1, 1:-1] = bad[k]
        rr = O.relres(cube, rho[k], gs, POI, 1.0)
        assert (rr > 0.05) == (k == 1), rr
    fell = (chk[3] & pg.STATUS_FALLBACK) != 0
    assert np.array_equal(fell, np.arange(n) % 2 == 1)
    assert not np.any(plain[3] & pg.STATUS_FALLBACK)
    for k in range(n):
        ref = zero if fell[k] else plain
        assert np.array_equal(chk[0][k], ref[0][k]) and np.array_equal(chk[1][k], ref[1][k])
        assert chk[2][k] == ref[2][k]
        assert (chk[3][k] & ~pg.STATUS_FALLBACK) == ref[3][k]


def test_autotune_keeps_results_and_roundtrips(tmp_path):
    torch = _torch()
    gs, n = 4, 512
    rho, cpot, flu, _ = I.plummer(n, gs, seed=9)
    args = pg.solver_args(dh=1.0, poi_coeff=POI, sor_min_iter=25, sor_max_iter=25, sor_omega=1.65)
    with pg.Context(0) as c:
        c.set_options(pot_ghost_size=gs, fuse_gravity=0)
        base = _solve_dev(c, rho, cpot, flu, args)
        G = pg.geometry(gs)["G"]
        d_out = torch.zeros((n, G, G, G), device="cuda")
        print("[case] autotune gs=4", flush=True)
        ppw, pad, ms = c.autotune(gs, _dev(rho), _dev(cpot), d_out, n, args, repeats=2)
        assert ms > 0 and c.get_layout(gs) == (ppw, pad)
        tuned = _solve_dev(c, rho, cpot, flu, args)
        path = str(tmp_path / "layouts.json")
        c.autotune_save(path)
        with open(path) as f:
            data = json.load(f)
        assert data["layouts"][str(gs)] == [ppw, pad]
    for u, v in zip(base, tuned):
        assert np.array_equal(u, v)
    with pg.Context(0) as c2:
        c2.autotune_load(path)
        assert c2.get_layout(gs) == (ppw, pad)
    assert os.path.exists(path)


def test_mixed_warm_start_matches_uniform():
    """pgamd_cnn_warm_start_mixed on a plan equals pgamd_cnn_warm_start per ghost size, bit for bit."""
    torch = _torch()
    rng = np.random.default_rng(12)
    gsz = rng.integers(1, 6, 40).astype(np.int32)
    rho, cpot = [None] * 40, [None] * 40
    for g in range(1, 6):
        idx = np.nonzero(gsz == g)[0]
        if idx.size:
            r, c, _, _ = I.plummer(idx.size, g, seed=60 + g)
            for k, i in enumerate(idx):
                rho[i], cpot[i] = r[k], c[k]
    args = pg.solver_args(dh=1.0, poi_coeff=POI)
    p = _random_params(21)
    with pg.Context(0) as c:
        net = pg.CNN(c)
        plan = c.plan(gsz)
        try:
            net.set_params(p)
            d_rho = _dev(pg.pack(rho, plan.rho_off, plan.rho_floats))
            d_cpot = _dev(pg.pack(cpot, plan.pot_in_off, plan.pot_in_floats))
            d_guess = torch.zeros(plan.rho_floats, device="cuda")
            print("[case] mixed warm start", flush=True)
            net.warm_start_mixed(plan, d_rho, d_cpot, args, d_guess)
            torch.cuda.synchronize()
            got = d_guess.cpu().numpy()
            for g in range(1, 6):
                idx = np.nonzero(gsz == g)[0]
                if idx.size == 0:
                    continue
                R = pg.geometry(g)["R"]
                c.set_options(pot_ghost_size=g)
                u = torch.empty((idx.size, R, R, R), device="cuda")
                net.warm_start(_dev(np.stack([rho[i] for i in idx])), _dev(np.stack([cpot[i] for i in idx])),
                               idx.size, args, u)
                torch.cuda.synchronize()
                u = u.cpu().numpy()
                for k, i in enumerate(idx):
                    o = plan.rho_off[i]
                    assert np.array_equal(got[o:o + R ** 3].reshape(R, R, R), u[k]), (g, i)
        finally:
            plan.close()
            net.close()


@pytest.mark.parametrize("nchunk", [1, 3, 16])
def test_pipelined_warm_solve_matches_serial(nchunk):
    """pgamd_cnn_solve_device (network of slice i+1 on a second stream while slice i is solved) gives the
    results of pgamd_cnn_warm_start followed by pgamd_solve_device, bit for bit."""
    torch = _torch()
    gs, n = 5, 40
    rho, cpot, flu, _ = I.plummer(n, gs, seed=70 + nchunk)
    args = pg.solver_args(dt=1e-3, sor_min_iter=10, sor_max_iter=1000, sor_omega=1.8, poi_coeff=POI)
    with pg.Context(0) as c:
        c.set_options(pot_ghost_size=gs, init_mode=pg.INIT_GUESS, term_mode=pg.TERM_RELRES, rel_tol=1e-5,
                      guess_fallback_tol=1e-2, fuse_gravity=1)
        net = pg.CNN(c)
        try:
            net.set_params(_random_params(4) * 0.05)
            d_rho, d_cpot = _dev(rho), _dev(cpot)
            res = []
            for pipelined in (False, True):
                guess = torch.zeros((n, 16, 16, 16), device="cuda")
                out = torch.zeros((n, 10, 10, 10), device="cuda")
                d_flu = _dev(flu)
                it = torch.zeros(n, dtype=torch.int32, device="cuda")
                st = torch.zeros(n, dtype=torch.int32, device="cuda")
                print(f"[case] warm solve nchunk={nchunk} pipelined={pipelined}", flush=True)
                if pipelined:
                    net.solve_device(d_rho, d_cpot, guess, out, d_flu, n, args, iters=it, status=st, nchunk=nchunk)
                else:
                    net.warm_start(d_rho, d_cpot, n, args, guess)
                    c.solve_device(d_rho, d_cpot, out, d_flu, n, args, guess=guess, iters=it, status=st)
                torch.cuda.synchronize()
                res.append([x.cpu().numpy() for x in (guess, out, d_flu, it, st)])
            with pytest.raises(pg.PgError):
                net.solve_device(d_rho, d_cpot, guess, out, d_flu, n, args, nchunk=17)
        finally:
            net.close()
    for a, b in zip(*res):
        assert np.array_equal(a, b)
    assert np.all((res[1][4] & pg.STATUS_CONVERGED) != 0)
