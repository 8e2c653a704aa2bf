"""Mixed ghost sizes and schemes (per-ghost-size launches, and one persistent work-queue launch per scheme),
and HIP graph replay (needs an MI355X).

Each patch of a mixed batch must give exactly what the oracle gives for that patch alone (fixed iteration
counts: bit for bit), whatever the other patches are; graph replays must reproduce the eager solve bit for
bit, replay after replay, with unrelated eager work in between.
"""
import numpy as np
import pytest

import gpu_testbed_poisson_gravity_amd as pg
import oracle as O
from gpu_testbed_poisson_gravity_amd import inputs as I

pytestmark = pytest.mark.gpu
POI = 4.0 * np.pi * 1e-5
OMEGA = [1.49, 1.57, 1.62, 1.65, 1.69]


def _batch(npatch, seed, schemes=(pg.SOR, pg.MG)):
    rng = np.random.default_rng(seed)
    gs = rng.integers(1, 6, npatch).astype(np.int32)
    sc = rng.choice(np.array(schemes, np.int32), npatch)
    rho, cpot, flu = [None] * npatch, [None] * npatch, np.zeros((npatch, 5, 8, 8, 8), np.float32)
    for g in range(1, 6):
        idx = np.nonzero(gs == g)[0]
        if idx.size == 0:
            continue
        r, c, f, _ = I.lognormal(idx.size, g, seed=seed * 10 + g)
        for k, p in enumerate(idx):
            rho[p], cpot[p] = r[k], c[k]
            flu[p] = f[k]
    flu[:, 1:4] = 0.02 * np.cos(np.arange(flu[:, 1:4].size)).reshape(flu[:, 1:4].shape)
    return gs, sc, rho, cpot, flu


def _device(a):
    import torch

    return torch.as_tensor(np.ascontiguousarray(a)).to("cuda")


def _oracle_per_patch(gs, sc, rho, cpot, flu, args_kw, init_mode, term_mode, omega_by_gs):
    n = len(gs)
    pots, flus, its = [None] * n, np.zeros_like(flu), np.zeros(n, np.int32)
    for g in range(1, 6):
        for s in (pg.SOR, pg.MG):
            idx = np.nonzero((gs == g) & (sc == s))[0]
            if idx.size == 0:
                continue
            p = O.make_params(gs=g, scheme=s, init_mode=init_mode, term_mode=term_mode, poi_coeff=POI,
                              sor_omega=omega_by_gs[g - 1], **args_kw)
            po, fo, it, _ = O.solve(np.stack([rho[i] for i in idx]), np.stack([cpot[i] for i in idx]), flu[idx], p)
            for k, i in enumerate(idx):
                pots[i], flus[i], its[i] = po[k], fo[k], it[k]
    return np.stack(pots), flus, its


def _run_mixed(ctx, plan, rho, cpot, flu, args, omega):
    import torch

    n = plan.npatch
    d_rho = _device(pg.pack(rho, plan.rho_off, plan.rho_floats))
    d_cpot = _device(pg.pack(cpot, plan.pot_in_off, plan.pot_in_floats))
    d_flu = _device(flu)
    d_out = torch.zeros((n, 10, 10, 10), device="cuda")
    d_it = torch.zeros(n, dtype=torch.int32, device="cuda")
    d_st = torch.zeros(n, dtype=torch.int32, device="cuda")
    ctx.solve_mixed_device(plan, d_rho, d_cpot, d_out, d_flu, args, omega_by_ghost=omega, iters=d_it, status=d_st)
    torch.cuda.synchronize()
    return d_out.cpu().numpy(), d_flu.cpu().numpy(), d_it.cpu().numpy(), d_st.cpu().numpy()


@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("ppw1", [4, 2])
def test_mixed_fixed_iterations_bit_exact(ppw1, persistent):
    gs, sc, rho, cpot, flu = _batch(60, seed=1)
    print(f"[case] mixed fixed ppw1={ppw1} persistent={persistent} gs counts {np.bincount(gs)} "
          f"mg={int((sc == pg.MG).sum())}", flush=True)
    kw = dict(sor_min_iter=30, sor_max_iter=30, mg_max_iter=3, mg_tol=0.0, mg_npre=2, mg_npost=2, dt=0.3)
    with pg.Context(0) as ctx:
        ctx.set_options(init_mode=pg.INIT_INTERP, term_mode=pg.TERM_REFERENCE, fuse_gravity=1, gra_ghost_size=1,
                        mixed_persistent=persistent)
        ctx.set_layout(1, ppw1, 0)
        ctx.set_layout(2, 2 if ppw1 == 4 else 1, 0)
        plan = ctx.plan(gs, sc)
        args = pg.solver_args(poi_coeff=POI, **{k: v for k, v in kw.items()})
        pot, fl, it, st = _run_mixed(ctx, plan, rho, cpot, flu, args, OMEGA)
        plan.close()
    o_pot, o_flu, o_it = _oracle_per_patch(gs, sc, rho, cpot, flu, dict(sor_min_iter=30, sor_max_iter=30,
                                           mg_max_iter=3, mg_tol=0.0, mg_npre=2, mg_npost=2, dt=0.3),
                                           O.INIT_INTERP, O.TERM_REFERENCE, OMEGA)
    assert np.array_equal(it, o_it)
    assert np.array_equal(pot, o_pot)
    assert np.array_equal(fl, o_flu)


def test_mixed_relres_zero_guess():
    gs, sc, rho, cpot, flu = _batch(80, seed=2, schemes=(pg.SOR,))
    omega = [1.8] * 5
    kw = dict(sor_min_iter=10, sor_max_iter=1000, dt=0.3)
    with pg.Context(0) as ctx:
        ctx.set_options(init_mode=pg.INIT_ZERO, term_mode=pg.TERM_RELRES, check_every=2, rel_tol=1e-5)
        plan = ctx.plan(gs, sc)
        pot, fl, it, st = _run_mixed(ctx, plan, rho, cpot, flu, pg.solver_args(poi_coeff=POI, **kw), omega)
        plan.close()
    o_pot, _, o_it = _oracle_per_patch(gs, sc, rho, cpot, flu, dict(kw, check_every=2, rel_tol=1e-5),
                                       O.INIT_ZERO, O.TERM_RELRES, omega)
    assert np.all(st == pg.STATUS_CONVERGED)
    assert np.abs(it - o_it).max() <= 2
    same = it == o_it
    assert same.mean() > 0.8
    err = np.abs(pot[same] - o_pot[same]).max() / np.abs(o_pot[same]).max()
    assert err <= 1e-5


@pytest.mark.parametrize("persistent", [0, 1])
def test_graph_replay_mixed_matches_eager(persistent):
    import torch

    gs, sc, rho, cpot, flu = _batch(48, seed=3)
    n = len(gs)
    print(f"[case] graph replay mixed persistent={persistent}", flush=True)
    with pg.Context(0) as ctx:
        ctx.set_options(init_mode=pg.INIT_ZERO, term_mode=pg.TERM_RELRES, check_every=1, rel_tol=1e-5,
                        mixed_persistent=persistent)
        plan = ctx.plan(gs, sc)
        args = pg.solver_args(poi_coeff=POI, sor_min_iter=10, sor_max_iter=1000, sor_omega=1.8, mg_max_iter=20,
                              mg_tol=1e-5, mg_npre=2, mg_npost=2, dt=0.1)
        d_rho = _device(pg.pack(rho, plan.rho_off, plan.rho_floats))
        d_cpot = _device(pg.pack(cpot, plan.pot_in_off, plan.pot_in_floats))
        d_flu0 = _device(flu)
        d_flu = d_flu0.clone()
        d_out = torch.zeros((n, 10, 10, 10), device="cuda")
        d_it = torch.zeros(n, dtype=torch.int32, device="cuda")
        d_st = torch.zeros(n, dtype=torch.int32, device="cuda")
        ctx.solve_mixed_device(plan, d_rho, d_cpot, d_out, d_flu, args, iters=d_it, status=d_st)
        torch.cuda.synchronize()
        ref = [t.cpu().numpy() for t in (d_out, d_flu, d_it, d_st)]
        graph = ctx.graph_mixed(plan, d_rho, d_cpot, d_out, d_flu, args, iters=d_it, status=d_st)
        for rep in range(3):
            d_out.zero_()
            d_it.fill_(-1)
            d_flu.copy_(d_flu0)
            junk = torch.randn(1 << 20, device="cuda").sum()  # unrelated eager work between replays
            torch.cuda.synchronize()
            graph.launch(None)
            torch.cuda.synchronize()
            got = [t.cpu().numpy() for t in (d_out, d_flu, d_it, d_st)]
            for a, b in zip(got, ref):
                assert np.array_equal(a, b), f"replay {rep} differs"
            assert float(junk) == float(junk)
        graph.close()
        plan.close()


def test_graph_replay_uniform_matches_eager():
    import torch

    gs = 5
    rho, cpot, flu, _ = I.plummer(64, gs, seed=4)
    with pg.Context(0) as ctx:
        ctx.set_options(pot_ghost_size=gs, init_mode=pg.INIT_ZERO, term_mode=pg.TERM_RELRES, check_every=4)
        args = pg.solver_args(poi_coeff=POI, sor_min_iter=10, sor_max_iter=1000, sor_omega=1.8, dt=0.1)
        d_rho, d_cpot, d_flu0 = _device(rho), _device(cpot), _device(flu)
        d_flu = d_flu0.clone()
        d_out = torch.zeros((64, 10, 10, 10), device="cuda")
        d_it = torch.zeros(64, dtype=torch.int32, device="cuda")
        ctx.solve_device(d_rho, d_cpot, d_out, d_flu, 64, args, iters=d_it)
        torch.cuda.synchronize()
        ref = [t.cpu().numpy() for t in (d_out, d_flu, d_it)]
        graph = ctx.graph_device(d_rho, d_cpot, d_out, d_flu, 64, args, iters=d_it)
        for rep in range(3):
            d_out.zero_()
            d_flu.copy_(d_flu0)
            torch.cuda.synchronize()
            graph.launch(None)
            torch.cuda.synchronize()
            for a, b in zip((d_out.cpu().numpy(), d_flu.cpu().numpy(), d_it.cpu().numpy()), ref):
                assert np.array_equal(a, b), f"replay {rep} differs"
        graph.close()


def test_plan_validation():
    with pg.Context(0) as ctx:
        with pytest.raises(pg.PgError):
            ctx.plan([1, 2, 6])
        with pytest.raises(pg.PgError):
            ctx.plan([1, 2], [1, 3])
        plan = ctx.plan([1, 2, 3, 4, 5])
        sizes = [8 ** 3, 10 ** 3, 12 ** 3, 14 ** 3, 16 ** 3]
        assert plan.rho_floats == sum(sizes)
        assert list(plan.rho_off) == [0] + list(np.cumsum(sizes)[:-1])
        ctx.set_options(gra_ghost_size=2, pot_ghost_size=5)
        with pytest.raises(pg.PgError):  # GRA_GHOST_SIZE 2 > ghost size 1 of patch 0
            ctx.solve_mixed_device(plan, 1, 1, 1, 1, pg.solver_args())
        plan.close()
