"""What the trained warm-start network gains (needs an MI355X and the committed weights).

Held-out seeded patches (ghost size 5, Plummer and lognormal densities, seeds the training never drew) are
solved to relative residual 1e-5 (omega 1.8, min 10 / max 1000 iterations) three ways: from the zero
guess (the plain method), from the interpolated coarse potential (the reference's start) and from the
network's guess with the SOR kernel's guess check at 1e-2.  The network must cut the iteration count
against both, every patch must still reach the tolerance, and all three solutions must agree to the
accuracy the tolerance implies.  The CPU oracle's relative residual of the guesses themselves is compared
on a subset: the network's guess must be closer to the discrete solution than the interpolation.
"""
import os

import numpy as np
import pytest

import gpu_testbed_poisson_gravity_amd as pg
import oracle as O
from gpu_testbed_poisson_gravity_amd import inputs as I

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(pg.__file__), "weights", "warm_start_gs5.npy")
GS, POI = 5, 4.0 * np.pi * 1e-5


def _solve(ctx, d_rho, d_cpot, n, mode, guess=None, fallback=0.0):
    import torch

    ctx.set_options(pot_ghost_size=GS, init_mode=mode, term_mode=pg.TERM_RELRES, rel_tol=1e-5, check_every=1,
                    fuse_gravity=0, guess_fallback_tol=fallback)
    out = torch.empty((n, 10, 10, 10), device="cuda")
    it = torch.empty(n, dtype=torch.int32, device="cuda")
    st = torch.empty(n, dtype=torch.int32, device="cuda")
    args = pg.solver_args(dt=1e-3, sor_min_iter=10, sor_max_iter=1000, sor_omega=1.8, poi_coeff=POI, graacc=False)
    ctx.solve_device(d_rho, d_cpot, out, None, n, args, guess=guess, iters=it, status=st)
    torch.cuda.synchronize()
    return out.cpu().numpy(), it.cpu().numpy(), st.cpu().numpy()


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="trained weights not present")
@pytest.mark.parametrize("kind", ["plummer", "lognormal"])
def test_trained_network_reduces_iterations(kind):
    import torch

    n = 1024
    if kind == "plummer":
        rho, cpot, _, _ = I.plummer(n, GS, seed=987_654)
    else:
        rho, cpot, _, _ = I.lognormal(n, GS, seed=876_543)
    d_rho, d_cpot = torch.as_tensor(rho).cuda(), torch.as_tensor(cpot).cuda()
    with pg.Context(0) as ctx:
        net = pg.CNN(ctx)
        try:
            net.load_params(WEIGHTS)
            guess = torch.empty((n, 16, 16, 16), device="cuda")
            args = pg.solver_args(dt=1e-3, sor_omega=1.8, poi_coeff=POI, graacc=False)
            ctx.set_options(pot_ghost_size=GS)
            print(f"[case] warm start {kind} n={n}", flush=True)
            net.warm_start(d_rho, d_cpot, n, args, guess)
            torch.cuda.synchronize()
            zero = _solve(ctx, d_rho, d_cpot, n, pg.INIT_ZERO)
            interp = _solve(ctx, d_rho, d_cpot, n, pg.INIT_INTERP)
            warm = _solve(ctx, d_rho, d_cpot, n, pg.INIT_GUESS, guess=guess, fallback=1e-2)
            g_host = guess.cpu().numpy()
        finally:
            net.close()
    iz, ii, iw = zero[1].mean(), interp[1].mean(), warm[1].mean()
    fb = int(((warm[2] & pg.STATUS_FALLBACK) != 0).sum())
    print(f"[gain] {kind}: mean iterations zero {iz:.1f}, interp {ii:.1f}, network {iw:.1f} "
          f"(fallbacks {fb}/{n}); ratio vs zero {iw / iz:.2f}", flush=True)
    for r in (zero, interp, warm):
        assert np.all((r[2] & pg.STATUS_CONVERGED) != 0)
    assert iw < 0.8 * iz and iw < ii
    # same discrete solution to the accuracy a 1e-5 relative residual implies
    scale = np.abs(zero[0]).max(axis=(1, 2, 3))
    diff = np.abs(warm[0] - zero[0]).max(axis=(1, 2, 3)) / scale
    assert np.percentile(diff, 99) < 1e-3, np.percentile(diff, 99)
    # the network's guess is closer to the solution than the interpolation (oracle residuals, subset)
    cube = O.interp(cpot[:64], GS, pg.INT_CQUAD, dtype=np.float32)
    rr_i = np.array([O.relres(cube[k], rho[k], GS, POI, 1.0) for k in range(64)])
    cube[:, 1:-1, 1:-1, 1:-1] = g_host[:64]
    rr_w = np.array([O.relres(cube[k], rho[k], GS, POI, 1.0) for k in range(64)])
    print(f"[gain] {kind}: median relative residual of the guess: interp {np.median(rr_i):.2e}, "
          f"network {np.median(rr_w):.2e}", flush=True)
    assert np.median(rr_w) < np.median(rr_i)
