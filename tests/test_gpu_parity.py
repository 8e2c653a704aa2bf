"""HIP path against the CPU oracle on the same seeded inputs (needs an MI355X).

Gates (FP32):
* fixed iteration counts (Min_Iter = Max_Iter, or MG with an unreachable tolerance): the kernels keep the
  reference's operation order, so potential and fluid must be bit-identical to the oracle; the
  normwise relative L-infinity bound 1e-5 of the specification is asserted as well.
* data-dependent stopping rules: iteration counts may differ by a check when a floating-point sum sits on
  the threshold; patches with equal counts must agree within 1e-5 (normwise), and every patch must meet
  its own stopping rule.
"""
import numpy as np
import pytest

import gpu_testbed_poisson_gravity_amd as pg
import oracle as O
from gpu_testbed_poisson_gravity_amd import inputs as I

pytestmark = pytest.mark.gpu
POI = 4.0 * np.pi * 1e-5
OMEGA = [1.49, 1.57, 1.62, 1.65, 1.69]
TOL = 1e-5


def _rel_linf(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


def _mixed_inputs(n, gs, seed):
    r1, c1, f1 = I.reference_ic(n // 2, gs)
    r2, c2, f2, _ = I.plummer(n - n // 2, gs, seed=seed)
    flu = np.concatenate([f1, f2])
    flu[:, 1:4] = 0.05 * np.sin(np.arange(flu[:, 1:4].size, dtype=np.float64)).reshape(flu[:, 1:4].shape)
    return np.concatenate([r1, r2]), np.concatenate([c1, c2]), flu.astype(np.float32)


def _oracle_params(gs, gg, scheme, ints, init_mode, term_mode, check_every, kw):
    return O.make_params(gs=gs, gg=gg, scheme=scheme, int_scheme=ints, init_mode=init_mode, term_mode=term_mode,
                         check_every=check_every, sor_min_iter=kw["sor_min_iter"], sor_max_iter=kw["sor_max_iter"],
                         sor_omega=kw["sor_omega"], mg_max_iter=kw["mg_max_iter"], mg_npre=kw["mg_npre"],
                         mg_npost=kw["mg_npost"], mg_tol=kw["mg_tol"], rel_tol=kw.get("rel_tol", 1e-5),
                         poi_coeff=POI, dt=kw["dt"], p5=kw.get("p5", False))


def _both(rho, cpot, flu, gs, gg=1, scheme=pg.SOR, ints=pg.INT_CQUAD, init_mode=pg.INIT_INTERP,
          term_mode=pg.TERM_REFERENCE, check_every=1, guess=None, fuse=True, **kw):
    base = dict(sor_min_iter=50, sor_max_iter=50, sor_omega=OMEGA[gs - 1], mg_max_iter=10, mg_npre=3, mg_npost=3,
                mg_tol=1e-6, dt=0.5)
    base.update(kw)
    print(f"[case] gs={gs} gg={gg} scheme={scheme} ints={ints} init={init_mode} term={term_mode} "
          f"n={rho.shape[0]} {base}", flush=True)
    pot_g, flu_g, it_g, st_g = pg.solve_host(
        rho, cpot, flu, gs=gs, gg=gg, scheme=scheme, init_mode=init_mode, term_mode=term_mode,
        check_every=check_every, rel_tol=base.get("rel_tol", 1e-5), fuse_gravity=fuse, guess=guess,
        dt=base["dt"], sor_min_iter=base["sor_min_iter"], sor_max_iter=base["sor_max_iter"],
        sor_omega=base["sor_omega"], mg_max_iter=base["mg_max_iter"], mg_npre=base["mg_npre"],
        mg_npost=base["mg_npost"], mg_tol=base["mg_tol"], poi_coeff=POI, int_scheme=ints, p5=base.get("p5", False))
    p = _oracle_params(gs, gg, scheme, ints, init_mode, term_mode, check_every, base)
    pot_o, flu_o, it_o, _ = O.solve(rho, cpot, flu, p, guess=guess)
    return (pot_g, flu_g, it_g, st_g), (pot_o, flu_o, it_o)


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("ints", [pg.INT_CQUAD, pg.INT_QUAD])
def test_sor_fixed_iterations_bit_exact(gs, ints):
    rho, cpot, flu = _mixed_inputs(24, gs, seed=gs)
    (pg_pot, pg_flu, it, st), (o_pot, o_flu, o_it) = _both(rho, cpot, flu, gs, ints=ints)
    assert np.all(it == 50) and np.all(o_it == 50)
    assert _rel_linf(pg_pot, o_pot) <= TOL
    assert np.array_equal(pg_pot, o_pot), f"max rel diff {_rel_linf(pg_pot, o_pot):.3e}"
    assert np.array_equal(pg_flu, o_flu)
    # the stall rule may also fire on the last of the fixed iterations
    assert np.all((st == pg.STATUS_MAX_ITER) | (st == pg.STATUS_CONVERGED))


@pytest.mark.parametrize("gs", [1, 3, 5])
def test_config1_softened_powerlaw(gs):
    """Config-1 inputs: softened power law, omega 1.69, 50 iterations, gravity dt 1e-3."""
    rho, cpot, flu, _ = I.softened_powerlaw(32, gs, seed=7)
    (pg_pot, pg_flu, _, _), (o_pot, o_flu, _) = _both(rho, cpot, flu, gs, sor_omega=1.69, dt=1e-3)
    assert np.array_equal(pg_pot, o_pot)
    assert np.array_equal(pg_flu, o_flu)


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
def test_sor_reference_stall_rule(gs):
    rho, cpot, flu = _mixed_inputs(16, gs, seed=10 + gs)
    (pg_pot, pg_flu, it, st), (o_pot, o_flu, o_it) = _both(rho, cpot, flu, gs, sor_min_iter=10, sor_max_iter=60)
    assert np.all(it >= 10) and np.all(it <= 60)
    assert np.abs(it - o_it).max() <= 4
    same = it == o_it
    assert same.mean() >= 0.5
    # stalled solutions sit at FP32 round-off: even differing counts agree to the tolerance
    assert _rel_linf(pg_pot, o_pot) <= TOL
    assert _rel_linf(pg_pot[same], o_pot[same]) <= TOL


@pytest.mark.parametrize("gs", [1, 3, 5])
@pytest.mark.parametrize("check_every", [1, 2])
def test_sor_relres_zero_guess(gs, check_every):
    rho, cpot, flu, _ = I.plummer(16, gs, seed=30 + gs)
    (pg_pot, _, it, st), (o_pot, _, o_it) = _both(rho, cpot, flu, gs, init_mode=pg.INIT_ZERO,
                                                  term_mode=pg.TERM_RELRES, check_every=check_every,
                                                  sor_min_iter=10, sor_max_iter=1000, sor_omega=1.8,
                                                  rel_tol=1e-5)
    assert np.all(st == pg.STATUS_CONVERGED)
    assert np.abs(it - o_it).max() <= check_every
    same = it == o_it
    assert same.mean() >= 0.75
    assert _rel_linf(pg_pot[same], o_pot[same]) <= TOL
    # every returned field meets the tolerance when re-measured in FP64 on the CPU
    for P in range(rho.shape[0]):
        p = O.make_params(gs=gs, init_mode=O.INIT_ZERO, sor_min_iter=int(it[P]), sor_max_iter=int(it[P]),
                          sor_omega=1.8, poi_coeff=POI, graacc=False)
        _, _, _, rr = O.solve(rho[P:P + 1], cpot[P:P + 1], None, p)
        assert rr[0] <= 1.05e-5


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("ints", [pg.INT_CQUAD, pg.INT_QUAD])
def test_mg_fixed_vcycles_bit_exact(gs, ints):
    rho, cpot, flu = _mixed_inputs(16, gs, seed=50 + gs)
    (pg_pot, pg_flu, it, _), (o_pot, o_flu, o_it) = _both(rho, cpot, flu, gs, scheme=pg.MG, ints=ints, mg_max_iter=4,
                                                          mg_tol=0.0, mg_npre=2, mg_npost=2)
    assert np.all(it == 4) and np.all(o_it == 4)
    assert _rel_linf(pg_pot, o_pot) <= TOL
    assert np.array_equal(pg_pot, o_pot), f"max rel diff {_rel_linf(pg_pot, o_pot):.3e}"
    assert np.array_equal(pg_flu, o_flu)


@pytest.mark.parametrize("gs", [2, 5])
def test_mg_reference_tolerance(gs):
    rho, cpot, flu = _mixed_inputs(16, gs, seed=60 + gs)
    (pg_pot, _, it, st), (o_pot, _, o_it) = _both(rho, cpot, flu, gs, scheme=pg.MG, mg_max_iter=10, mg_tol=1e-5)
    assert np.all(st == pg.STATUS_CONVERGED)
    same = it == o_it
    assert same.mean() >= 0.75
    assert _rel_linf(pg_pot[same], o_pot[same]) <= TOL


@pytest.mark.parametrize("scheme", [pg.SOR, pg.MG])
def test_init_modes_zero_and_guess(scheme):
    gs = 4
    rho, cpot, flu = _mixed_inputs(16, gs, seed=70)
    guess = (np.random.default_rng(0).standard_normal(rho.shape) * 1e-5).astype(np.float32)
    for mode in (pg.INIT_ZERO, pg.INIT_GUESS):
        (pg_pot, _, _, _), (o_pot, _, _) = _both(rho, cpot, flu, gs, scheme=scheme, init_mode=mode,
                                                 guess=guess if mode == pg.INIT_GUESS else None,
                                                 sor_min_iter=20, sor_max_iter=20, mg_max_iter=2, mg_tol=0.0)
        assert np.array_equal(pg_pot, o_pot)


@pytest.mark.parametrize("gs", [2, 5])
@pytest.mark.parametrize("fuse", [True, False])
def test_gravity_five_point_and_unfused(gs, fuse):
    rho, cpot, flu = _mixed_inputs(16, gs, seed=80)
    (pg_pot, pg_flu, _, _), (o_pot, o_flu, _) = _both(rho, cpot, flu, gs, gg=2, p5=True, fuse=fuse,
                                                      sor_min_iter=20, sor_max_iter=20)
    assert np.array_equal(pg_pot, o_pot)
    assert np.array_equal(pg_flu, o_flu)


def test_gravity_only_path():
    """Poisson = false: the caller's potential drives the gravity kernel (CUAPI_Asyn...cu:192-200)."""
    gs = 3
    rho, cpot, flu = _mixed_inputs(16, gs, seed=90)
    pot_in = np.random.default_rng(2).standard_normal((16, 10, 10, 10)).astype(np.float32)
    pot_g, flu_g, _, _ = pg.solve_host(None, None, flu, gs=gs, pot_out=pot_in, poisson=False, dt=0.3, poi_coeff=POI)
    _, flu_o, _, _ = O.solve(None, None, flu, O.make_params(gs=gs, poisson=False, dt=0.3), pot_out=pot_in)
    assert np.array_equal(pot_g, pot_in)
    assert np.array_equal(flu_g, flu_o)


def test_reference_call_shape_host_arrays_multi_stream():
    """pgamd_asyn_poisson_gravity_solver with pinned host buffers, 3 streams, then synchronize."""
    gs = 5
    rho, cpot, flu = _mixed_inputs(24, gs, seed=95)
    with pg.Context(0) as ctx:
        ctx.set_options(pot_ghost_size=gs, gra_ghost_size=1)
        h_rho, h_cpot, h_out, h_flu = ctx.mem_allocate(3)
        h_rho[:] = rho
        h_cpot[:] = cpot
        h_flu[:] = flu
        ctx.asyn_poisson_gravity_solver(h_rho, h_cpot, h_out, h_flu, 3, 0.5, 1.0, 10, 60, OMEGA[4], 10, 3, 3, 1e-6,
                                        POI, pg.INT_CQUAD, False, True, True, gpu_nstream=3)
        ctx.synchronize()
        pot_g, flu_g = h_out.copy(), h_flu.copy()
        it, st = ctx.patch_stats(24)
        ctx.mem_free()
    p = O.make_params(gs=gs, sor_min_iter=10, sor_max_iter=60, sor_omega=OMEGA[4], poi_coeff=POI, dt=0.5)
    pot_o, flu_o, it_o, _ = O.solve(rho, cpot, flu, p)
    assert _rel_linf(pot_g, pot_o) <= TOL
    assert np.all((it >= 10) & (it <= 60))
    assert np.all(st & (pg.STATUS_CONVERGED | pg.STATUS_MAX_ITER))


def _dev_batch(gs, n, gg=1):
    """Zeroed, correctly sized device arrays (rho, coarse potential, output window, fluid with rho = 1): a call
    that host validation should reject would still only run a harmless solve if that validation regressed."""
    import torch

    g = pg.geometry(gs, gg)
    dev = torch.device("cuda", 0)
    flu = torch.zeros((n, 5, 8, 8, 8), device=dev)
    flu[:, 0] = 1.0
    return (torch.zeros((n, g["R"], g["R"], g["R"]), device=dev), torch.zeros((n, g["C"], g["C"], g["C"]), device=dev),
            torch.zeros((n, g["G"], g["G"], g["G"]), device=dev), flu)


def test_error_paths():
    with pg.Context(0) as ctx:
        with pytest.raises(pg.PgError):
            ctx.set_options(pot_ghost_size=6)
        ctx.set_options(pot_ghost_size=2, gra_ghost_size=1)
        rho, cpot, out, flu = _dev_batch(2, 8)
        with pytest.raises(pg.PgError):  # 5-point gradient needs GRA_GHOST_SIZE = 2
            ctx.solve_device(rho, cpot, out, flu, 8, pg.solver_args(p5=True))
        with pytest.raises(pg.PgError):  # bad interpolation scheme
            ctx.solve_device(rho, cpot, out, flu, 8, pg.solver_args(int_scheme=3))
        with pytest.raises(pg.PgError):  # solve before allocation
            ctx.asyn_poisson_gravity_solver(None, None, None, None, 1, 0.1, 1.0, 1, 1, 1.5, 1, 1, 1, 1e-6, POI, 4,
                                            False, True, True)
        ctx.solve_device(None, None, out, None, 0, pg.solver_args())  # zero patches: no-op


def test_asyn_refuses_ghost_sizes_changed_after_allocation():
    """The host-array entry's device buffers are sized at allocation; solving after pgamd_set_options changed
    a ghost size is refused with PGAMD_ERR_STATE before any copy or launch."""
    with pg.Context(0) as ctx:
        ctx.set_options(pot_ghost_size=1, gra_ghost_size=1)
        h_rho, h_cpot, h_out, h_flu = ctx.mem_allocate(1)      # 8 patches at ghost size 1
        h_rho[:] = 0.0
        h_cpot[:] = 0.0
        h_flu[:] = 1.0
        call = lambda: ctx.asyn_poisson_gravity_solver(h_rho, h_cpot, h_out, h_flu, 1, 1e-3, 1.0, 5, 5, 1.5, 1, 1,  # noqa
                                                        1, 1e-6, POI, pg.INT_CQUAD, False, True, True)
        ctx.set_options(pot_ghost_size=5)
        with pytest.raises(pg.PgError) as e:
            call()
        assert e.value.code == pg.ERR_STATE
        ctx.set_options(pot_ghost_size=2, gra_ghost_size=2)
        with pytest.raises(pg.PgError) as e:
            call()
        assert e.value.code == pg.ERR_STATE
        ctx.set_options(pot_ghost_size=1, gra_ghost_size=1)   # back to the allocated sizes: runs
        print("[case] asyn solve at the allocated ghost size", flush=True)
        call()
        ctx.synchronize()
        assert np.all(h_out == 0.0)                         # zero density and boundary give a zero potential
        ctx.mem_free()


def test_guess_start_without_a_guess_array_is_refused():
    """init_mode GUESS with no guess array: the reference-shaped entry (which has no guess argument) and the
    graph capture refuse it with PGAMD_ERR_ARG; the autotuner times a guess start as the zero start."""
    import torch

    gs, n = 3, 64
    with pg.Context(0) as ctx:
        ctx.set_options(pot_ghost_size=gs, gra_ghost_size=1, init_mode=pg.INIT_GUESS)
        h_rho, h_cpot, h_out, h_flu = ctx.mem_allocate(1)
        h_flu[:] = 1.0
        with pytest.raises(pg.PgError) as e:
            ctx.asyn_poisson_gravity_solver(h_rho, h_cpot, h_out, h_flu, 1, 1e-3, 1.0, 5, 5, 1.5, 1, 1, 1, 1e-6, POI,
                                            pg.INT_CQUAD, False, True, True)
        assert e.value.code == pg.ERR_ARG
        ctx.mem_free()
        rho, cpot, out, flu = _dev_batch(gs, n)
        with pytest.raises(pg.PgError) as e:
            ctx.graph_device(rho, cpot, out, flu, n, pg.solver_args())
        assert e.value.code == pg.ERR_ARG
        r, c, f, _ = I.plummer(n, gs, seed=4)
        rho.copy_(torch.from_numpy(r))
        cpot.copy_(torch.from_numpy(c))
        print("[case] autotune with init GUESS", flush=True)
        ppw, pad, ms = ctx.autotune(gs, rho, cpot, out, n, pg.solver_args(sor_min_iter=10, sor_max_iter=10,
                                                                           poi_coeff=POI, graacc=False), repeats=1)
        assert ms > 0 and ppw in (1, 2, 4) and 0 <= pad <= 3
        zero = pg.solve_host(r, c, None, gs=gs, init_mode=pg.INIT_ZERO, sor_min_iter=10, sor_max_iter=10,
                             poi_coeff=POI, graacc=False)[0]
        assert np.array_equal(out.cpu().numpy(), zero)        # the candidates ran the zero start


@pytest.mark.parametrize("gs", [2, 5])
def test_nonfinite_density_flagged(gs):
    """A NaN density cell makes its patch run to Max_Iter with the non-finite flag; the other patches of the
    batch converge as usual (ghost size 5 runs the compact layout, with the cell in an LDS-held Const*rho)."""
    rho, cpot, flu = _mixed_inputs(16, gs, seed=99)
    rho[5, 3, 3, 3] = np.nan
    if gs == 5:
        rho[9, 14, 2, 7] = np.inf
    pot_g, _, it, st = pg.solve_host(rho, cpot, flu, gs=gs, term_mode=pg.TERM_RELRES, init_mode=pg.INIT_ZERO,
                                     sor_min_iter=5, sor_max_iter=80, sor_omega=1.6, poi_coeff=POI, graacc=False)
    bad = [5, 9] if gs == 5 else [5]
    for b in bad:
        assert st[b] & pg.STATUS_NONFINITE and it[b] == 80
    assert np.all(st[~np.isin(np.arange(16), bad)] == pg.STATUS_CONVERGED)


def test_repeat_is_bitwise_deterministic():
    gs = 5
    rho, cpot, flu, _ = I.lognormal(16, gs, seed=3)
    a = pg.solve_host(rho, cpot, flu, gs=gs, term_mode=pg.TERM_RELRES, init_mode=pg.INIT_ZERO, sor_omega=1.8,
                      sor_min_iter=10, sor_max_iter=1000, poi_coeff=POI)
    b = pg.solve_host(rho, cpot, flu, gs=gs, term_mode=pg.TERM_RELRES, init_mode=pg.INIT_ZERO, sor_omega=1.8,
                      sor_min_iter=10, sor_max_iter=1000, poi_coeff=POI)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("scheme", [pg.SOR, pg.MG])
def test_full_cube_output_window(scheme):
    """A Poisson-only call may widen the output window up to the whole solve cube (GRA_GHOST_SIZE = ghost
    size): its centre equals the usual window bit for bit and its outer layer is the interpolated
    Dirichlet boundary (oracle interpolation).  With the gravity update, the window is limited to 2."""
    gs, n = 4, 12
    rho, cpot, flu = _mixed_inputs(n, gs, seed=7)
    kw = dict(gs=gs, scheme=scheme, sor_min_iter=30, sor_max_iter=30, sor_omega=OMEGA[3], mg_max_iter=4,
              mg_tol=1e-12, poi_coeff=POI, graacc=False)
    print(f"[case] full-cube window scheme={scheme}", flush=True)
    full = pg.solve_host(rho, cpot, None, gg=gs, **kw)[0]
    win = pg.solve_host(rho, cpot, None, gg=1, **kw)[0]
    N = 8 + 2 * gs
    assert full.shape == (n, N, N, N)
    o = gs - 1
    assert np.array_equal(full[:, o:o + 10, o:o + 10, o:o + 10], win)
    cube = O.interp(cpot, gs, pg.INT_CQUAD, dtype=np.float32)
    face = np.ones((N, N, N), bool)
    face[1:-1, 1:-1, 1:-1] = False
    assert np.array_equal(full[:, face], cube[:, face])
    with pg.Context(0) as ctx:
        ctx.set_options(pot_ghost_size=gs, gra_ghost_size=3)
        rho_d, cpot_d, out_d, flu_d = _dev_batch(gs, 8, gg=3)
        with pytest.raises(pg.PgError):  # the gravity stencil needs GRA_GHOST_SIZE <= 2
            ctx.solve_device(rho_d, cpot_d, out_d, flu_d, 8, pg.solver_args(graacc=True))


@pytest.mark.parametrize("init_mode,term_mode,gg,fuse,tol", [
    (pg.INIT_INTERP, pg.TERM_REFERENCE, 1, True, 0.0),
    (pg.INIT_ZERO, pg.TERM_RELRES, 1, True, 0.0),
    (pg.INIT_INTERP, pg.TERM_RELRES, 2, True, 0.0),
    (pg.INIT_ZERO, pg.TERM_REFERENCE, 4, False, 0.0),
    (pg.INIT_ZERO, pg.TERM_REFERENCE, 5, False, 0.0),          # whole-cube window: runs in the split layout
    (pg.INIT_GUESS, pg.TERM_RELRES, 1, True, 1e-2),            # guess check, bad guesses rejected
    (pg.INIT_GUESS_CUBE, pg.TERM_RELRES, 1, True, 0.0),        # no guess check
    (pg.INIT_GUESS_CUBE, pg.TERM_RELRES, 1, True, 1e30),       # guess check, every guess kept
    (pg.INIT_GUESS_CUBE, pg.TERM_REFERENCE, 1, True, 1e-2),    # guess check, every guess rejected
])
def test_compact_layout_matches_split(init_mode, term_mode, gg, fuse, tol):
    """The compact single-buffer LDS cube of ghost size 5 (PGAMD_LAYOUT_COMPACT) and the colour-split cube give
    bitwise the same potential window, fluid, iteration counts and status flags, for every start (the compact
    kernel interpolates by coarse column and face, the split kernel by coarse cell), both stopping rules, output
    windows 1..5, the guess check kept, rejected or off, and both interpolation schemes."""
    import torch

    gs, n = 5, 192
    rho, cpot, flu = _mixed_inputs(n, gs, seed=11)
    g = pg.geometry(gs, gg)
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(5)
    gshape = (n, g["N"], g["N"], g["N"]) if init_mode == pg.INIT_GUESS_CUBE else (n, g["R"], g["R"], g["R"])
    guess = None
    if init_mode in (pg.INIT_GUESS, pg.INIT_GUESS_CUBE):
        # half the patches get a plausible guess (a scaled interpolated cube), half a zero guess
        cube = O.interp(cpot, gs, pg.INT_CQUAD, dtype=np.float32)
        guess = cube if init_mode == pg.INIT_GUESS_CUBE else np.ascontiguousarray(cube[:, 1:-1, 1:-1, 1:-1])
        guess = (guess * rng.uniform(0.9, 1.1, size=(n, 1, 1, 1))).astype(np.float32)
        guess[::2] = 0.0
        assert guess.shape == gshape
    results = {}
    for ints in (pg.INT_CQUAD, pg.INT_QUAD):
        args = pg.solver_args(dt=0.5, sor_min_iter=10, sor_max_iter=300, sor_omega=1.8, poi_coeff=POI,
                              int_scheme=ints, graacc=fuse)
        for lay in (1, pg.LAYOUT_COMPACT):
            print(f"[case] compact vs split: init={init_mode} term={term_mode} gg={gg} fuse={fuse} tol={tol} "
                  f"ints={ints} layout pad {lay}", flush=True)
            with pg.Context(0) as ctx:
                ctx.set_options(pot_ghost_size=gs, gra_ghost_size=gg, init_mode=init_mode, term_mode=term_mode,
                                check_every=2, rel_tol=1e-5, fuse_gravity=int(fuse), guess_fallback_tol=tol)
                ctx.set_layout(gs, 1, lay)
                d_rho = torch.from_numpy(rho).to(dev)
                d_cpot = torch.from_numpy(cpot).to(dev)
                d_flu = torch.from_numpy(flu).to(dev)
                d_out = torch.zeros((n, g["G"], g["G"], g["G"]), device=dev)
                d_guess = torch.from_numpy(guess).to(dev) if guess is not None else None
                d_it = torch.zeros(n, dtype=torch.int32, device=dev)
                d_st = torch.zeros(n, dtype=torch.int32, device=dev)
                ctx.solve_device(d_rho, d_cpot, d_out, d_flu, n, args, guess=d_guess, iters=d_it, status=d_st)
                torch.cuda.synchronize()
                results[(ints, lay)] = [t.cpu().numpy() for t in (d_out, d_flu, d_it, d_st)]
        for a, b in zip(results[(ints, 1)], results[(ints, pg.LAYOUT_COMPACT)]):
            assert np.array_equal(a, b)
    st = results[(pg.INT_CQUAD, pg.LAYOUT_COMPACT)][3]
    if tol == 1e-2:
        assert ((st & pg.STATUS_FALLBACK) != 0).sum() >= n // 2      # every zero guess is rejected
    if tol == 1e30:   # only the zero guesses (relative residual infinite) are rejected
        assert ((st & pg.STATUS_FALLBACK) != 0).sum() == n // 2


def test_autotune_ghost_size_5_considers_the_compact_layout(tmp_path):
    """The autotuner's search at ghost size 5 includes the compact cube (PGAMD_LAYOUT_COMPACT); whichever layout
    it keeps, the solve is bitwise the default layout's, and a saved table with the compact layout loads back."""
    import torch

    gs, n = 5, 512
    rho, cpot, flu, _ = I.plummer(n, gs, seed=21)
    dev = torch.device("cuda", 0)
    d_rho, d_cpot = torch.from_numpy(rho).to(dev), torch.from_numpy(cpot).to(dev)
    g = pg.geometry(gs)
    args = pg.solver_args(sor_min_iter=20, sor_max_iter=20, poi_coeff=POI, graacc=False)

    def solve(ctx):
        out = torch.zeros((n, g["G"], g["G"], g["G"]), device=dev)
        ctx.solve_device(d_rho, d_cpot, out, None, n, args)
        torch.cuda.synchronize()
        return out.cpu().numpy()

    with pg.Context(0) as ctx:
        ctx.set_options(pot_ghost_size=gs, init_mode=pg.INIT_INTERP, fuse_gravity=0)
        assert ctx.get_layout(gs) == (1, pg.LAYOUT_COMPACT)      # the default at ghost size 5
        base = solve(ctx)
        out = torch.zeros((n, g["G"], g["G"], g["G"]), device=dev)
        print("[case] autotune gs=5", flush=True)
        ppw, pad, ms = ctx.autotune(gs, d_rho, d_cpot, out, n, args, repeats=2)
        assert ms > 0 and ppw == 1 and pad in (0, 1, 2, 3, pg.LAYOUT_COMPACT)
        assert np.array_equal(solve(ctx), base)
        ctx.set_layout(gs, 1, pg.LAYOUT_COMPACT)
        path = str(tmp_path / "layouts.json")
        ctx.autotune_save(path)
    with pg.Context(0) as c2:
        c2.set_layout(gs, 1, 0)
        c2.autotune_load(path)
        assert c2.get_layout(gs) == (1, pg.LAYOUT_COMPACT)


def test_layout_occupancy():
    """pgamd_layout_occupancy: resident SOR workgroups per CU at every ghost size and layout, consistent with the
    layout's patches per wavefront; at ghost size 5 and one patch per wavefront, LDS (the colour-split cube) and
    registers allow between 4 and 8 resident workgroups, and the compact cube (16 KB) reaches the register
    limit of 8.  The compact layout exists only for ghost size 5 with one patch per wavefront."""
    with pg.Context(0) as ctx:
        for gs in range(1, 6):
            ppw, pad = ctx.get_layout(gs)
            wg, patches = ctx.layout_occupancy(gs)
            print(f"[occupancy] gs={gs} layout ppw={ppw} pad={pad}: {wg} workgroups, {patches} patches per CU",
                  flush=True)
            assert 1 <= wg <= 32 and patches == wg * ppw
        ctx.set_layout(5, 1, 0)
        wg, patches = ctx.layout_occupancy(5)
        assert wg == patches and 4 <= wg <= 8
        ctx.set_layout(5, 1, pg.LAYOUT_COMPACT)
        wg, patches = ctx.layout_occupancy(5)
        print(f"[occupancy] gs=5 compact layout: {wg} workgroups per CU", flush=True)
        assert wg == patches == 8
        for gs, ppw in ((4, 1), (3, 1), (1, 4)):
            with pytest.raises(pg.PgError):
                ctx.set_layout(gs, ppw, pg.LAYOUT_COMPACT)
        with pytest.raises(pg.PgError):
            ctx.layout_occupancy(6)
