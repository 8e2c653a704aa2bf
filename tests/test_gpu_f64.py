"""FP64 device path (pgamd_solve_device_f64, the reference's FLOAT8 build) against the FP64 oracle and the
reference's own FP64 outputs (needs an MI355X).

* Fixed iteration counts (SOR Min_Iter = Max_Iter, MG with an unreachable tolerance): potential and fluid must
  be bitwise those of the FP64 oracle, which is bitwise the reference's FP64 CPU solver.
* tests/golden/ref_gs*.npz: every FP64 entry (SOR and MG, CQUAD and QUAD) re-run at each patch's reference
  iteration count must equal the reference's potential bit for bit.
* Data-dependent stopping (stall rule, relative residual): counts within a few iterations of the oracle's,
  equal counts bitwise equal.
"""
import os

import numpy as np
import pytest

import gpu_testbed_poisson_gravity_amd as pg
import oracle as O
from gpu_testbed_poisson_gravity_amd import inputs as I

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
POI = 4.0 * np.pi * 1e-5
OMEGA = [1.49, 1.57, 1.62, 1.65, 1.69]


def _inputs(n, gs, seed):
    r1, c1, f1 = I.reference_ic(n // 2, gs, dtype=np.float64)
    r2, c2, f2, _ = I.plummer(n - n // 2, gs, seed=seed, dtype=np.float64)
    flu = np.concatenate([f1, f2])
    flu[:, 1:4] = 0.05 * np.sin(np.arange(flu[:, 1:4].size, dtype=np.float64)).reshape(flu[:, 1:4].shape)
    return np.concatenate([r1, r2]), np.concatenate([c1, c2]), flu


def _both(rho, cpot, flu, gs, gg=1, scheme=pg.SOR, ints=pg.INT_CQUAD, init_mode=pg.INIT_INTERP,
          term_mode=pg.TERM_REFERENCE, guess=None, fuse=True, **kw):
    base = dict(sor_min_iter=50, sor_max_iter=50, sor_omega=OMEGA[gs - 1], mg_max_iter=4, mg_npre=3, mg_npost=3,
                mg_tol=0.0, dt=0.5, rel_tol=1e-5, p5=False)
    base.update(kw)
    print(f"[case] f64 gs={gs} gg={gg} scheme={scheme} ints={ints} init={init_mode} term={term_mode} "
          f"n={rho.shape[0]} {base}", flush=True)
    g = pg.solve_host_f64(rho, cpot, flu, gs=gs, gg=gg, scheme=scheme, init_mode=init_mode, term_mode=term_mode,
                          rel_tol=base["rel_tol"], fuse_gravity=fuse, guess=guess, dt=base["dt"],
                          sor_min_iter=base["sor_min_iter"], sor_max_iter=base["sor_max_iter"],
                          sor_omega=base["sor_omega"], mg_max_iter=base["mg_max_iter"], mg_npre=base["mg_npre"],
                          mg_npost=base["mg_npost"], mg_tol=base["mg_tol"], poi_coeff=POI, int_scheme=ints,
                          p5=base["p5"])
    p = O.make_params(gs=gs, gg=gg, scheme=scheme, int_scheme=ints, init_mode=init_mode, term_mode=term_mode,
                      sor_min_iter=base["sor_min_iter"], sor_max_iter=base["sor_max_iter"],
                      sor_omega=base["sor_omega"], mg_max_iter=base["mg_max_iter"], mg_npre=base["mg_npre"],
                      mg_npost=base["mg_npost"], mg_tol=base["mg_tol"], rel_tol=base["rel_tol"], poi_coeff=POI,
                      dt=base["dt"], p5=base["p5"])
    o = O.solve(rho, cpot, flu, p, guess=guess, dtype=np.float64)
    return g, o


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("ints", [pg.INT_CQUAD, pg.INT_QUAD])
@pytest.mark.parametrize("scheme", [pg.SOR, pg.MG])
def test_f64_fixed_counts_bit_exact(gs, ints, scheme):
    rho, cpot, flu = _inputs(12, gs, seed=200 + gs)
    (pot, fl, it, st), (o_pot, o_fl, o_it, _) = _both(rho, cpot, flu, gs, scheme=scheme, ints=ints)
    want = 50 if scheme == pg.SOR else 4
    assert pot.dtype == np.float64 and np.all(it == want) and np.all(o_it == want)
    assert np.array_equal(pot, o_pot), f"max diff {np.abs(pot - o_pot).max():.3e}"
    assert np.array_equal(fl, o_fl)


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
def test_f64_reference_fixtures_at_reference_counts(gs):
    """The reference's FP64 outputs (SOR: stall rule, Min_Iter 10, Max_Iter 100; MG: tolerance 1e-15, at most 20
    V-cycles), re-run with every patch's reference count fixed."""
    d = np.load(os.path.join(GOLDEN, f"ref_gs{gs}.npz"))
    rho, cpot, flu = (d[k].astype(np.float64) for k in ("rho", "cpot", "flu"))
    for name in ("SOR_cquad_f64", "SOR_quad_f64", "MG_cquad_f64", "MG_quad_f64"):
        scheme = pg.SOR if name.startswith("SOR") else pg.MG
        ints = pg.INT_QUAD if "_quad_" in name else pg.INT_CQUAD
        p = O.make_params(gs=gs, scheme=scheme, int_scheme=ints, sor_min_iter=10, sor_max_iter=100,
                          sor_omega=OMEGA[gs - 1], mg_max_iter=20, mg_npre=3, mg_npost=3, mg_tol=1e-15, poi_coeff=POI,
                          dh=1.0, dt=1e-2)
        o_pot, _, o_it, _ = O.solve(rho, cpot, flu, p, dtype=np.float64)
        ref = d[f"{name}__pot"]
        assert np.array_equal(o_pot, ref)
        pot = np.empty_like(ref)
        for k in np.unique(o_it):
            sel = np.nonzero(o_it == k)[0]
            print(f"[case] f64 fixture gs={gs} {name} patches {sel.tolist()} count {int(k)}", flush=True)
            kw = (dict(sor_min_iter=int(k), sor_max_iter=int(k)) if scheme == pg.SOR else
                  dict(mg_max_iter=int(k), mg_tol=0.0))
            g_pot, _, g_it, _ = pg.solve_host_f64(rho[sel], cpot[sel], flu[sel], gs=gs, scheme=scheme, dt=1e-2, dh=1.0,
                                                  sor_omega=OMEGA[gs - 1], mg_npre=3, mg_npost=3, poi_coeff=POI,
                                                  int_scheme=ints, **kw)
            assert np.all(g_it == k)
            pot[sel] = g_pot
        assert np.array_equal(pot, ref), f"gs={gs} {name}: differs from the reference's FP64 output"


@pytest.mark.parametrize("gs", [2, 5])
def test_f64_stopping_rules(gs):
    rho, cpot, flu = _inputs(16, gs, seed=300 + gs)
    (pot, _, it, st), (o_pot, _, o_it, _) = _both(rho, cpot, flu, gs, sor_min_iter=10, sor_max_iter=100)
    assert np.abs(it - o_it).max() <= 4 and np.mean(it == o_it) >= 0.5
    same = it == o_it
    assert np.array_equal(pot[same], o_pot[same])
    (pot, _, it, st), (o_pot, _, o_it, _) = _both(rho, cpot, flu, gs, init_mode=pg.INIT_ZERO,
                                                  term_mode=pg.TERM_RELRES, sor_min_iter=10, sor_max_iter=1000,
                                                  sor_omega=1.8, rel_tol=1e-10)
    assert np.all(st == pg.STATUS_CONVERGED)
    assert np.abs(it - o_it).max() <= 1
    same = it == o_it
    assert same.mean() >= 0.75 and np.array_equal(pot[same], o_pot[same])
    (pot, _, it, st), (o_pot, _, o_it, _) = _both(rho, cpot, flu, gs, scheme=pg.MG, mg_max_iter=20, mg_tol=1e-12)
    assert np.all(st == pg.STATUS_CONVERGED) and np.abs(it - o_it).max() <= 1
    same = it == o_it
    assert np.array_equal(pot[same], o_pot[same])


@pytest.mark.parametrize("scheme", [pg.SOR, pg.MG])
def test_f64_init_modes_gravity_paths(scheme):
    gs = 4
    rho, cpot, flu = _inputs(12, gs, seed=400)
    guess = np.random.default_rng(1).standard_normal(rho.shape) * 1e-4
    for mode in (pg.INIT_ZERO, pg.INIT_GUESS):
        (pot, fl, _, _), (o_pot, o_fl, _, _) = _both(rho, cpot, flu, gs, scheme=scheme, init_mode=mode,
                                                     guess=guess if mode == pg.INIT_GUESS else None,
                                                     sor_min_iter=20, sor_max_iter=20, mg_max_iter=2)
        assert np.array_equal(pot, o_pot) and np.array_equal(fl, o_fl)
    for fuse in (True, False):   # 5-point gradient, fused and separate
        (pot, fl, _, _), (o_pot, o_fl, _, _) = _both(rho, cpot, flu, gs, gg=2, scheme=scheme, fuse=fuse, p5=True,
                                                     sor_min_iter=20, sor_max_iter=20, mg_max_iter=2)
        assert np.array_equal(pot, o_pot) and np.array_equal(fl, o_fl)


def test_f64_gravity_only_path():
    gs = 3
    _, _, flu = _inputs(8, gs, seed=500)
    pot_in = np.random.default_rng(2).standard_normal((8, 10, 10, 10))
    _, fl, _, _ = pg.solve_host_f64(None, None, flu, gs=gs, pot_out=pot_in, poisson=False, dt=0.3, poi_coeff=POI)
    _, o_fl, _, _ = O.solve(None, None, flu, O.make_params(gs=gs, poisson=False, dt=0.3), pot_out=pot_in,
                            dtype=np.float64)
    assert np.array_equal(fl, o_fl)


def test_f64_error_against_fp32():
    """The FP32 device solve at relative residual 1e-5 against the FP64 device solve converged far below it: the
    per-patch relative L-infinity error is of the order of the tolerance."""
    gs = 5
    rho, cpot, flu, _ = I.plummer(32, gs, seed=9)
    kw = dict(gs=gs, init_mode=pg.INIT_ZERO, term_mode=pg.TERM_RELRES, sor_min_iter=10, sor_max_iter=4000,
              sor_omega=1.8, poi_coeff=POI, graacc=False)
    p64, _, _, st64 = pg.solve_host_f64(rho.astype(np.float64), cpot.astype(np.float64), None, rel_tol=1e-12, **kw)
    p32, _, _, st32 = pg.solve_host(rho, cpot, None, rel_tol=1e-5, **kw)
    assert np.all(st64 == pg.STATUS_CONVERGED) and np.all(st32 == pg.STATUS_CONVERGED)
    err = np.abs(p32 - p64).reshape(32, -1).max(1) / np.abs(p64).reshape(32, -1).max(1)
    print(f"[f64] FP32 at 1e-5 vs FP64 at 1e-12: relative Linf max {err.max():.2e} p99 {np.percentile(err, 99):.2e}",
          flush=True)
    assert err.max() < 1e-4
