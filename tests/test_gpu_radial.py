"""The device solver's converged potential against closed-form solutions (needs an MI355X).

Output_Radial of the reference testbed (Main.cpp:593-710) compares the solver's potential on the output
window with the analytic power law A r^(2-n) of its initial condition (phi_anal, :679-708).  Here the HIP
kernels solve in FP32 to the workloads' tolerance (relative residual 1e-5 for SOR, MG error 1e-5) and the potential of the 8^3 patch proper is held to
the same bounds the FP64 oracle meets (tests/test_oracle_math.py): the discretisation error of the 7-point
stencil and of the interpolated boundary, not solver error, sets them.  Plummer spheres and the softened
power law of the other workloads are checked against their own closed forms the same way.
"""
import numpy as np
import pytest

import gpu_testbed_poisson_gravity_amd as pg
from gpu_testbed_poisson_gravity_amd import inputs as I

pytestmark = pytest.mark.gpu
POI = 4.0 * np.pi * I.NEWTON_G


def _solve(rho, cpot, flu, gs, scheme=pg.SOR):
    print(f"[case] radial gs={gs} scheme={scheme} n={rho.shape[0]}", flush=True)
    pot, _, it, st = pg.solve_host(rho, cpot, flu, gs=gs, scheme=scheme, init_mode=pg.INIT_INTERP,
                                   term_mode=pg.TERM_RELRES, rel_tol=1e-5, sor_min_iter=10, sor_max_iter=4000,
                                   sor_omega=1.7, mg_max_iter=40, mg_npre=3, mg_npost=3, mg_tol=1e-5, poi_coeff=POI,
                                   graacc=False)
    assert np.all(st & pg.STATUS_CONVERGED), (it, st)
    return pot.astype(np.float64)


@pytest.mark.parametrize("scheme", [pg.SOR, pg.MG])
@pytest.mark.parametrize("gs", [1, 3, 5])
def test_reference_power_law_matches_analytic(gs, scheme):
    """rho = (r/W)^-n gives phi = A r^(2-n), A = 4 pi G W^n / ((n-3)(n-2)) (Main.cpp:679-708)."""
    rho, cpot, flu = I.reference_ic(8, gs)
    pot = _solve(rho, cpot, flu, gs, scheme)
    err = []
    for P in range(8):
        r, an = I.reference_ic_analytic(gs, 1, P)
        err.append(np.abs(pot[P, 1:9, 1:9, 1:9] - an) / np.abs(an))
    err = np.array(err)
    print(f"[radial] power law gs={gs} scheme={scheme}: per-cell relative error median {np.median(err):.2e} "
          f"max {err.max():.2e}", flush=True)
    assert np.median(err) < 1e-2 and err.max() < 6e-2


@pytest.mark.parametrize("gs", [1, 3, 5])
def test_plummer_and_softened_power_law_match_closed_forms(gs):
    ax = np.arange(8) + 0.5
    zz, yy, xx = np.meshgrid(ax, ax, ax, indexing="ij")
    pts = np.stack([xx, yy, zz], -1)
    rho, cpot, flu, meta = I.plummer(8, gs, seed=3)
    pot = _solve(rho, cpot, flu, gs)
    for P in range(8):
        an = I.plummer_phi(pts, meta, P)
        assert np.abs(pot[P, 1:9, 1:9, 1:9] - an).max() < 0.15 * np.abs(an).max()
    rho, cpot, flu, meta = I.softened_powerlaw(8, gs, seed=3)
    pot = _solve(rho, cpot, flu, gs)
    for P in range(8):
        r = np.sqrt(((pts - meta["centre"][P]) ** 2).sum(-1))
        an = I.softened_powerlaw_phi(r, meta["rho0"][P], 0.5)
        assert np.abs(pot[P, 1:9, 1:9, 1:9] - an).max() < 0.15 * np.abs(an).max()
