"""Patch-group preparation and scatter (pgamd_prepare_patch_groups / pgamd_scatter_patch_groups): the GAMER-side
step that the reference testbed replaces by synthetic arrays (GPU_API/CUAPI_Asyn_PoissonGravitySolver.cu:37-41,
CPU_Solver/CPU_PoissonSolver_SOR.cpp:5-7, 70-76).  Host code; runs without a GPU.

The packed arrays are checked against two formulations that share nothing with the neighbour-table gather:
* slices of the global grids the levels were cut from (zero padding outside the refined box or the domain);
* linear fields, whose values at the solver's cell centres (inputs.fine_centres / coarse_centres, the frame the
  interpolation and the analytic tests assume) are known in closed form.
"""
import numpy as np
import pytest

import gpu_testbed_poisson_gravity_amd as pg
from gpu_testbed_poisson_gravity_amd import inputs as I


def _level(ncp, m, origin, seed, fine=None, coarse=None):
    rng = np.random.default_rng(seed)
    L, F = 8 * ncp, 16 * m
    if fine is None:
        fine = rng.standard_normal((5, F, F, F)).astype(np.float32)
    if coarse is None:
        coarse = rng.standard_normal((L, L, L)).astype(np.float32)
    return fine, coarse, I.two_level_box(fine, coarse, origin, seed=seed)


def _prepare(gs, d, **kw):
    return pg.prepare_patch_groups(gs, d["group_first"], d["group_father"], d["fine_flu"], d["fine_sib"],
                                   d["coarse_pot"], d["coarse_sib"], **kw)


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
def test_prepare_matches_global_slices(gs):
    ncp, m, origin = 3, 2, (0, 1, 0)       # refined box on the domain edge in z and x: coarse cells go missing
    fine, coarse, d = _level(ncp, m, origin, seed=gs)
    rho, cpot, flu, st = _prepare(gs, d)
    g = pg.geometry(gs)
    gf, gc = gs - 1, (g["C"] - 4) // 2
    F, L = fine.shape[1], coarse.shape[0]
    fpad = np.pad(fine[0], 8)
    cpad = np.pad(coarse, 8)
    o = np.asarray(origin) * 8
    for p, (z, y, x) in enumerate(d["fine_coords"]):
        lo = np.array([z, y, x]) * 8 - gf
        want = fpad[8 + lo[0]:8 + lo[0] + g["R"], 8 + lo[1]:8 + lo[1] + g["R"], 8 + lo[2]:8 + lo[2] + g["R"]]
        assert np.array_equal(rho[p], want)
        clo = o + np.array([z, y, x]) * 4 - gc
        cwant = cpad[8 + clo[0]:8 + clo[0] + g["C"], 8 + clo[1]:8 + clo[1] + g["C"], 8 + clo[2]:8 + clo[2] + g["C"]]
        assert np.array_equal(cpot[p], cwant)
        assert np.array_equal(flu[p], fine[:, 8 * z:8 * z + 8, 8 * y:8 * y + 8, 8 * x:8 * x + 8])
        miss_f = bool(np.any(lo < 0) or np.any(lo + g["R"] > F))
        miss_c = bool(np.any(clo < 0) or np.any(clo + g["C"] > L))
        assert st[p] == (pg.PACK_MISSING_FINE if miss_f else 0) | (pg.PACK_MISSING_COARSE if miss_c else 0)
    assert np.any(st & pg.PACK_MISSING_COARSE) and (gs == 1 or np.any(st & pg.PACK_MISSING_FINE))
    assert gs > 1 or not np.any(st & pg.PACK_MISSING_FINE)   # no fine ghost cells at gs = 1


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
def test_prepare_linear_fields_at_solver_centres(gs):
    """Density a + b.x on the fine level and potential c + d.x on the coarse level (x the global cell centres,
    spacings 1 and 2): every packed cell equals the field at its solver-frame centre plus the patch origin."""
    ncp, m, origin = 4, 2, (1, 1, 1)
    F, L = 16 * m, 8 * ncp
    a, b = 0.25, np.array([0.5, -0.75, 1.25])     # b over (z, y, x)
    c, dv = -1.5, np.array([0.125, 0.375, -0.625])
    fz = (np.arange(F) + 0.5)
    fine = np.zeros((5, F, F, F), np.float32)
    fine[0] = a + b[0] * fz[:, None, None] + b[1] * fz[None, :, None] + b[2] * fz[None, None, :]
    cz = (np.arange(L) + 0.5) * 2.0 - np.asarray(origin)[:, None] * 16.0    # coarse centres in the fine box frame
    coarse = (c + dv[0] * cz[0][:, None, None] + dv[1] * cz[1][None, :, None] + dv[2] * cz[2][None, None, :])
    coarse = coarse.astype(np.float32)
    _, _, d = _level(ncp, m, origin, seed=10 + gs, fine=fine, coarse=coarse)
    rho, cpot, _, st = _prepare(gs, d)
    assert not np.any(st & pg.PACK_MISSING_COARSE)
    fcen, ccen = I.fine_centres(gs), I.coarse_centres(gs)
    for p, org in enumerate(d["fine_coords"] * 8.0):
        pz, py, px = (org[k] + fcen for k in range(3))
        want = a + b[0] * pz[:, None, None] + b[1] * py[None, :, None] + b[2] * px[None, None, :]
        inside = ((pz >= 0) & (pz < F))[:, None, None] & ((py >= 0) & (py < F))[None, :, None] & \
                 ((px >= 0) & (px < F))[None, None, :]
        np.testing.assert_allclose(rho[p][inside], want[inside], rtol=0, atol=1e-4)
        assert np.all(rho[p][~inside] == 0.0)
        qz, qy, qx = (org[k] + ccen for k in range(3))
        cwant = c + dv[0] * qz[:, None, None] + dv[1] * qy[None, :, None] + dv[2] * qx[None, None, :]
        np.testing.assert_allclose(cpot[p], cwant, rtol=0, atol=1e-4)


def test_prepare_independent_of_storage_order_and_threads():
    gs = 5
    fine, coarse, d1 = _level(3, 2, (1, 0, 1), seed=3)
    d2 = I.two_level_box(fine, coarse, (1, 0, 1), seed=99)
    out1 = _prepare(gs, d1, nthreads=1)
    out8 = _prepare(gs, d1, nthreads=8)
    for a, b in zip(out1, out8):
        assert np.array_equal(a, b)
    out2 = _prepare(gs, d2)
    key = lambda c: (c[:, 0] * 64 + c[:, 1]) * 64 + c[:, 2]   # noqa: E731
    i1, i2 = np.argsort(key(d1["fine_coords"])), np.argsort(key(d2["fine_coords"]))
    assert not np.array_equal(i1, i2)
    for a, b in zip(out1, out2):
        assert np.array_equal(a[i1], b[i2])


def test_prepare_into_given_buffers():
    gs = 4
    _, _, d = _level(2, 1, (0, 0, 0), seed=4)
    g = pg.geometry(gs)
    n = 8 * d["group_first"].size
    hr = np.full(n * g["R"] ** 3 + 7, -9.0, np.float32)       # flat, larger than needed (pinned-buffer shape)
    hc = np.full(n * g["C"] ** 3, -9.0, np.float32)
    hf = np.full(n * 5 * 512, -9.0, np.float32)
    _prepare(gs, d, h_rho=hr, h_pot_in=hc, h_flu=hf)
    rho, cpot, flu, _ = _prepare(gs, d)
    assert np.array_equal(hr[:rho.size], rho.ravel()) and np.all(hr[rho.size:] == -9.0)
    assert np.array_equal(hc, cpot.ravel()) and np.array_equal(hf, flu.ravel())


@pytest.mark.parametrize("gg", [1, 2])
def test_scatter_round_trip(gg):
    gs = 5
    fine, coarse, d = _level(3, 2, (0, 0, 1), seed=7)
    _, _, h_flu, _ = _prepare(gs, d)
    G = pg.geometry(gs, gg)["G"]
    nfine = d["fine_flu"].shape[0]
    # potential windows cut from a global fine potential; the scatter must return the patch parts
    rng = np.random.default_rng(8)
    F = fine.shape[1]
    phi = rng.standard_normal((F, F, F)).astype(np.float32)
    ppad = np.pad(phi, 8)
    pot_out = np.stack([ppad[8 + 8 * z - gg:8 + 8 * z + 8 + gg, 8 + 8 * y - gg:8 + 8 * y + 8 + gg,
                             8 + 8 * x - gg:8 + 8 * x + 8 + gg] for z, y, x in d["fine_coords"]])
    assert pot_out.shape == (nfine, G, G, G)
    h_flu = h_flu + 1.0                         # the "updated" fluid
    # only every other group is scattered; the rest of the level must stay as it was
    groups = d["group_first"][::2].copy()
    fine_pot = np.full((nfine, 8, 8, 8), 7.0, np.float32)
    fine_flu = d["fine_flu"].copy()
    sel = np.concatenate([np.arange(8 * i, 8 * i + 8) for i in range(0, d["group_first"].size, 2)])
    pg.scatter_patch_groups(gg, groups, np.ascontiguousarray(pot_out[sel]), np.ascontiguousarray(h_flu[sel]),
                            fine_pot=fine_pot, fine_flu=fine_flu)
    for p, (z, y, x) in enumerate(d["fine_coords"]):
        if p in set(sel.tolist()):
            assert np.array_equal(fine_pot[p], phi[8 * z:8 * z + 8, 8 * y:8 * y + 8, 8 * x:8 * x + 8])
            assert np.array_equal(fine_flu[p], d["fine_flu"][p] + 1.0)
        else:
            assert np.all(fine_pot[p] == 7.0) and np.array_equal(fine_flu[p], d["fine_flu"][p])


def test_pack_error_paths():
    lib = pg.load()
    gs = 3
    _, _, d = _level(2, 1, (0, 0, 0), seed=5)
    with pytest.raises(pg.PgError, match="POT_GHOST_SIZE"):
        _prepare(6, d)
    with pytest.raises(pg.PgError, match="POT_GHOST_SIZE"):
        _prepare(0, d)
    bad = dict(d, group_first=np.array([d["fine_flu"].shape[0] - 7], np.int32), group_father=np.array([0], np.int32))
    with pytest.raises(pg.PgError, match="outside the level"):
        _prepare(gs, bad)
    bad = dict(d, group_father=np.full_like(d["group_father"], d["coarse_pot"].shape[0]))
    with pytest.raises(pg.PgError, match="outside the level"):
        _prepare(gs, bad)
    sib = d["fine_sib"].copy()
    sib[sib < 0] = d["fine_flu"].shape[0] + 3
    with pytest.raises(pg.PgError, match="neighbour table"):
        _prepare(gs, dict(d, fine_sib=sib))
    with pytest.raises(TypeError):
        _prepare(gs, dict(d, fine_flu=d["fine_flu"].astype(np.float64)))
    # null arrays straight through the C ABI
    assert lib.pgamd_prepare_patch_groups(gs, 1, None, None, 8, None, None, 1, None, None, None, None, None, None,
                                          1) == pg.ERR_ARG
    assert b"null" in lib.pgamd_last_error()
    assert lib.pgamd_scatter_patch_groups(1, 1, d["group_first"].ctypes.data, None, None, 8, None, None, 1) == pg.ERR_ARG
    assert lib.pgamd_scatter_patch_groups(6, 1, d["group_first"].ctypes.data, None, None, 8, None, None,
                                          1) == pg.ERR_ARG
    assert b"GRA_GHOST_SIZE" in lib.pgamd_last_error()
    # empty batches are no-ops
    assert lib.pgamd_prepare_patch_groups(gs, 0, None, None, 0, None, None, 0, None, None, None, None, None, None,
                                          1) == 0
    assert lib.pgamd_scatter_patch_groups(1, 0, None, None, None, 0, None, None, 1) == 0


def test_python_binding_checks_buffers():
    """The binding refuses buffers the C side would overrun: too small, non-contiguous, or of the wrong type."""
    gs = 2
    _, _, d = _level(2, 1, (0, 0, 0), seed=6)
    g = pg.geometry(gs)
    n = 8 * d["group_first"].size
    with pytest.raises(ValueError):
        _prepare(gs, d, h_rho=np.zeros(n * g["R"] ** 3 - 1, np.float32))
    with pytest.raises(ValueError):
        _prepare(gs, d, h_rho=np.zeros((n * g["R"] ** 3, 2), np.float32)[:, 0])    # strided view
    with pytest.raises(ValueError):
        _prepare(gs, d, h_flu=np.zeros(n * 5 * 512, np.float64))
    fine_pot = np.zeros((d["fine_flu"].shape[0], 8, 8, 8), np.float32)
    with pytest.raises(ValueError):
        pg.scatter_patch_groups(1, d["group_first"], np.zeros(n * 1000 - 1, np.float32), None, fine_pot=fine_pot)
    with pytest.raises(ValueError):
        pg.scatter_patch_groups(1, d["group_first"], None, None)
    pg.scatter_patch_groups(1, d["group_first"], np.zeros(n * 1000, np.float32), None, fine_pot=fine_pot)
    assert np.all(fine_pot == 0)
