"""Two AMR levels through the whole host-side cycle (needs an MI355X): pgamd_prepare_patch_groups gathers the
patch groups into the pinned buffers, pgamd_asyn_poisson_gravity_solver solves them, pgamd_scatter_patch_groups
writes the potential and the updated fluid back into the fine level.

The level is a Plummer sphere (M = 1, a = 4 cells) at the centre of a 64^3 domain: the coarse level holds the
closed-form potential -G M / sqrt(r^2 + a^2) at its cell centres, the refined 32^3 box around the centre the
closed-form density.  At fixed SOR counts the device result equals the FP32 oracle run on the packed arrays bit
for bit; converged (relative residual 1e-5), the assembled fine potential matches the closed form to the
discretisation error the oracle shows on the same problem (max 0.92 % of max|phi|, median 6e-4 relative).
"""
import numpy as np
import pytest

import gpu_testbed_poisson_gravity_amd as pg
import oracle as O
from gpu_testbed_poisson_gravity_amd import inputs as I

pytestmark = pytest.mark.gpu
POI = 4.0 * np.pi * I.NEWTON_G
OMEGA = [1.49, 1.57, 1.62, 1.65, 1.69]


def _plummer_level(ncp=4, m=2, origin=(1, 1, 1), M=1.0, a=4.0):
    L, F = 8 * ncp, 16 * m
    c = 8.0 * ncp                                            # domain centre, fine units
    f = 16.0 * np.asarray(origin)[:, None] + np.arange(F)[None, :] + 0.5
    r2 = (f[0][:, None, None] - c) ** 2 + (f[1][None, :, None] - c) ** 2 + (f[2][None, None, :] - c) ** 2
    fine = np.zeros((5, F, F, F))
    fine[0] = 3.0 * M / (4.0 * np.pi * a ** 3) * (1.0 + r2 / a ** 2) ** -2.5
    fine[1:4] = 0.0                                          # at rest: any gravity update shows
    fine[4] = 1.0
    X = 2.0 * np.arange(L) + 1.0
    R2 = (X[:, None, None] - c) ** 2 + (X[None, :, None] - c) ** 2 + (X[None, None, :] - c) ** 2
    coarse = -I.NEWTON_G * M / np.sqrt(R2 + a ** 2)
    phi = -I.NEWTON_G * M / np.sqrt(r2 + a ** 2)
    return fine.astype(np.float32), coarse.astype(np.float32), phi


def _cycle(gs, d, term_mode, sor_min, sor_max, dt=0.5):
    ng = d["group_first"].size
    print(f"[case] pack cycle gs={gs} groups={ng} term={term_mode} sor {sor_min}..{sor_max}", flush=True)
    nfine = d["fine_flu"].shape[0]
    fine_pot = np.zeros((nfine, 8, 8, 8), np.float32)
    fine_flu = d["fine_flu"].copy()
    with pg.Context(0) as ctx:
        ctx.set_options(pot_ghost_size=gs, gra_ghost_size=1, term_mode=term_mode, rel_tol=1e-5)
        h_rho, h_cpot, h_out, h_flu = ctx.mem_allocate(ng)
        _, _, _, st = pg.prepare_patch_groups(gs, d["group_first"], d["group_father"], d["fine_flu"], d["fine_sib"],
                                              d["coarse_pot"], d["coarse_sib"], h_rho=h_rho, h_pot_in=h_cpot,
                                              h_flu=h_flu)
        packed = (h_rho.copy(), h_cpot.copy(), h_flu.copy())
        ctx.asyn_poisson_gravity_solver(h_rho, h_cpot, h_out, h_flu, ng, dt, 1.0, sor_min, sor_max, OMEGA[gs - 1], 10,
                                        3, 3, 1e-6, POI, pg.INT_CQUAD, False, True, True, gpu_nstream=2)
        ctx.synchronize()
        it, pst = ctx.patch_stats(8 * ng)
        pg.scatter_patch_groups(1, d["group_first"], h_out, h_flu, fine_pot=fine_pot, fine_flu=fine_flu)
        ctx.mem_free()
    return packed, st, fine_pot, fine_flu, it, pst


def _assemble(d, fine_pot):
    F = 16 * round((d["fine_coords"].shape[0] // 8) ** (1 / 3))
    glob = np.zeros((F, F, F), np.float32)
    for q, (z, y, x) in enumerate(d["fine_coords"]):
        glob[8 * z:8 * z + 8, 8 * y:8 * y + 8, 8 * x:8 * x + 8] = fine_pot[q]
    return glob


@pytest.mark.parametrize("gs", [1, 3, 5])
def test_pack_solve_scatter_fixed_counts_bitwise(gs):
    fine, coarse, _ = _plummer_level()
    d = I.two_level_box(fine, coarse, (1, 1, 1), seed=gs)
    (rho, cpot, flu), st, fine_pot, fine_flu, it, _ = _cycle(gs, d, pg.TERM_REFERENCE, 40, 40)
    assert np.all(it == 40)
    assert np.any(st & pg.PACK_MISSING_FINE) == (gs > 1) and not np.any(st & pg.PACK_MISSING_COARSE)
    p = O.make_params(gs=gs, sor_min_iter=40, sor_max_iter=40, sor_omega=OMEGA[gs - 1], poi_coeff=POI, dt=0.5)
    o_pot, o_flu, _, _ = O.solve(rho, cpot, flu, p)
    assert np.array_equal(fine_pot, o_pot[:, 1:9, 1:9, 1:9])
    assert np.array_equal(fine_flu, o_flu)
    assert not np.array_equal(fine_flu, d["fine_flu"])     # the gravity update did move the fluid


@pytest.mark.parametrize("gs", [1, 3, 5])
def test_pack_solve_scatter_converged_matches_closed_form(gs):
    fine, coarse, phi = _plummer_level()
    d = I.two_level_box(fine, coarse, (1, 1, 1), seed=10 + gs)
    _, _, fine_pot, _, it, pst = _cycle(gs, d, pg.TERM_RELRES, 10, 4000)
    assert np.all(pst & pg.STATUS_CONVERGED)
    err = np.abs(_assemble(d, fine_pot) - phi)
    print(f"[pack] gs={gs}: iterations mean {it.mean():.1f}; closed-form error max {err.max() / np.abs(phi).max():.2e} "
          f"of max|phi|, median relative {np.median(err / np.abs(phi)):.2e}", flush=True)
    assert err.max() < 0.02 * np.abs(phi).max() and np.median(err / np.abs(phi)) < 2e-3


def _to_dev(d):
    import torch

    dev = torch.device("cuda", 0)
    return {k: torch.as_tensor(np.ascontiguousarray(v)).to(dev) for k, v in d.items() if k not in ("fine_coords",
                                                                                               "coarse_coords")}


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
def test_device_prepare_and_scatter_equal_host(gs):
    """The device gather and scatter against the host ones, bitwise, on a level with missing neighbours
    (refined box on the domain edge) and a second launch on a smaller batch into the same buffers."""
    import torch

    rng = np.random.default_rng(gs)
    fine = rng.standard_normal((5, 32, 32, 32)).astype(np.float32)
    coarse = rng.standard_normal((24, 24, 24)).astype(np.float32)
    d = I.two_level_box(fine, coarse, (0, 1, 0), seed=gs)
    print(f"[case] device pack gs={gs} groups={d['group_first'].size}", flush=True)
    h = pg.prepare_patch_groups(gs, d["group_first"], d["group_father"], d["fine_flu"], d["fine_sib"],
                                d["coarse_pot"], d["coarse_sib"])
    dd = _to_dev(d)
    g = pg.geometry(gs)
    ng, nfine, ncoarse = d["group_first"].size, d["fine_flu"].shape[0], d["coarse_pot"].shape[0]
    dev = torch.device("cuda", 0)
    rho = torch.full((8 * ng, g["R"], g["R"], g["R"]), -7.0, device=dev)
    pin = torch.full((8 * ng, g["C"], g["C"], g["C"]), -7.0, device=dev)
    flu = torch.full((8 * ng, 5, 8, 8, 8), -7.0, device=dev)
    st = torch.full((8 * ng,), -1, dtype=torch.int32, device=dev)
    for groups in (ng, ng // 2):
        for a in (rho, pin, flu, st):
            a.fill_(-7)
        pg.prepare_patch_groups_device(gs, groups, dd["group_first"], dd["group_father"], nfine, dd["fine_flu"],
                                       dd["fine_sib"], ncoarse, dd["coarse_pot"], dd["coarse_sib"], rho, pin, flu, st)
        torch.cuda.synchronize()
        n = 8 * groups
        for a, b in zip((rho, pin, flu, st), h):
            assert np.array_equal(a[:n].cpu().numpy(), b[:n])
            assert np.all(a[n:].cpu().numpy() == -7)
    assert np.any(h[3] & pg.PACK_MISSING_COARSE)
    # scatter: potential windows and fluid back into device-resident level arrays
    G = pg.geometry(gs, 1)["G"]
    pot_out = torch.as_tensor(rng.standard_normal((8 * ng, G, G, G)).astype(np.float32)).to(dev)
    fine_pot = torch.zeros((nfine, 8, 8, 8), device=dev)
    fine_flu = dd["fine_flu"].clone()
    pg.scatter_patch_groups_device(1, ng, dd["group_first"], pot_out, flu, nfine, fine_pot, fine_flu)
    torch.cuda.synchronize()
    h_pot = np.zeros((nfine, 8, 8, 8), np.float32)
    h_flu = d["fine_flu"].copy()
    pg.scatter_patch_groups(1, d["group_first"], pot_out.cpu().numpy(), flu.cpu().numpy(), fine_pot=h_pot,
                            fine_flu=h_flu)
    assert np.array_equal(fine_pot.cpu().numpy(), h_pot) and np.array_equal(fine_flu.cpu().numpy(), h_flu)


def test_device_prepare_flags_bad_indices():
    """Index errors the device cannot return: a group outside the level is zero-filled, a neighbour index beyond
    the level zero-fills its cells; both flagged PGAMD_PACK_BAD_INDEX, the rest of the batch unaffected.  The
    scatter skips the bad group."""
    import torch

    gs = 3
    rng = np.random.default_rng(5)
    d = I.two_level_box(rng.standard_normal((5, 32, 32, 32)).astype(np.float32),
                        rng.standard_normal((16, 16, 16)).astype(np.float32), (0, 0, 0), seed=5)
    nfine, ncoarse = d["fine_flu"].shape[0], d["coarse_pot"].shape[0]
    d["group_first"][1] = nfine - 4                    # group 1 runs past the level
    d["fine_sib"][d["group_first"][2] + 3, 0] = nfine + 100   # one neighbour index of patch 8*2+3 beyond the level
    dd = _to_dev(d)
    g = pg.geometry(gs)
    ng = d["group_first"].size
    dev = torch.device("cuda", 0)
    rho = torch.full((8 * ng, g["R"], g["R"], g["R"]), -7.0, device=dev)
    pin = torch.full((8 * ng, g["C"], g["C"], g["C"]), -7.0, device=dev)
    flu = torch.full((8 * ng, 5, 8, 8, 8), -7.0, device=dev)
    st = torch.zeros(8 * ng, dtype=torch.int32, device=dev)
    print(f"[case] device pack bad indices gs={gs} groups={ng}", flush=True)
    pg.prepare_patch_groups_device(gs, ng, dd["group_first"], dd["group_father"], nfine, dd["fine_flu"],
                                   dd["fine_sib"], ncoarse, dd["coarse_pot"], dd["coarse_sib"], rho, pin, flu, st)
    torch.cuda.synchronize()
    s = st.cpu().numpy()
    assert np.all(s[8:16] & pg.PACK_BAD_INDEX) and np.all(rho[8:16].cpu().numpy() == 0)
    assert np.all(flu[8:16].cpu().numpy() == 0) and np.all(pin[8:16].cpu().numpy() == 0)
    assert s[19] & pg.PACK_BAD_INDEX
    ok = np.ones(8 * ng, bool)
    ok[8:16] = False
    ok[19] = False
    assert not np.any(s[ok] & pg.PACK_BAD_INDEX)
    # the host gather reports the neighbour error as an error and flags the same patch
    d2 = dict(d)
    d2["group_first"] = d["group_first"].copy()
    d2["group_first"][1] = d["group_first"][0]
    with pytest.raises(pg.PgError, match="neighbour table"):
        pg.prepare_patch_groups(gs, d2["group_first"], d2["group_father"], d2["fine_flu"], d2["fine_sib"],
                                d2["coarse_pot"], d2["coarse_sib"])
    fine_pot = torch.full((nfine, 8, 8, 8), 3.0, device=dev)
    pot_out = torch.ones((8 * ng, 10, 10, 10), device=dev)
    pg.scatter_patch_groups_device(1, ng, dd["group_first"], pot_out, None, nfine, fine_pot, None)
    torch.cuda.synchronize()
    fp = fine_pot.cpu().numpy()
    touched = np.zeros(nfine, bool)
    for gi, f in enumerate(d["group_first"]):
        if gi != 1:
            touched[f:f + 8] = True
    assert np.all(fp[touched] == 1.0) and np.all(fp[~touched] == 3.0)
