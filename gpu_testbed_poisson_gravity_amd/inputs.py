"""Seeded input generators for batched per-patch Poisson-gravity solves, with their analytic potentials.

Every generator returns host arrays in the solver's layouts (x fastest):
    rho  [npatch][R][R][R]      density on the solve interior, R = 6 + 2*gs
    cpot [npatch][C][C][C]      coarse potential (spacing 2*dh), C = 4 + 2*((gs+3)//2)
    flu  [npatch][5][8][8][8]   fluid (rho, px, py, pz, E) of the patch proper

Frame: the patch proper covers [0, 8*dh)^3.  Fine cell i of the density array is centred at
(i - gs + 1.5)*dh, solve-cube cell c at (c - gs + 0.5)*dh and coarse cell i at (i - (gs+3)//2 + 0.5)*2*dh,
which is the alignment the interpolation assumes (Header/Macro.h:53-61, CPU_PoissonSolver_SOR.cpp:11-12).

Units: dh = 1 and NEWTON_G = 1e-5 as in the reference harness (Makefile:41,47), so Poi_Coeff = 4*pi*1e-5.
"""
from __future__ import annotations

import numpy as np

PATCH = 8
NEWTON_G = 1.0e-5
POI_COEFF = 4.0 * np.pi * NEWTON_G


def geometry(gs: int, gg: int = 1) -> dict:
    """Size algebra of Header/Macro.h:47-61 with run-time ghost sizes."""
    if not (1 <= gs <= 5 and 1 <= gg <= 2 and gg <= gs):
        raise ValueError(f"unsupported ghost sizes gs={gs} gg={gg}")
    n = PATCH + 2 * gs
    pot_int = (gs + 3) // 2
    c = PATCH // 2 + 2 * pot_int
    return {"gs": gs, "gg": gg, "N": n, "R": n - 2, "C": c, "G": PATCH + 2 * gg, "U": gs % 2, "POT_INT": pot_int}


def fine_centres(gs: int, dh: float = 1.0) -> np.ndarray:
    """1-D centres of the density-array cells."""
    r = geometry(gs)["R"]
    return (np.arange(r) - gs + 1.5) * dh


def cube_centres(gs: int, dh: float = 1.0) -> np.ndarray:
    n = geometry(gs)["N"]
    return (np.arange(n) - gs + 0.5) * dh


def coarse_centres(gs: int, dh: float = 1.0) -> np.ndarray:
    g = geometry(gs)
    return (np.arange(g["C"]) - g["POT_INT"] + 0.5) * 2.0 * dh


def _r2(axes: np.ndarray, centres: np.ndarray) -> np.ndarray:
    """|x - c|^2 on the 3-D grid axes^3 for each centre: centres [..., 3] -> [..., n, n, n]."""
    cz = centres[..., 2, None, None, None]
    cy = centres[..., 1, None, None, None]
    cx = centres[..., 0, None, None, None]
    z = axes[:, None, None]
    y = axes[None, :, None]
    x = axes[None, None, :]
    return (z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2


def _fluid_from(rho_patch: np.ndarray, energy: float) -> np.ndarray:
    npatch = rho_patch.shape[0]
    flu = np.zeros((npatch, 5, PATCH, PATCH, PATCH), dtype=np.float64)
    flu[:, 0] = rho_patch
    flu[:, 4] = energy
    return flu


def _patch_proper(rho: np.ndarray, gs: int) -> np.ndarray:
    o = gs - 1
    return rho[:, o:o + PATCH, o:o + PATCH, o:o + PATCH]


# --- reference power-law IC (Main.cpp:524-580) ---------------------------------------------------------

def reference_ic(npatch: int, gs: int, width: float = 410.0, index: float = 2.1, offset: float = 0.09,
                 dh: float = 1.0, newton_g: float = NEWTON_G, dtype=np.float32):
    """rho = (r/W)^-n about the array centre, each patch shifted by offset*P*dh in z; coarse potential
    A r^(2-n) with A = 4 pi G W^n / ((n-3)(n-2)); fluid 2 + sin(2 pi/33.3 (x+y+z+v))."""
    g = geometry(gs)
    R, Cn = g["R"], g["C"]
    P = np.arange(npatch, dtype=np.float64)[:, None, None, None]
    ax = (np.arange(R) + 0.5) * dh
    cen = 0.5 * R * dh
    z = ax[None, :, None, None] + offset * P * dh
    y = ax[None, None, :, None]
    x = ax[None, None, None, :]
    r = np.sqrt((x - cen) ** 2 + (y - cen) ** 2 + (z - cen) ** 2)
    rho = (r / width) ** (-index)
    dh2 = 2.0 * dh
    A = 4.0 * np.pi * newton_g * width ** index / (index - 3.0) / (index - 2.0)
    axc = (np.arange(Cn) + 0.5) * dh2
    cenc = 0.5 * Cn * dh2
    zc = axc[None, :, None, None] + offset * P * dh
    rc = np.sqrt((axc[None, None, None, :] - cenc) ** 2 + (axc[None, None, :, None] - cenc) ** 2 + (zc - cenc) ** 2)
    cpot = A * rc ** (2.0 - index)
    axf = (np.arange(PATCH) + 0.5) * dh
    v = np.arange(5, dtype=np.float64)[None, :, None, None, None]
    zf = axf[None, None, :, None, None] + offset * P[:, :, :, :, None] * dh
    flu = 2.0 + np.sin(2.0 * np.pi / 33.3 * (axf[None, None, None, None, :] + axf[None, None, None, :, None] + zf + v))
    return rho.astype(dtype), cpot.astype(dtype), flu.astype(dtype)


def reference_ic_analytic(gs: int, gg: int, P: int, width=410.0, index=2.1, offset=0.09, dh=1.0, newton_g=NEWTON_G):
    """Analytic potential on the output window of patch P, as Output_Radial computes it (Main.cpp:679-708):
    returns (radius, phi_analytic) over the patch-proper cells of the G^3 window."""
    G = PATCH + 2 * gg
    A = 4.0 * np.pi * newton_g * width ** index / (index - 3.0) / (index - 2.0)
    cen = 0.5 * G * dh
    ax = (np.arange(gg, gg + PATCH) + 0.5) * dh
    z = ax[:, None, None] + offset * P * dh
    r = np.sqrt((ax[None, None, :] - cen) ** 2 + (ax[None, :, None] - cen) ** 2 + (z - cen) ** 2)
    return r, A * r ** (2.0 - index)


# --- config generators ----------------------------------------------------------------------------------

def softened_powerlaw(npatch: int, gs: int, seed: int, eps: float = 0.5, rho0_range=(0.1, 10.0), dh: float = 1.0,
                      newton_g: float = NEWTON_G, energy: float = 1.0, dtype=np.float32):
    """rho = rho0/(r^2+eps^2) about a seeded centre in the patch, rho0 log-uniform.  The exact solution of
    lap(phi) = 4 pi G rho is phi = 4 pi G rho0 [eps atan(r/eps)/r + ln(r^2+eps^2)/2] (+ const), which gives
    the coarse potential.  Fluid: rho from the same field, zero momentum, uniform E."""
    rng = np.random.default_rng(seed)
    rho0 = np.exp(rng.uniform(np.log(rho0_range[0]), np.log(rho0_range[1]), npatch))
    centre = rng.uniform(0.0, PATCH * dh, (npatch, 3))
    rho = rho0[:, None, None, None] / (_r2(fine_centres(gs, dh), centre) + eps ** 2)
    cpot = softened_powerlaw_phi(np.sqrt(_r2(coarse_centres(gs, dh), centre)), rho0[:, None, None, None], eps, newton_g)
    flu = _fluid_from(_patch_proper(rho, gs), energy)
    meta = {"rho0": rho0, "centre": centre, "eps": eps}
    return rho.astype(dtype), cpot.astype(dtype), flu.astype(dtype), meta


def softened_powerlaw_phi(r, rho0, eps, newton_g=NEWTON_G):
    r = np.asarray(r, dtype=np.float64)
    safe = np.where(r > 0, r, 1.0)
    core = np.where(r > 0, eps * np.arctan(safe / eps) / safe, 1.0)
    return 4.0 * np.pi * newton_g * rho0 * (core + 0.5 * np.log(r * r + eps * eps))


def plummer(npatch: int, gs: int, seed: int, nmax: int = 8, mass_range=(0.01, 1.0), a_range=(0.5, 4.0),
            dh: float = 1.0, newton_g: float = NEWTON_G, energy: float = 1.0, dtype=np.float32):
    """Sum of 1..nmax Plummer spheres per patch: rho = 3M/(4 pi a^3) (1 + r^2/a^2)^(-5/2), with
    phi = -G M / sqrt(r^2 + a^2) in closed form for the coarse potential.  Masses log-uniform, scale radii
    uniform (in cells), centres uniform over the solve cube."""
    rng = np.random.default_rng(seed)
    g = geometry(gs)
    nsph = rng.integers(1, nmax + 1, npatch)
    mass = np.exp(rng.uniform(np.log(mass_range[0]), np.log(mass_range[1]), (npatch, nmax)))
    mass *= (np.arange(nmax)[None, :] < nsph[:, None])
    a = rng.uniform(a_range[0], a_range[1], (npatch, nmax)) * dh
    lo, hi = -gs * dh, (PATCH + gs) * dh
    centre = rng.uniform(lo, hi, (npatch, nmax, 3))
    R, Cn = g["R"], g["C"]
    rho = np.zeros((npatch, R, R, R))
    cpot = np.zeros((npatch, Cn, Cn, Cn))
    fa, ca = fine_centres(gs, dh), coarse_centres(gs, dh)
    for s in range(nmax):
        m = mass[:, s, None, None, None]
        aa = a[:, s, None, None, None]
        rho += 3.0 * m / (4.0 * np.pi * aa ** 3) * (1.0 + _r2(fa, centre[:, s]) / aa ** 2) ** -2.5
        cpot += -newton_g * m / np.sqrt(_r2(ca, centre[:, s]) + aa ** 2)
    flu = _fluid_from(_patch_proper(rho, gs), energy)
    meta = {"nsph": nsph, "mass": mass, "a": a, "centre": centre}
    return rho.astype(dtype), cpot.astype(dtype), flu.astype(dtype), meta


def plummer_phi(points: np.ndarray, meta: dict, p: int, newton_g: float = NEWTON_G) -> np.ndarray:
    """Analytic potential of patch p's spheres at points [..., 3] (x, y, z)."""
    phi = np.zeros(points.shape[:-1])
    for s in range(meta["mass"].shape[1]):
        m = meta["mass"][p, s]
        if m == 0:
            continue
        d2 = np.sum((points - meta["centre"][p, s]) ** 2, axis=-1)
        phi += -newton_g * m / np.sqrt(d2 + meta["a"][p, s] ** 2)
    return phi


_CUBE_SELF = 2.3800772543    # integral of 1/|r| over the unit cube centred on the origin


def _direct_sum_kernel(gs: int, dh: float) -> np.ndarray:
    """K[i, j] = 1/|x_i - x_j| between coarse cell centres, with the uniform-cube self term on the diagonal."""
    ca = coarse_centres(gs, dh)
    zz, yy, xx = np.meshgrid(ca, ca, ca, indexing="ij")
    pts = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
    d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    L = 2.0 * dh
    with np.errstate(divide="ignore"):
        k = 1.0 / d
    np.fill_diagonal(k, _CUBE_SELF / L)
    return k


def _smooth_axis(f: np.ndarray, w: np.ndarray, axis: int) -> np.ndarray:
    """'valid' correlation of f with w along one axis, as a sum of shifted slices."""
    m = f.shape[axis] - len(w) + 1
    out = np.zeros(f.shape[:axis] + (m,) + f.shape[axis + 1:])
    for t, wt in enumerate(w):
        out += wt * np.take(f, np.arange(t, t + m), axis=axis)
    return out


def lognormal(npatch: int, gs: int, seed: int, sigma: float = 1.5, corr: float = 3.0, mean_rho: float = 1.0,
              dh: float = 1.0, newton_g: float = NEWTON_G, energy: float = 1.0, dtype=np.float32,
              chunk: int = 512):
    """Lognormal density: a unit-variance Gaussian random field (white noise smoothed by a Gaussian of
    width corr cells) exponentiated as mean_rho*exp(sigma g - sigma^2/2), generated over the whole coarse
    region.  The coarse potential is the free-space FP64 direct sum over the coarse-averaged density
    (uniform-cube self term)."""
    rng = np.random.default_rng(seed)
    g = geometry(gs)
    Cn, R = g["C"], g["R"]
    nf = 2 * Cn                 # fine cells spanned by the coarse cube
    pad = int(np.ceil(3 * corr))
    n = nf + 2 * pad
    t = np.arange(-pad, pad + 1)
    w = np.exp(-0.5 * (t / corr) ** 2)
    w /= np.sqrt((w ** 2).sum())   # unit variance after three separable passes
    o = 2 * g["POT_INT"] - gs + 1  # field index of the first density cell
    kern = _direct_sum_kernel(gs, dh)
    rho = np.empty((npatch, R, R, R), dtype=dtype)
    cpot = np.empty((npatch, Cn, Cn, Cn), dtype=dtype)
    for b in range(0, npatch, chunk):
        m = min(chunk, npatch - b)
        f = rng.standard_normal((m, n, n, n))
        for axis in (1, 2, 3):
            f = _smooth_axis(f, w, axis)
        field = mean_rho * np.exp(sigma * f - 0.5 * sigma ** 2)      # [m, nf, nf, nf]
        rho[b:b + m] = field[:, o:o + R, o:o + R, o:o + R]
        coarse_mean = field.reshape(m, Cn, 2, Cn, 2, Cn, 2).mean(axis=(2, 4, 6))
        mass = coarse_mean.reshape(m, -1) * (2.0 * dh) ** 3
        cpot[b:b + m] = (-newton_g * (mass @ kern)).reshape(m, Cn, Cn, Cn)
    flu = _fluid_from(_patch_proper(rho.astype(np.float64), gs), energy)
    return rho, cpot, flu.astype(dtype), {}


def uniform_quadratic(npatch: int, gs: int, rho_c: float = 1.0, dh: float = 1.0, newton_g: float = NEWTON_G,
                      dtype=np.float64):
    """Constant density with the quadratic potential phi = (2/3) pi G rho r^2 about the patch centre.
    lap_h(r^2) = 6 exactly on any grid, so this phi is the exact discrete solution."""
    g = geometry(gs)
    c = 0.5 * PATCH * dh
    centre = np.full((npatch, 3), c)
    rho = np.full((npatch, g["R"], g["R"], g["R"]), rho_c)
    cpot = (2.0 / 3.0) * np.pi * newton_g * rho_c * _r2(coarse_centres(gs, dh), centre)
    flu = _fluid_from(_patch_proper(rho, gs), 1.0)
    return rho.astype(dtype), cpot.astype(dtype), flu.astype(dtype)


_KINDS = {"powerlaw": lambda n, gs, seed: softened_powerlaw(n, gs, seed),
          "plummer": lambda n, gs, seed: plummer(n, gs, seed),
          "lognormal": lambda n, gs, seed: lognormal(n, gs, seed, chunk=128)}


def _mixed_job(job):
    kind, gs, n, seed = job
    r, c, f, _ = _KINDS[kind](n, gs, seed)
    return r, c, f


def mixed_batch(npatch: int, seed: int, kinds=("lognormal",), gs_choices=(1, 2, 3, 4, 5), workers: int = 1,
                chunk: int = 256):
    """Seeded batch of npatch patches with ghost sizes drawn uniformly from gs_choices and density types
    assigned in equal shares (kinds cycled over a seeded permutation).  Returns (gs [n], kind [n], rho list,
    cpot list, flu [n,5,8,8,8]); generation of the (kind, gs) groups runs on `workers` host processes."""
    rng = np.random.default_rng(seed)
    gs = rng.choice(np.asarray(gs_choices, np.int32), npatch)
    kind = np.asarray(kinds)[rng.permutation(np.arange(npatch) % len(kinds))]
    jobs, where = [], []
    for ki, k in enumerate(kinds):
        for g in gs_choices:
            idx = np.nonzero((gs == g) & (kind == k))[0]
            for j, b in enumerate(range(0, idx.size, chunk)):
                sel = idx[b:b + chunk]
                jobs.append((k, int(g), int(sel.size), seed * 7919 + ki * 1000 + g * 100 + j))
                where.append(sel)
    if workers > 1 and len(jobs) > 1:
        import multiprocessing as mp

        with mp.get_context("fork").Pool(min(workers, len(jobs))) as pool:
            parts = pool.map(_mixed_job, jobs, chunksize=1)
    else:
        parts = [_mixed_job(j) for j in jobs]
    rho, cpot = [None] * npatch, [None] * npatch
    flu = np.zeros((npatch, 5, PATCH, PATCH, PATCH), np.float32)
    for sel, (r, c, f) in zip(where, parts):
        for k, i in enumerate(sel):
            rho[i], cpot[i] = r[k], c[k]
        flu[sel] = f
    return gs, kind, rho, cpot, flu


def _sib_table(coords: np.ndarray, extent: int) -> np.ndarray:
    """Neighbour table [n, 27] of patches at integer coords [n, 3] (z, y, x) on a box of extent^3 patch slots:
    entry (dz+1)*9 + (dy+1)*3 + (dx+1) is the index of the patch at coords + (dz, dy, dx), -1 where there is
    none (outside the box; no periodicity)."""
    slot = -np.ones((extent,) * 3, np.int64)
    slot[coords[:, 0], coords[:, 1], coords[:, 2]] = np.arange(coords.shape[0])
    padded = -np.ones((extent + 2,) * 3, np.int64)
    padded[1:-1, 1:-1, 1:-1] = slot
    sib = np.empty((coords.shape[0], 27), np.int32)
    k = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                sib[:, k] = padded[coords[:, 0] + 1 + dz, coords[:, 1] + 1 + dy, coords[:, 2] + 1 + dx]
                k += 1
    return sib


def two_level_box(fine_flu_grid: np.ndarray, coarse_pot_grid: np.ndarray, refined_origin, seed: int = 0) -> dict:
    """Two AMR levels in the patch-group format of pgamd_prepare_patch_groups, cut from global grids.

    coarse_pot_grid [L, L, L] (L = 8 * coarse patches per side, spacing 2h) is the coarse level, covering the
    whole domain.  The refined region is the box of m^3 coarse patches starting at coarse patch
    refined_origin = (z, y, x); fine_flu_grid [5, 16m, 16m, 16m] (spacing h) is its fluid.  Each coarse patch
    of the box has one patch group (8 fine patches).  Patch groups and coarse patches are stored in a seeded
    random order, so nothing may rely on storage order.  Returns group_first, group_father, fine_flu
    [8G, 5, 8, 8, 8], fine_sib, fine_coords [8G, 3] (fine patch coords inside the box, z y x), coarse_pot
    [Nc, 8, 8, 8], coarse_sib, coarse_coords [Nc, 3]."""
    rng = np.random.default_rng(seed)
    L = coarse_pot_grid.shape[0]
    if coarse_pot_grid.shape != (L, L, L) or L % PATCH:
        raise ValueError("coarse_pot_grid must be [8k, 8k, 8k]")
    ncp = L // PATCH
    F = fine_flu_grid.shape[1]
    if fine_flu_grid.shape != (5, F, F, F) or F % (2 * PATCH):
        raise ValueError("fine_flu_grid must be [5, 16m, 16m, 16m]")
    m = F // (2 * PATCH)
    o = np.asarray(refined_origin, np.int64)
    if np.any(o < 0) or np.any(o + m > ncp):
        raise ValueError("refined box outside the coarse level")
    # coarse patches in random order
    cc = np.stack(np.meshgrid(*(np.arange(ncp),) * 3, indexing="ij"), -1).reshape(-1, 3)[rng.permutation(ncp ** 3)]
    cpot = np.stack([coarse_pot_grid[z * 8:z * 8 + 8, y * 8:y * 8 + 8, x * 8:x * 8 + 8] for z, y, x in cc])
    cindex = {tuple(c): i for i, c in enumerate(cc)}
    # one patch group per coarse patch of the box, groups in random order, children in l = (lz*2+ly)*2+lx order
    gb = np.stack(np.meshgrid(*(np.arange(m),) * 3, indexing="ij"), -1).reshape(-1, 3)[rng.permutation(m ** 3)]
    fc = np.empty((8 * gb.shape[0], 3), np.int64)
    for g, (z, y, x) in enumerate(gb):
        for l in range(8):
            fc[8 * g + l] = (2 * z + (l >> 2), 2 * y + ((l >> 1) & 1), 2 * x + (l & 1))
    flu = np.stack([fine_flu_grid[:, z * 8:z * 8 + 8, y * 8:y * 8 + 8, x * 8:x * 8 + 8] for z, y, x in fc])
    return {
        "group_first": np.arange(0, 8 * gb.shape[0], 8, dtype=np.int32),
        "group_father": np.asarray([cindex[tuple(b + o)] for b in gb], np.int32),
        "fine_flu": np.ascontiguousarray(flu, np.float32),
        "fine_sib": _sib_table(fc, 2 * m),
        "fine_coords": fc,
        "coarse_pot": np.ascontiguousarray(cpot, np.float32),
        "coarse_sib": _sib_table(cc, ncp),
        "coarse_coords": cc,
    }
