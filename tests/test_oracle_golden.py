"""The CPU oracle against the reference's own outputs.

* Committed fixtures (tests/golden/ref_gs*.npz, written by make_golden.py from a g++ build of the reference
  CPU solver) must be reproduced bit for bit, for SOR and MG, CQUAD and QUAD, FP32 and FP64, the
  reference's stall rule, fixed iteration counts and the 5-point gradient.
* Where oracle/_ref is built (this container), fresh seeded inputs are compared with the reference build
  directly.
"""
import os

import numpy as np
import pytest

import oracle as O
from gpu_testbed_poisson_gravity_amd import inputs as I

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
POI = 4.0 * np.pi * 1e-5
OMEGA = [1.49, 1.57, 1.62, 1.65, 1.69]


def _params(gs, name, prec):
    scheme = O.SOR if name.startswith("SOR") else O.MG
    ints = O.INT_QUAD if "_quad_" in name else O.INT_CQUAD
    f32 = prec == "f32"
    sor_min, sor_max = (50, 50) if name.endswith("fixed50") else (10, 60 if f32 else 100)
    p5 = name.endswith("p5")
    return O.make_params(gs=gs, gg=2 if p5 else 1, scheme=scheme, int_scheme=ints, sor_min_iter=sor_min,
                         sor_max_iter=sor_max, sor_omega=OMEGA[gs - 1], mg_max_iter=10 if f32 else 20,
                         mg_npre=3, mg_npost=3, mg_tol=1e-6 if f32 else 1e-15, poi_coeff=POI, dh=1.0, dt=1e-2,
                         p5=p5)


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
def test_oracle_reproduces_reference_fixtures(gs):
    d = np.load(os.path.join(GOLDEN, f"ref_gs{gs}.npz"))
    names = sorted({k.split("__")[0] for k in d.files if "__" in k})
    assert len(names) >= 9
    for name in names:
        prec = "f64" if "_f64" in name else "f32"
        dt = np.float64 if prec == "f64" else np.float32
        p = _params(gs, name, prec)
        pot, flu, iters, _ = O.solve(d["rho"].astype(dt), d["cpot"].astype(dt), d["flu"].astype(dt), p, dtype=dt)
        ref_pot = d[f"{name}__pot"]
        assert pot.dtype == ref_pot.dtype
        assert np.array_equal(pot, ref_pot), f"{name}: potential differs from the reference"
        if f"{name}__flu" in d.files:
            assert np.array_equal(flu, d[f"{name}__flu"]), f"{name}: fluid differs from the reference"
        if name.endswith("fixed50"):
            assert np.all(iters == 50)


ref_built = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference tree)")


@ref_built
@pytest.mark.parametrize("scheme", ["SOR", "MG"])
@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_oracle_matches_reference_build(scheme, gs, prec):
    dt = np.float32 if prec == "f32" else np.float64
    rho, cpot, flu, _ = I.lognormal(8, gs, seed=11 * gs)
    rho2, cpot2, flu2, _ = I.softened_powerlaw(8, gs, seed=13 * gs)
    rho, cpot, flu = (np.concatenate(a).astype(dt) for a in ((rho, rho2), (cpot, cpot2), (flu, flu2)))
    flu[:, 1:4] = 0.1 * np.sin(np.arange(flu[:, 1:4].size)).reshape(flu[:, 1:4].shape)  # non-zero momentum
    for ints in (O.INT_CQUAD, O.INT_QUAD):
        ref = O.RefSolver(scheme, gs, 1, prec)
        pot_r, flu_r = ref.solve(rho, cpot, flu, 2, 0.5, 1.0, 7, 40, 1.7, 6, 2, 2, 1e-5, POI, ints)
        p = O.make_params(gs=gs, scheme=O.SOR if scheme == "SOR" else O.MG, int_scheme=ints, sor_min_iter=7,
                          sor_max_iter=40, sor_omega=1.7, mg_max_iter=6, mg_npre=2, mg_npost=2, mg_tol=1e-5,
                          poi_coeff=POI, dt=0.5)
        pot_o, flu_o, _, _ = O.solve(rho, cpot, flu, p, dtype=dt)
        assert np.array_equal(pot_o, pot_r)
        assert np.array_equal(flu_o, flu_r)


@ref_built
def test_reference_build_only_gravity_path():
    """Poisson=false, GraAcc=true re-uses the caller's potential (CUAPI_Asyn_PoissonGravitySolver.cu:192-200)."""
    rho, cpot, flu, _ = I.plummer(8, 3, seed=5)
    pot_in = np.random.default_rng(1).standard_normal((8, 10, 10, 10)).astype(np.float32)
    ref = O.RefSolver("SOR", 3, 1, "f32")
    pot_r, flu_r = ref.solve(rho, cpot, flu, 1, 0.3, 1.0, 10, 60, 1.6, 10, 3, 3, 1e-6, POI, 4, poisson=False,
                             pot_out=pot_in)
    p = O.make_params(gs=3, dt=0.3, poisson=False)
    pot_o, flu_o, _, _ = O.solve(rho, cpot, flu, p, pot_out=pot_in)
    assert np.array_equal(pot_r, pot_in) and np.array_equal(pot_o, pot_in)
    assert np.array_equal(flu_o, flu_r)
