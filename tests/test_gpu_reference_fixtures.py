"""The HIP path against the reference's own outputs (needs an MI355X).

tests/golden/ref_gs{1..5}.npz hold inputs and outputs of the reference CPU solver itself (a g++ build of
CPU_Solver/*.cpp, written by tests/golden/make_golden.py).  The oracle reproduces every entry bit for bit
(tests/test_oracle_golden.py); here the device kernels are held to the same entries directly:

* Equal iteration counts: each FP32 entry is re-run on the GPU with every patch's reference iteration count
  (the stall rule's count for SOR, the V-cycle count for MG, read from the oracle, which matches the
  reference bitwise) as a fixed count.  Potential and fluid must equal the reference's bit for bit.
* The stall rule itself (CPU_PoissonSolver_SOR.cpp:256, 275), as the drop-in entry runs it: the device sums
  |R| as a wavefront tree and the reference sequentially in FP32, so a stall that sits on a rounding tie can
  be detected at another iteration.  The per-patch agreement of the iteration counts is printed and bounded.
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
F32_ENTRIES = ["SOR_cquad_f32", "SOR_quad_f32", "SOR_cquad_f32_fixed50", "SOR_cquad_f32_p5", "MG_cquad_f32",
               "MG_quad_f32"]


def _entry(gs, name):
    """Settings of a fixture entry (make_golden.py): scheme, GRA_GHOST_SIZE, interpolation, iteration limits."""
    scheme = pg.SOR if name.startswith("SOR") else pg.MG
    ints = pg.INT_QUAD if "_quad_" in name else pg.INT_CQUAD
    p5 = name.endswith("p5")
    sor_min, sor_max = (50, 50) if name.endswith("fixed50") else (10, 60)
    return dict(scheme=scheme, ints=ints, p5=p5, gg=2 if p5 else 1, sor_min=sor_min, sor_max=sor_max)


def _oracle_iters(d, gs, e):
    p = O.make_params(gs=gs, gg=e["gg"], scheme=e["scheme"], int_scheme=e["ints"], sor_min_iter=e["sor_min"],
                      sor_max_iter=e["sor_max"], sor_omega=OMEGA[gs - 1], mg_max_iter=10, mg_npre=3, mg_npost=3,
                      mg_tol=1e-6, poi_coeff=POI, dh=1.0, dt=1e-2, p5=e["p5"])
    pot, flu, it, _ = O.solve(d["rho"], d["cpot"], d["flu"], p)
    return pot, flu, it


def _gpu(rho, cpot, flu, gs, e, sor_min, sor_max, mg_iters=10, mg_tol=1e-6, term=pg.TERM_REFERENCE):
    return pg.solve_host(rho, cpot, flu, gs=gs, gg=e["gg"], scheme=e["scheme"], init_mode=pg.INIT_INTERP,
                         term_mode=term, dt=1e-2, dh=1.0, sor_min_iter=sor_min, sor_max_iter=sor_max,
                         sor_omega=OMEGA[gs - 1], mg_max_iter=mg_iters, mg_npre=3, mg_npost=3, mg_tol=mg_tol,
                         poi_coeff=POI, int_scheme=e["ints"], p5=e["p5"])


@pytest.mark.parametrize("gs", [1, 2, 3, 4, 5])
def test_fixtures_bitwise_at_reference_iteration_counts(gs):
    d = np.load(os.path.join(GOLDEN, f"ref_gs{gs}.npz"))
    for name in F32_ENTRIES:
        if f"{name}__pot" not in d.files:
            continue                                          # the 5-point entry exists for gs >= 2 only
        e = _entry(gs, name)
        o_pot, _, it = _oracle_iters(d, gs, e)
        assert np.array_equal(o_pot, d[f"{name}__pot"])      # the count source reproduces the reference
        pot = np.empty_like(d[f"{name}__pot"])
        flu = np.empty_like(d["flu"])
        for k in np.unique(it):
            sel = np.nonzero(it == k)[0]
            print(f"[case] fixture gs={gs} {name} patches {sel.tolist()} count {int(k)}", flush=True)
            if e["scheme"] == pg.SOR:
                p_g, f_g, it_g, _ = _gpu(d["rho"][sel], d["cpot"][sel], d["flu"][sel], gs, e, int(k), int(k))
            else:
                p_g, f_g, it_g, _ = _gpu(d["rho"][sel], d["cpot"][sel], d["flu"][sel], gs, e, 10, 60,
                                         mg_iters=int(k), mg_tol=0.0)
            assert np.all(it_g == k)
            pot[sel], flu[sel] = p_g, f_g
        assert np.array_equal(pot, d[f"{name}__pot"]), f"gs={gs} {name}: potential differs from the reference"
        if f"{name}__flu" in d.files:
            assert np.array_equal(flu, d[f"{name}__flu"]), f"gs={gs} {name}: fluid differs from the reference"


def test_stall_rule_iteration_agreement():
    """Stall rule (Min_Iter 10, Max_Iter 60) on the fixtures and on 256 seeded patches per ghost size: the
    iteration counts of the device and of the reference differ by at most 4, and where they agree the
    potentials are bitwise equal; the agreement rate is printed."""
    agree = total = bitwise = 0
    per_gs = {}
    for gs in range(1, 6):
        d = np.load(os.path.join(GOLDEN, f"ref_gs{gs}.npz"))
        rl, cl, fl, _ = I.lognormal(128, gs, seed=500 + gs)
        rp, cp, fp, _ = I.plummer(128, gs, seed=600 + gs)
        sets = [("fixture", d["rho"], d["cpot"], d["flu"], d["SOR_cquad_f32__pot"]),
                ("seeded", np.concatenate([rl, rp]), np.concatenate([cl, cp]), np.concatenate([fl, fp]), None)]
        e = _entry(gs, "SOR_cquad_f32")
        for label, rho, cpot, flu, ref_pot in sets:
            p = O.make_params(gs=gs, sor_min_iter=10, sor_max_iter=60, sor_omega=OMEGA[gs - 1], poi_coeff=POI,
                              dh=1.0, dt=1e-2)
            o_pot, _, o_it, _ = O.solve(rho, cpot, flu, p)
            if ref_pot is not None:
                assert np.array_equal(o_pot, ref_pot)
            print(f"[case] stall rule gs={gs} {label} n={rho.shape[0]}", flush=True)
            pot, _, it, _ = _gpu(rho, cpot, flu, gs, e, 10, 60)
            assert np.abs(it - o_it).max() <= 4
            same = it == o_it
            assert np.array_equal(pot[same], o_pot[same])
            err = np.abs(pot.astype(np.float64) - o_pot).max() / np.abs(o_pot).max()
            assert err <= 1e-5
            agree += int(same.sum())
            total += same.size
            bitwise += int(sum(np.array_equal(pot[k], o_pot[k]) for k in range(same.size)))
            a, t = per_gs.get(gs, (0, 0))
            per_gs[gs] = (a + int(same.sum()), t + same.size)
    rates = {g: round(a / t, 3) for g, (a, t) in per_gs.items()}
    print(f"[stall-rule agreement] equal iteration counts {agree}/{total} = {agree / total:.3f}; bitwise equal "
          f"potentials {bitwise}/{total}; by ghost size {rates}", flush=True)
    assert agree / total >= 0.5
