"""Write the golden fixtures under tests/golden/ from a g++ build of the reference CPU solver.

Run in a container that has the reference tree (make -C oracle builds oracle/_ref from it):
    python tests/golden/make_golden.py

Each fixture holds inputs and the reference's outputs only (data, no source).  Inputs are one FP32 batch of
8 patches (one patch group, as the reference requires) per ghost size: patches 0-3 follow the reference
power-law initial condition (Main.cpp:524-580), patches 4-7 are seeded Plummer spheres.  FP64 runs use the
same FP32 values widened to FP64.  Outputs are recorded for every (scheme, interpolation, precision)
combination with the reference defaults (Main.cpp:1142-1223), plus a fixed-50-iteration SOR run and a
5-point-gradient run.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from gpu_testbed_poisson_gravity_amd import inputs as I  # noqa: E402

POI = 4.0 * np.pi * 1e-5
OMEGA = [1.49, 1.57, 1.62, 1.65, 1.69]


def batch(gs, gg=1):
    r1, c1, f1 = I.reference_ic(4, gs)
    r2, c2, f2, _ = I.plummer(4, gs, seed=100 + gs)
    return (np.concatenate([r1, r2]).astype(np.float32), np.concatenate([c1, c2]).astype(np.float32),
            np.concatenate([f1, f2]).astype(np.float32))


def runs(gs):
    """(name, scheme, gg, int_scheme, prec, kwargs) of the reference runs recorded for one ghost size."""
    out = []
    for scheme in ("SOR", "MG"):
        for ints, iname in ((4, "cquad"), (5, "quad")):
            for prec in ("f32", "f64"):
                out.append((f"{scheme}_{iname}_{prec}", scheme, 1, ints, prec, {}))
    out.append(("SOR_cquad_f32_fixed50", "SOR", 1, 4, "f32", {"sor_min": 50, "sor_max": 50}))
    if gs >= 2:
        out.append(("SOR_cquad_f32_p5", "SOR", 2, 4, "f32", {"p5": True}))
    return out


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref is not built: run `make -C oracle` where the reference tree exists")
    for gs in range(1, 6):
        rho, cpot, flu = batch(gs)
        rec = {"rho": rho, "cpot": cpot, "flu": flu}
        for name, scheme, gg, ints, prec, kw in runs(gs):
            ref = O.RefSolver(scheme, gs, gg, prec)
            dt = np.float32 if prec == "f32" else np.float64
            sor_max = kw.get("sor_max", 60 if prec == "f32" else 100)
            sor_min = kw.get("sor_min", 10)
            mg_max, mg_tol = (10, 1e-6) if prec == "f32" else (20, 1e-15)
            pot, fl = ref.solve(rho.astype(dt), cpot.astype(dt), flu.astype(dt), 1, 1e-2, 1.0, sor_min, sor_max,
                                OMEGA[gs - 1], mg_max, 3, 3, mg_tol, POI, ints, p5=kw.get("p5", False))
            rec[f"{name}__pot"] = pot
            if name.endswith(("cquad_f32", "p5")):
                rec[f"{name}__flu"] = fl
        path = os.path.join(HERE, f"ref_gs{gs}.npz")
        np.savez_compressed(path, **rec)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
