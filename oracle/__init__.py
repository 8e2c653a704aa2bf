"""ctypes front end of the CPU oracle (liboracle_pg.so) and of the reference builds under oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline
leg, always as the checker.  The product package never imports it.

The oracle restates the reference CPU solvers (CPU_Solver/*.cpp); see pg_oracle.cpp for the line-by-line
citations.  ``RefSolver`` drives a g++ build of the reference sources themselves (built by
``make -C oracle ref`` where the reference tree exists) and is used to pin the oracle.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_pg.so")
REF_DIR = os.path.join(HERE, "_ref")

SOR, MG = 1, 2
INT_CQUAD, INT_QUAD = 4, 5
INIT_INTERP, INIT_ZERO, INIT_GUESS = 0, 1, 2
TERM_REFERENCE, TERM_RELRES = 0, 1


class OrcParams(C.Structure):
    _fields_ = [
        ("ghost_size", C.c_int),
        ("gra_ghost_size", C.c_int),
        ("scheme", C.c_int),
        ("int_scheme", C.c_int),
        ("init_mode", C.c_int),
        ("term_mode", C.c_int),
        ("check_every", C.c_int),
        ("sor_min_iter", C.c_int),
        ("sor_max_iter", C.c_int),
        ("sor_omega", C.c_double),
        ("mg_max_iter", C.c_int),
        ("mg_npre_smooth", C.c_int),
        ("mg_npost_smooth", C.c_int),
        ("mg_tolerated_error", C.c_double),
        ("rel_tol", C.c_double),
        ("poi_coeff", C.c_double),
        ("dh", C.c_double),
        ("dt", C.c_double),
        ("p5_gradient", C.c_int),
        ("poisson", C.c_int),
        ("graacc", C.c_int),
    ]


def geometry(gs: int, gg: int = 1) -> dict:
    """Size algebra of Header/Macro.h:47-61 (PATCH_SIZE = 8)."""
    n = 8 + 2 * gs
    c = 4 + 2 * ((gs + 3) // 2)
    return {"N": n, "R": n - 2, "C": c, "G": 8 + 2 * gg, "U": gs % 2}


def build(quiet: bool = True) -> None:
    """Compile liboracle_pg.so (and oracle/_ref when the reference tree is present)."""
    out = subprocess.run(["make", "-C", HERE, "-j8"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
        l = C.CDLL(LIB_PATH)
        for name, ft in (("orc_poisson_gravity_f32", C.c_float), ("orc_poisson_gravity_f64", C.c_double)):
            fn = getattr(l, name)
            P = C.POINTER(ft)
            fn.argtypes = [P, P, P, P, P, C.c_int, C.POINTER(OrcParams), C.POINTER(C.c_int),
                           C.POINTER(C.c_double), C.c_int]
            fn.restype = C.c_int
        l.orc_interp_f64.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int]
        l.orc_interp_f32.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.c_int, C.c_int]
        l.orc_restrict_f64.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_int]
        l.orc_prolongate_add_f64.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_int]
        l.orc_mg_levels.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        l.orc_relres_f64.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_double, C.c_double]
        l.orc_relres_f64.restype = C.c_double
        l.orc_relres_f32.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.c_double, C.c_double]
        l.orc_relres_f32.restype = C.c_double
        _lib = l
    return _lib


def _ptr(a, ctype):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


def make_params(gs=1, gg=1, scheme=SOR, int_scheme=INT_CQUAD, init_mode=INIT_INTERP, term_mode=TERM_REFERENCE,
                check_every=1, sor_min_iter=10, sor_max_iter=60, sor_omega=1.69, mg_max_iter=10, mg_npre=3,
                mg_npost=3, mg_tol=1e-6, rel_tol=1e-5, poi_coeff=4.0 * np.pi * 1e-5, dh=1.0, dt=1e-2,
                p5=False, poisson=True, graacc=True) -> OrcParams:
    return OrcParams(gs, gg, scheme, int_scheme, init_mode, term_mode, check_every, sor_min_iter, sor_max_iter,
                     sor_omega, mg_max_iter, mg_npre, mg_npost, mg_tol, rel_tol, poi_coeff, dh, dt, int(p5),
                     int(poisson), int(graacc))


def solve(rho, cpot, flu, params: OrcParams, guess=None, pot_out=None, dtype=np.float32, nthreads=None):
    """Run the oracle on a batch.  Returns (pot_out, flu_out, iters, relres).  Inputs are not modified."""
    dtype = np.dtype(dtype)
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().orc_poisson_gravity_f32 if dtype == np.float32 else lib().orc_poisson_gravity_f64
    g = geometry(params.ghost_size, params.gra_ghost_size)
    npatch = (rho if rho is not None else pot_out).shape[0]
    rho_c = None if rho is None else np.ascontiguousarray(rho, dtype=dtype)
    cpot_c = None if cpot is None else np.ascontiguousarray(cpot, dtype=dtype)
    flu_c = None if flu is None else np.array(flu, dtype=dtype, copy=True, order="C")
    guess_c = None if guess is None else np.ascontiguousarray(guess, dtype=dtype)
    out = (np.zeros((npatch, g["G"], g["G"], g["G"]), dtype=dtype) if pot_out is None
           else np.array(pot_out, dtype=dtype, copy=True, order="C"))
    iters = np.zeros(npatch, dtype=np.int32)
    relres = np.zeros(npatch, dtype=np.float64)
    nt = nthreads if nthreads is not None else (os.cpu_count() or 1)
    rc = fn(_ptr(rho_c, ct), _ptr(cpot_c, ct), _ptr(out, ct), _ptr(flu_c, ct), _ptr(guess_c, ct), npatch,
            C.byref(params), _ptr(iters, C.c_int), _ptr(relres, C.c_double), nt)
    if rc != 0:
        raise ValueError(f"oracle returned error {rc}")
    return out, flu_c, iters, relres


def interp(cpot, gs, int_scheme, dtype=np.float64):
    g = geometry(gs)
    dtype = np.dtype(dtype)
    cp = np.ascontiguousarray(cpot, dtype=dtype)
    npatch = cp.shape[0]
    cube = np.zeros((npatch, g["N"], g["N"], g["N"]), dtype=dtype)
    if dtype == np.float64:
        rc = lib().orc_interp_f64(_ptr(cp, C.c_double), _ptr(cube, C.c_double), gs, int_scheme, npatch)
    else:
        rc = lib().orc_interp_f32(_ptr(cp, C.c_float), _ptr(cube, C.c_float), gs, int_scheme, npatch)
    if rc != 0:
        raise ValueError(f"oracle interp error {rc}")
    return cube


def restrict(fine):
    fine = np.ascontiguousarray(fine, dtype=np.float64)
    nf = fine.shape[0]
    ng = (C.c_int * 10)()
    nb = C.c_int()
    lib().orc_mg_levels(nf, ng, C.byref(nb))
    nc = ng[1]
    coarse = np.zeros((nc, nc, nc))
    lib().orc_restrict_f64(_ptr(fine, C.c_double), _ptr(coarse, C.c_double), nf, nc)
    return coarse


def restrict_to(fine, nc):
    fine = np.ascontiguousarray(fine, dtype=np.float64)
    coarse = np.zeros((nc, nc, nc))
    lib().orc_restrict_f64(_ptr(fine, C.c_double), _ptr(coarse, C.c_double), fine.shape[0], nc)
    return coarse


def prolongate_add(coarse, fine):
    coarse = np.ascontiguousarray(coarse, dtype=np.float64)
    f = np.array(fine, dtype=np.float64, copy=True, order="C")
    lib().orc_prolongate_add_f64(_ptr(coarse, C.c_double), _ptr(f, C.c_double), coarse.shape[0], f.shape[0])
    return f


def mg_levels(n0):
    ng = (C.c_int * 10)()
    nb = C.c_int()
    nlv = lib().orc_mg_levels(n0, ng, C.byref(nb))
    return [ng[i] for i in range(nlv)], nb.value


def relres(cube, rho, gs, poi_coeff, dh):
    """FP64 sum|R|/sum|phi| over the interior of one N^3 cube."""
    if cube.dtype == np.float32:
        return lib().orc_relres_f32(_ptr(np.ascontiguousarray(cube), C.c_float),
                                    _ptr(np.ascontiguousarray(rho, np.float32), C.c_float), gs, poi_coeff, dh)
    return lib().orc_relres_f64(_ptr(np.ascontiguousarray(cube, np.float64), C.c_double),
                                _ptr(np.ascontiguousarray(rho, np.float64), C.c_double), gs, poi_coeff, dh)


# ---- reference builds ---------------------------------------------------------------------------------

def ref_path(scheme: str, gs: int, gg: int, prec: str) -> str:
    return os.path.join(REF_DIR, f"libref_{scheme}_gs{gs}_gg{gg}_{prec}.so")


def ref_available(scheme="SOR", gs=1, gg=1, prec="f32") -> bool:
    return os.path.exists(ref_path(scheme, gs, gg, prec))


class RefSolver:
    """The reference's own CPU_PoissonGravitySolver, compiled for one configuration."""

    def __init__(self, scheme: str, gs: int, gg: int = 1, prec: str = "f32"):
        self.path = ref_path(scheme, gs, gg, prec)
        self.lib = C.CDLL(self.path)
        self.dtype = np.float32 if prec == "f32" else np.float64
        self.ct = C.c_float if prec == "f32" else C.c_double
        geo = (C.c_int * 7)()
        self.lib.ref_geometry(geo)
        self.geo = list(geo)
        P = C.POINTER(self.ct)
        self.lib.ref_poisson_gravity.argtypes = [P, P, P, P, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                                 C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                                 C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]

    def solve(self, rho, cpot, flu, npatch_group, dt, dh, sor_min, sor_max, omega, mg_max, npre, npost, mg_tol,
              poi_coeff, int_scheme, p5=False, poisson=True, graacc=True, pot_out=None, nthreads=1):
        npatch = 8 * npatch_group
        G = self.geo[2]
        rho_c = np.ascontiguousarray(rho, dtype=self.dtype)
        cpot_c = np.array(cpot, dtype=self.dtype, copy=True, order="C")
        flu_c = np.array(flu, dtype=self.dtype, copy=True, order="C")
        out = (np.zeros((npatch, G, G, G), dtype=self.dtype) if pot_out is None
               else np.array(pot_out, dtype=self.dtype, copy=True, order="C"))
        self.lib.ref_poisson_gravity(_ptr(rho_c, self.ct), _ptr(cpot_c, self.ct), _ptr(out, self.ct),
                                     _ptr(flu_c, self.ct), npatch_group, dt, dh, sor_min, sor_max, omega, mg_max,
                                     npre, npost, mg_tol, poi_coeff, int_scheme, int(p5), int(poisson),
                                     int(graacc), nthreads)
        return out, flu_c
