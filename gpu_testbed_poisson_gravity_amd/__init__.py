"""MI355X-native batched per-patch Poisson-gravity solver.

The product is the C ABI in include/pg_amd.h, implemented by libpgamd.so (hand-written HIP kernels for
gfx950, built by ``make -C gpu_testbed_poisson_gravity_amd``).  This module is a thin ctypes binding of that
ABI for Python callers (tests, bench.py); it has no compute path of its own and raises if the library is
missing.  Device arrays are passed as raw pointers, e.g. ``tensor.data_ptr()`` of a torch tensor.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpgamd.so")

SOR, MG = 1, 2
INT_CQUAD, INT_QUAD = 4, 5
INIT_INTERP, INIT_ZERO, INIT_GUESS, INIT_GUESS_CUBE = 0, 1, 2, 3
TERM_REFERENCE, TERM_RELRES = 0, 1
STATUS_CONVERGED, STATUS_MAX_ITER, STATUS_NONFINITE, STATUS_FALLBACK = 1, 2, 4, 8
PACK_MISSING_FINE, PACK_MISSING_COARSE, PACK_BAD_INDEX = 1, 2, 4
LAYOUT_COMPACT = 4   # SOR layout code of the compact single-buffer cube (ghost size 5, one patch per wavefront)
ERR_ARG, ERR_DEVICE, ERR_STATE, ERR_NOMEM = -1, -2, -3, -4


class PgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pgamd error {code}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [
        ("pot_ghost_size", C.c_int),
        ("gra_ghost_size", C.c_int),
        ("pot_scheme", C.c_int),
        ("init_mode", C.c_int),
        ("term_mode", C.c_int),
        ("check_every", C.c_int),
        ("rel_tol", C.c_float),
        ("fuse_gravity", C.c_int),
        ("guess_fallback_tol", C.c_float),
        ("mixed_persistent", C.c_int),
    ]


class SolverArgs(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("dh", C.c_float),
        ("sor_min_iter", C.c_int),
        ("sor_max_iter", C.c_int),
        ("sor_omega", C.c_float),
        ("mg_max_iter", C.c_int),
        ("mg_npre_smooth", C.c_int),
        ("mg_npost_smooth", C.c_int),
        ("mg_tolerated_error", C.c_float),
        ("poi_coeff", C.c_float),
        ("int_scheme", C.c_int),
        ("p5_gradient", C.c_int),
        ("poisson", C.c_int),
        ("graacc", C.c_int),
    ]


class SolverArgsF64(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("dh", C.c_double),
        ("sor_min_iter", C.c_int),
        ("sor_max_iter", C.c_int),
        ("sor_omega", C.c_double),
        ("mg_max_iter", C.c_int),
        ("mg_npre_smooth", C.c_int),
        ("mg_npost_smooth", C.c_int),
        ("mg_tolerated_error", C.c_double),
        ("poi_coeff", C.c_double),
        ("int_scheme", C.c_int),
        ("p5_gradient", C.c_int),
        ("poisson", C.c_int),
        ("graacc", C.c_int),
    ]


# Every entry point of include/pg_amd.h with its ctypes signature.
_FP = C.POINTER(C.c_float)
_FPP = C.POINTER(_FP)
_IP = C.POINTER(C.c_int)
_CTX = C.c_void_p
SIGNATURES = {
    "pgamd_default_options": (None, [C.POINTER(Options)]),
    "pgamd_set_diagnose_device": (C.c_int, [C.c_int, _IP, _IP, C.POINTER(C.c_void_p)]),
    "pgamd_destroy": (C.c_int, [_CTX]),
    "pgamd_set_options": (C.c_int, [_CTX, C.POINTER(Options)]),
    "pgamd_last_error": (C.c_char_p, []),
    "pgamd_device_info": (C.c_int, [_CTX, C.c_char_p, C.c_int]),
    "pgamd_measure_peaks": (C.c_int, [_CTX, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pgamd_mem_allocate_poisson_gravity": (C.c_int, [_CTX, C.c_int, _FPP, _FPP, _FPP, _FPP]),
    "pgamd_mem_free_poisson_gravity": (C.c_int, [_CTX, _FPP, _FPP, _FPP, _FPP]),
    "pgamd_asyn_poisson_gravity_solver": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                                    C.c_float, C.c_float, C.c_int, C.c_int, C.c_float, C.c_int,
                                                    C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int,
                                                    C.c_int, C.c_int, C.c_int]),
    "pgamd_synchronize": (C.c_int, [_CTX]),
    "pgamd_get_patch_stats": (C.c_int, [_CTX, _IP, _IP, C.c_int]),
    "pgamd_prepare_patch_groups": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_int]),
    "pgamd_scatter_patch_groups": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                             C.c_void_p, C.c_void_p, C.c_int]),
    "pgamd_prepare_patch_groups_device": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                                    C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pgamd_scatter_patch_groups_device": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    "pgamd_set_layout": (C.c_int, [_CTX, C.c_int, C.c_int, C.c_int]),
    "pgamd_get_layout": (C.c_int, [_CTX, C.c_int, _IP, _IP]),
    "pgamd_layout_occupancy": (C.c_int, [_CTX, C.c_int, _IP, _IP]),
    "pgamd_solve_device": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                     C.POINTER(SolverArgs), C.c_void_p, C.c_void_p, C.c_void_p]),
    "pgamd_solve_device_f64": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                         C.POINTER(SolverArgsF64), C.c_void_p, C.c_void_p, C.c_void_p]),
    "pgamd_autotune": (C.c_int, [_CTX, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(SolverArgs),
                                 C.c_int, _IP, _IP, _FP]),
    "pgamd_autotune_save": (C.c_int, [_CTX, C.c_char_p]),
    "pgamd_autotune_load": (C.c_int, [_CTX, C.c_char_p]),
    "pgamd_plan_create": (C.c_int, [_CTX, _IP, _IP, C.c_int, C.POINTER(C.c_void_p)]),
    "pgamd_plan_sizes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "pgamd_plan_offsets": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "pgamd_plan_destroy": (C.c_int, [C.c_void_p]),
    "pgamd_solve_mixed_device": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.POINTER(SolverArgs), _FP, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "pgamd_graph_capture_mixed": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.POINTER(SolverArgs), _FP, C.c_void_p, C.c_void_p,
                                            C.POINTER(C.c_void_p)]),
    "pgamd_graph_capture_device": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_int, C.POINTER(SolverArgs), C.c_void_p, C.c_void_p,
                                             C.POINTER(C.c_void_p)]),
    "pgamd_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pgamd_graph_destroy": (C.c_int, [C.c_void_p]),
    "pgamd_cnn_num_params": (C.c_int, []),
    "pgamd_cnn_set_loss": (C.c_int, [C.c_void_p, C.c_float]),
    "pgamd_cnn_create": (C.c_int, [_CTX, C.c_uint64, C.POINTER(C.c_void_p)]),
    "pgamd_cnn_destroy": (C.c_int, [C.c_void_p]),
    "pgamd_cnn_get_params": (C.c_int, [C.c_void_p, _FP, C.c_int]),
    "pgamd_cnn_set_params": (C.c_int, [C.c_void_p, _FP, C.c_int]),
    "pgamd_cnn_prepare": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(SolverArgs),
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    "pgamd_cnn_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "pgamd_cnn_forward_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "pgamd_cnn_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _FP, C.c_void_p,
                                      C.c_void_p]),
    "pgamd_cnn_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float, _FP,
                                       C.c_void_p]),
    "pgamd_cnn_warm_start": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(SolverArgs),
                                       C.c_void_p, C.c_void_p]),
    "pgamd_cnn_warm_start_cube": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(SolverArgs),
                                            C.c_void_p, C.c_void_p]),
    "pgamd_cnn_warm_start_mixed": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.POINTER(SolverArgs), C.c_void_p, C.c_void_p]),
    "pgamd_cnn_solve_device": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_int, C.POINTER(SolverArgs), C.c_void_p, C.c_void_p,
                                         C.c_int, C.c_void_p]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libpgamd.so; raises FileNotFoundError when it has not been built (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `make -C {HERE}` (no CPU fallback exists)")
    # PyTorch wheels bundle their own HIP runtime under the same soname.  Whichever copy is loaded first
    # serves the whole process; torch only initialises on its own copy, so load torch (when present)
    # before the library resolves libamdhip64.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def geometry(gs: int, gg: int = 1) -> dict:
    n = 8 + 2 * gs
    return {"N": n, "R": n - 2, "C": 4 + 2 * ((gs + 3) // 2), "G": 8 + 2 * gg}


def _check(rc):
    if rc != 0:
        raise PgError(rc, load().pgamd_last_error().decode())


def _vp(a):
    """Raw pointer of a numpy array, torch tensor (data_ptr) or int; None stays NULL."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def solver_args(dt=1e-2, dh=1.0, sor_min_iter=10, sor_max_iter=60, sor_omega=1.69, mg_max_iter=10, mg_npre=3,
                mg_npost=3, mg_tol=1e-6, poi_coeff=4.0 * np.pi * 1e-5, int_scheme=INT_CQUAD, p5=False, poisson=True,
                graacc=True) -> SolverArgs:
    return SolverArgs(dt, dh, sor_min_iter, sor_max_iter, sor_omega, mg_max_iter, mg_npre, mg_npost, mg_tol,
                      poi_coeff, int_scheme, int(p5), int(poisson), int(graacc))


class Context:
    """One device context (pgamd_set_diagnose_device ... pgamd_destroy)."""

    def __init__(self, device: int = 0, nstream: int = 1):
        lib = load()
        npg, ns = C.c_int(0), C.c_int(nstream)
        h = C.c_void_p()
        _check(lib.pgamd_set_diagnose_device(device, C.byref(npg), C.byref(ns), C.byref(h)))
        self.h = h
        self.default_npatch_group = npg.value
        self.nstream = ns.value
        self.opts = Options()
        lib.pgamd_default_options(C.byref(self.opts))
        self._host = None

    def close(self):
        if self.h:
            if self._host is not None:
                self.mem_free()
            _check(load().pgamd_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_options(self, **kw):
        for k, v in kw.items():
            setattr(self.opts, k, v)
        _check(load().pgamd_set_options(self.h, C.byref(self.opts)))

    def device_info(self) -> str:
        buf = C.create_string_buffer(256)
        _check(load().pgamd_device_info(self.h, buf, 256))
        return buf.value.decode()

    def measure_peaks(self):
        """(LDS read B/s, HBM copy B/s) measured on this context's device (pgamd_measure_peaks)."""
        lds, hbm = C.c_double(), C.c_double()
        _check(load().pgamd_measure_peaks(self.h, C.byref(lds), C.byref(hbm)))
        return lds.value, hbm.value

    # --- reference call shape ---------------------------------------------------------------------------
    def mem_allocate(self, npatch_group: int):
        """pgamd_mem_allocate_poisson_gravity; returns numpy views of the pinned host buffers."""
        ptrs = [_FP() for _ in range(4)]
        _check(load().pgamd_mem_allocate_poisson_gravity(self.h, npatch_group, *[C.byref(p) for p in ptrs]))
        g = geometry(self.opts.pot_ghost_size, self.opts.gra_ghost_size)
        n = 8 * npatch_group
        shapes = [(n, g["R"], g["R"], g["R"]), (n, g["C"], g["C"], g["C"]), (n, g["G"], g["G"], g["G"]),
                  (n, 5, 8, 8, 8)]
        views = [np.ctypeslib.as_array(p, shape=s) for p, s in zip(ptrs, shapes)]
        self._host = ptrs
        return views

    def mem_free(self):
        ptrs = self._host or [_FP() for _ in range(4)]
        _check(load().pgamd_mem_free_poisson_gravity(self.h, *[C.byref(p) for p in ptrs]))
        self._host = None

    def asyn_poisson_gravity_solver(self, h_rho, h_pot_in, h_pot_out, h_flu, npatch_group, dt, dh, sor_min_iter,
                                    sor_max_iter, sor_omega, mg_max_iter, mg_npre_smooth, mg_npost_smooth,
                                    mg_tolerated_error, poi_coeff, int_scheme, p5_gradient, poisson, graacc,
                                    gpu_nstream=1):
        for a in (h_rho, h_pot_in, h_pot_out, h_flu):
            if a is not None and not (a.dtype == np.float32 and a.flags.c_contiguous):
                raise TypeError("host arrays must be C-contiguous float32")
        _check(load().pgamd_asyn_poisson_gravity_solver(
            self.h, _vp(h_rho), _vp(h_pot_in), _vp(h_pot_out), _vp(h_flu), npatch_group, dt, dh, sor_min_iter,
            sor_max_iter, sor_omega, mg_max_iter, mg_npre_smooth, mg_npost_smooth, mg_tolerated_error, poi_coeff,
            int_scheme, int(p5_gradient), int(poisson), int(graacc), gpu_nstream))

    def synchronize(self):
        _check(load().pgamd_synchronize(self.h))

    def patch_stats(self, npatch: int):
        iters = np.zeros(npatch, np.int32)
        status = np.zeros(npatch, np.int32)
        _check(load().pgamd_get_patch_stats(self.h, iters.ctypes.data_as(_IP), status.ctypes.data_as(_IP), npatch))
        return iters, status

    def set_layout(self, gs: int, ppw: int, pad: int):
        _check(load().pgamd_set_layout(self.h, gs, ppw, pad))

    def get_layout(self, gs: int):
        a, b = C.c_int(), C.c_int()
        _check(load().pgamd_get_layout(self.h, gs, C.byref(a), C.byref(b)))
        return a.value, b.value

    def layout_occupancy(self, gs: int):
        """(workgroups per CU, patches per CU) of the SOR kernel at the current layout of ghost size gs."""
        a, b = C.c_int(), C.c_int()
        _check(load().pgamd_layout_occupancy(self.h, gs, C.byref(a), C.byref(b)))
        return a.value, b.value

    # --- device-resident ----------------------------------------------------------------------------------
    def solve_device(self, rho, cpot, pot_out, flu, npatch, args: SolverArgs, guess=None, iters=None, status=None,
                     stream=None):
        _check(load().pgamd_solve_device(self.h, _vp(rho), _vp(cpot), _vp(guess), _vp(pot_out), _vp(flu), npatch,
                                         C.byref(args), _vp(iters), _vp(status), stream))

    def solve_device_f64(self, rho, cpot, pot_out, flu, npatch, args: "SolverArgsF64", guess=None, iters=None,
                         status=None, stream=None):
        _check(load().pgamd_solve_device_f64(self.h, _vp(rho), _vp(cpot), _vp(guess), _vp(pot_out), _vp(flu), npatch,
                                             C.byref(args), _vp(iters), _vp(status), stream))

    def autotune(self, gs, rho, cpot, pot_out, npatch, args: SolverArgs, repeats=3):
        """Time every layout of ghost size gs on this batch and keep the fastest: (ppw, pad, ms)."""
        p, d, ms = C.c_int(), C.c_int(), C.c_float()
        _check(load().pgamd_autotune(self.h, gs, _vp(rho), _vp(cpot), _vp(pot_out), npatch, C.byref(args), repeats,
                                     C.byref(p), C.byref(d), C.byref(ms)))
        return p.value, d.value, ms.value

    def autotune_save(self, path):
        _check(load().pgamd_autotune_save(self.h, path.encode()))

    def autotune_load(self, path):
        _check(load().pgamd_autotune_load(self.h, path.encode()))

    # --- mixed batches and graphs ---------------------------------------------------------------------------
    def plan(self, ghost_sizes, schemes=None) -> "Plan":
        return Plan(self, ghost_sizes, schemes)

    def solve_mixed_device(self, plan, rho, cpot, pot_out, flu, args: SolverArgs, guess=None, omega_by_ghost=None,
                           iters=None, status=None, stream=None):
        om = None if omega_by_ghost is None else (C.c_float * 5)(*omega_by_ghost)
        _check(load().pgamd_solve_mixed_device(self.h, plan.h, _vp(rho), _vp(cpot), _vp(guess), _vp(pot_out),
                                               _vp(flu), C.byref(args), om, _vp(iters), _vp(status), stream))

    def graph_mixed(self, plan, rho, cpot, pot_out, flu, args: SolverArgs, guess=None, omega_by_ghost=None,
                    iters=None, status=None) -> "Graph":
        om = None if omega_by_ghost is None else (C.c_float * 5)(*omega_by_ghost)
        h = C.c_void_p()
        _check(load().pgamd_graph_capture_mixed(self.h, plan.h, _vp(rho), _vp(cpot), _vp(guess), _vp(pot_out),
                                                _vp(flu), C.byref(args), om, _vp(iters), _vp(status), C.byref(h)))
        return Graph(h)

    def graph_device(self, rho, cpot, pot_out, flu, npatch, args: SolverArgs, guess=None, iters=None,
                     status=None) -> "Graph":
        h = C.c_void_p()
        _check(load().pgamd_graph_capture_device(self.h, _vp(rho), _vp(cpot), _vp(guess), _vp(pot_out), _vp(flu),
                                                 npatch, C.byref(args), _vp(iters), _vp(status), C.byref(h)))
        return Graph(h)


class Plan:
    """pgamd_plan_*: per-patch ghost sizes and schemes, packed offsets and the persistent work queues."""

    def __init__(self, ctx: Context, ghost_sizes, schemes=None):
        gs = np.ascontiguousarray(ghost_sizes, dtype=np.int32)
        sc = None if schemes is None else np.ascontiguousarray(schemes, dtype=np.int32)
        self.npatch = int(gs.size)
        self.h = C.c_void_p()
        _check(load().pgamd_plan_create(ctx.h, gs.ctypes.data_as(_IP), None if sc is None else sc.ctypes.data_as(_IP),
                                        self.npatch, C.byref(self.h)))
        r, c = C.c_int64(), C.c_int64()
        _check(load().pgamd_plan_sizes(self.h, C.byref(r), C.byref(c)))
        self.rho_floats, self.pot_in_floats = r.value, c.value
        self.rho_off = np.zeros(self.npatch, np.int64)
        self.pot_in_off = np.zeros(self.npatch, np.int64)
        _check(load().pgamd_plan_offsets(self.h, self.rho_off.ctypes.data_as(C.POINTER(C.c_int64)),
                                         self.pot_in_off.ctypes.data_as(C.POINTER(C.c_int64))))

    def close(self):
        if self.h:
            _check(load().pgamd_plan_destroy(self.h))
            self.h = None


class Graph:
    def __init__(self, h):
        self.h = h

    def launch(self, stream=None):
        _check(load().pgamd_graph_launch(self.h, stream))

    def close(self):
        if self.h:
            _check(load().pgamd_graph_destroy(self.h))
            self.h = None


def pack(arrays, offsets, total):
    """Pack per-patch arrays (any shapes) into one float32 vector at the given element offsets."""
    out = np.zeros(total, np.float32)
    for a, o in zip(arrays, offsets):
        f = np.asarray(a, np.float32).ravel()
        out[o:o + f.size] = f
    return out


def _c_array(a, dtype, shape=None, name="array"):
    a = np.asarray(a)
    if a.dtype != dtype or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a C-contiguous {np.dtype(dtype).name} array")
    if shape is not None and a.shape != shape:
        raise ValueError(f"{name} has shape {a.shape}, expected {shape}")
    return a


def prepare_patch_groups(gs, group_first, group_father, fine_flu, fine_sib, coarse_pot, coarse_sib, h_rho=None,
                         h_pot_in=None, h_flu=None, nthreads=0):
    """pgamd_prepare_patch_groups: gather the solver's per-patch arrays for the given patch groups from a fine
    level (fluid [nfine,5,8,8,8], neighbour table [nfine,27]) and its coarse level (potential [ncoarse,8,8,8],
    neighbour table [ncoarse,27]).  Outputs may be given (e.g. the pinned buffers of Context.mem_allocate, as
    flat float32 arrays); returns (rho [8g,R,R,R], pot_in [8g,C,C,C], flu [8g,5,8,8,8], status [8g])."""
    g = geometry(gs)
    gf = _c_array(group_first, np.int32, name="group_first")
    ng = gf.shape[0]
    gfa = _c_array(group_father, np.int32, (ng,), "group_father")
    flu = _c_array(fine_flu, np.float32, name="fine_flu")
    nfine = flu.shape[0]
    _c_array(flu, np.float32, (nfine, 5, 8, 8, 8), "fine_flu")
    fsib = _c_array(fine_sib, np.int32, (nfine, 27), "fine_sib")
    cpot = _c_array(coarse_pot, np.float32, name="coarse_pot")
    ncoarse = cpot.shape[0]
    _c_array(cpot, np.float32, (ncoarse, 8, 8, 8), "coarse_pot")
    csib = _c_array(coarse_sib, np.int32, (ncoarse, 27), "coarse_sib")
    n = 8 * ng
    out = []
    for a, shp in ((h_rho, (n, g["R"], g["R"], g["R"])), (h_pot_in, (n, g["C"], g["C"], g["C"])),
                   (h_flu, (n, 5, 8, 8, 8))):
        if a is None:
            a = np.empty(shp, np.float32)
        elif not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous
                  and a.size >= int(np.prod(shp))):
            raise ValueError(f"output arrays must be C-contiguous float32 numpy arrays of at least {shp}")
        out.append(a)
    st = np.zeros(n, np.int32)
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    _check(load().pgamd_prepare_patch_groups(gs, ng, _vp(gf), _vp(gfa), nfine, _vp(flu), _vp(fsib), ncoarse,
                                             _vp(cpot), _vp(csib), _vp(out[0]), _vp(out[1]), _vp(out[2]), _vp(st),
                                             nthreads))
    return out[0], out[1], out[2], st


def scatter_patch_groups(gg, group_first, h_pot_out, h_flu, fine_pot=None, fine_flu=None, nthreads=0):
    """pgamd_scatter_patch_groups: write the patch part of the solver's potential output and the updated fluid
    back into the fine level's arrays (in place); either may be None."""
    gf = _c_array(group_first, np.int32, name="group_first")
    if fine_pot is None and fine_flu is None:
        raise ValueError("nothing to scatter: fine_pot and fine_flu are both None")
    nfine = (fine_pot if fine_pot is not None else fine_flu).shape[0]
    n, G = 8 * gf.shape[0], 8 + 2 * gg
    for a, need, name in ((h_pot_out if fine_pot is not None else None, n * G ** 3, "h_pot_out"),
                          (h_flu if fine_flu is not None else None, n * 5 * 512, "h_flu")):
        if a is not None and not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous
                                  and a.size >= need):
            raise ValueError(f"{name} must be a C-contiguous float32 array of at least {need} values")
    if fine_pot is not None:
        _c_array(fine_pot, np.float32, (nfine, 8, 8, 8), "fine_pot")
    if fine_flu is not None:
        _c_array(fine_flu, np.float32, (nfine, 5, 8, 8, 8), "fine_flu")
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    _check(load().pgamd_scatter_patch_groups(gg, gf.shape[0], _vp(gf), _vp(h_pot_out), _vp(h_flu), nfine,
                                             _vp(fine_pot), _vp(fine_flu), nthreads))


def prepare_patch_groups_device(gs, ngroup, group_first, group_father, nfine, fine_flu, fine_sib, ncoarse,
                                coarse_pot, coarse_sib, rho, pot_in, flu=None, status=None, stream=None):
    """pgamd_prepare_patch_groups_device: the gather on device-resident levels (device pointers or tensors, the
    layouts of prepare_patch_groups; int32 index tables).  Asynchronous on `stream`."""
    _check(load().pgamd_prepare_patch_groups_device(gs, ngroup, _vp(group_first), _vp(group_father), nfine,
                                                    _vp(fine_flu), _vp(fine_sib), ncoarse, _vp(coarse_pot),
                                                    _vp(coarse_sib), _vp(rho), _vp(pot_in), _vp(flu), _vp(status),
                                                    stream))


def scatter_patch_groups_device(gg, ngroup, group_first, pot_out, flu, nfine, fine_pot=None, fine_flu=None,
                                stream=None):
    """pgamd_scatter_patch_groups_device: the write-back into device-resident levels.  Asynchronous."""
    _check(load().pgamd_scatter_patch_groups_device(gg, ngroup, _vp(group_first), _vp(pot_out), _vp(flu), nfine,
                                                    _vp(fine_pot), _vp(fine_flu), stream))


def solver_args_f64(dt=1e-2, dh=1.0, sor_min_iter=10, sor_max_iter=60, sor_omega=1.69, mg_max_iter=10, mg_npre=3,
                    mg_npost=3, mg_tol=1e-6, poi_coeff=4.0 * np.pi * 1e-5, int_scheme=INT_CQUAD, p5=False,
                    poisson=True, graacc=True) -> SolverArgsF64:
    return SolverArgsF64(dt, dh, sor_min_iter, sor_max_iter, sor_omega, mg_max_iter, mg_npre, mg_npost, mg_tol,
                         poi_coeff, int_scheme, int(p5), int(poisson), int(graacc))


def solve_host_f64(rho, cpot, flu, *, gs, gg=1, scheme=SOR, init_mode=INIT_INTERP, term_mode=TERM_REFERENCE,
                   check_every=1, rel_tol=1e-5, fuse_gravity=True, guess=None, pot_out=None, device=0, **argkw):
    """FP64 twin of solve_host (pgamd_solve_device_f64): double inputs, double outputs."""
    import torch

    args = solver_args_f64(**argkw)
    g = geometry(gs, gg)
    npatch = (rho if rho is not None else pot_out).shape[0]
    with Context(device) as ctx:
        ctx.set_options(pot_ghost_size=gs, gra_ghost_size=gg, pot_scheme=scheme, init_mode=init_mode,
                        term_mode=term_mode, check_every=check_every, rel_tol=rel_tol, fuse_gravity=int(fuse_gravity))
        dev = torch.device("cuda", device)
        t = lambda a: None if a is None else torch.as_tensor(np.ascontiguousarray(a, np.float64)).to(dev)  # noqa
        d_rho, d_cpot, d_guess = t(rho), t(cpot), t(guess)
        d_flu = t(flu) if flu is not None else None
        d_out = (t(pot_out) if pot_out is not None else
                 torch.zeros((npatch, g["G"], g["G"], g["G"]), dtype=torch.float64, device=dev))
        d_it = torch.zeros(npatch, dtype=torch.int32, device=dev)
        d_st = torch.zeros(npatch, dtype=torch.int32, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        ctx.solve_device_f64(d_rho, d_cpot, d_out, d_flu, npatch, args, guess=d_guess, iters=d_it, status=d_st,
                             stream=stream)
        torch.cuda.synchronize(dev)
        return (d_out.cpu().numpy(), None if d_flu is None else d_flu.cpu().numpy(), d_it.cpu().numpy(),
                d_st.cpu().numpy())


def solve_host(rho, cpot, flu, *, gs, gg=1, scheme=SOR, init_mode=INIT_INTERP, term_mode=TERM_REFERENCE,
               check_every=1, rel_tol=1e-5, fuse_gravity=True, guess=None, pot_out=None, device=0, **argkw):
    """Convenience for tests: one batched solve through the reference-shaped host entry point (or the
    device entry point when a guess is given, since the host shape has no guess argument).  Returns
    (pot_out, flu_out, iters, status) as numpy arrays."""
    import torch

    args = solver_args(**argkw)
    g = geometry(gs, gg)
    npatch = (rho if rho is not None else pot_out).shape[0]
    with Context(device) as ctx:
        ctx.set_options(pot_ghost_size=gs, gra_ghost_size=gg, pot_scheme=scheme, init_mode=init_mode,
                        term_mode=term_mode, check_every=check_every, rel_tol=rel_tol, fuse_gravity=int(fuse_gravity))
        dev = torch.device("cuda", device)
        t = lambda a: None if a is None else torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(dev)  # noqa
        d_rho, d_cpot, d_guess = t(rho), t(cpot), t(guess)
        d_flu = t(flu) if flu is not None else None
        d_out = (t(pot_out) if pot_out is not None else torch.zeros((npatch, g["G"], g["G"], g["G"]), device=dev))
        d_it = torch.zeros(npatch, dtype=torch.int32, device=dev)
        d_st = torch.zeros(npatch, dtype=torch.int32, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        ctx.solve_device(d_rho, d_cpot, d_out, d_flu, npatch, args, guess=d_guess, iters=d_it, status=d_st,
                         stream=stream)
        torch.cuda.synchronize(dev)
        return (d_out.cpu().numpy(), None if d_flu is None else d_flu.cpu().numpy(), d_it.cpu().numpy(),
                d_st.cpu().numpy())


class CNN:
    """pgamd_cnn_*: the warm-start network (parameters on the device; BF16 tensors as torch.int16).

    Packed input x0: [n, 10^3*16] BF16 bits; output / target: [n, 512, 8] FP32."""

    NPARAMS = 41568
    INPUT = 10 ** 3 * 16
    OUTPUTS = 512 * 8

    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(load().pgamd_cnn_create(ctx.h, seed, C.byref(self.h)))

    def close(self):
        if self.h:
            _check(load().pgamd_cnn_destroy(self.h))
            self.h = None

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.NPARAMS, np.float32)
        _check(load().pgamd_cnn_get_params(self.h, p.ctypes.data_as(_FP), self.NPARAMS))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(load().pgamd_cnn_set_params(self.h, p.ctypes.data_as(_FP), p.size))

    def save(self, path):
        np.save(path, self.get_params())

    def load_params(self, path):
        self.set_params(np.load(path, allow_pickle=False))

    def prepare(self, rho, cpot, npatch, gs, args: SolverArgs, x0, scale, stream=None):
        _check(load().pgamd_cnn_prepare(self.ctx.h, _vp(rho), _vp(cpot), npatch, gs, C.byref(args), _vp(x0),
                                        _vp(scale), stream))

    def forward(self, x0, count, out, stream=None):
        _check(load().pgamd_cnn_forward(self.h, _vp(x0), count, _vp(out), stream))

    def forward_noalloc(self, x0, count, out, stream=None):
        """pgamd_cnn_forward without allocation (at most 64 samples; activations in the library's fixed array)."""
        _check(load().pgamd_cnn_forward_noalloc(self.h, _vp(x0), count, _vp(out), stream))

    def set_loss(self, res_weight: float = 0.0):
        """Loss = MSE + res_weight * mean squared residual the error leaves in the guess (0: plain MSE)."""
        _check(load().pgamd_cnn_set_loss(self.h, res_weight))

    def loss_grad(self, x0, target, count, grad=None, stream=None) -> float:
        loss = C.c_float()
        _check(load().pgamd_cnn_loss_grad(self.h, _vp(x0), _vp(target), count, C.byref(loss), _vp(grad), stream))
        return loss.value

    def train_step(self, x0, target, count, lr, want_loss=False, stream=None):
        loss = C.c_float()
        _check(load().pgamd_cnn_train_step(self.h, _vp(x0), _vp(target), count, lr,
                                           C.byref(loss) if want_loss else None, stream))
        return loss.value if want_loss else None

    def warm_start(self, rho, cpot, npatch, args: SolverArgs, guess, stream=None):
        _check(load().pgamd_cnn_warm_start(self.ctx.h, self.h, _vp(rho), _vp(cpot), npatch, C.byref(args),
                                           _vp(guess), stream))

    def warm_start_cube(self, rho, cpot, npatch, args: SolverArgs, guess_cube, stream=None):
        """The guess as the whole solve cube [n][N^3] (for init_mode INIT_GUESS_CUBE)."""
        _check(load().pgamd_cnn_warm_start_cube(self.ctx.h, self.h, _vp(rho), _vp(cpot), npatch, C.byref(args),
                                                _vp(guess_cube), stream))

    def warm_start_mixed(self, plan, rho, cpot, args: SolverArgs, guess, stream=None):
        _check(load().pgamd_cnn_warm_start_mixed(self.ctx.h, self.h, plan.h, _vp(rho), _vp(cpot), C.byref(args),
                                                 _vp(guess), stream))

    def solve_device(self, rho, cpot, guess, pot_out, flu, npatch, args: SolverArgs, iters=None, status=None,
                     nchunk=4, stream=None):
        """Warm start + solve, pipelined over nchunk slices (context init_mode INIT_GUESS or INIT_GUESS_CUBE)."""
        _check(load().pgamd_cnn_solve_device(self.ctx.h, self.h, _vp(rho), _vp(cpot), _vp(guess), _vp(pot_out),
                                             _vp(flu), npatch, C.byref(args), _vp(iters), _vp(status), nchunk,
                                             stream))
