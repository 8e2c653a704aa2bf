"""CPU-side checks of the C ABI: the library loads and exports every entry point include/*.h declares,
and the ctypes binding covers exactly that set.  No compute call is made (no GPU here)."""
import ctypes
import glob
import os
import re

import pytest

import gpu_testbed_poisson_gravity_amd as pg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(pgamd_[a-z0-9_]+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(pg.LIB_PATH):
        pytest.fail("libpgamd.so is not built: run `make -C gpu_testbed_poisson_gravity_amd`")
    return ctypes.CDLL(pg.LIB_PATH)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"pgamd_asyn_poisson_gravity_solver", "pgamd_mem_allocate_poisson_gravity", "pgamd_synchronize",
            "pgamd_set_diagnose_device", "pgamd_mem_free_poisson_gravity"} <= names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_header():
    assert set(pg.SIGNATURES) == declared_symbols()


def test_binding_loads_and_sets_signatures():
    lib = pg.load()
    assert lib.pgamd_asyn_poisson_gravity_solver.restype is ctypes.c_int


def test_default_options_without_device(lib):
    o = pg.Options()
    pg.load().pgamd_default_options(ctypes.byref(o))
    assert (o.pot_ghost_size, o.gra_ghost_size, o.pot_scheme, o.check_every) == (5, 1, pg.SOR, 1)


def test_null_context_is_an_error_not_a_crash(lib):
    L = pg.load()
    assert L.pgamd_synchronize(None) == pg.ERR_ARG
    assert L.pgamd_set_options(None, None) == pg.ERR_ARG
    assert b"null" in L.pgamd_last_error()
    assert L.pgamd_destroy(None) == 0
