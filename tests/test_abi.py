"""The C-ABI library builds, loads and exports every symbol that
include/mom5_gpu.h declares; argument errors come back as codes with a
message, never as a crash (no GPU needed for these checks)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mom_gpu_\w+)\s*\(", txt, re.M))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("mom_gpu_create", "mom_gpu_horz_advect_tracer", "mom_gpu_vert_advect_tracer",
                 "mom_gpu_advect_sweby_all", "mom_gpu_invtri", "mom_gpu_vert_diffuse_implicit",
                 "mom_gpu_update_ocean_barotropic", "mom_gpu_compute_tmask_limit",
                 "mom_gpu_update_tracer_explicit", "mom_gpu_tracer_content"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from mom5_amd import _lib
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTED) >= declared_symbols()


def test_null_context_is_an_error_not_a_crash():
    from mom5_amd import _lib
    lib = _lib.load()
    rc = lib.mom_gpu_sync(None)
    assert rc == -1
    assert b"NULL context" in lib.mom_gpu_last_error(None)
    rc = lib.mom_gpu_horz_advect_tracer(None, 9, *([None] * 7), 0, 1.0, *([None] * 5))
    assert rc == -1


def test_create_rejects_wide_halo_before_touching_the_gpu():
    from mom5_amd import _lib
    lib = _lib.load()
    dom = _lib.Domain(isd=-1, ied=12, jsd=0, jed=9, isc=1, iec=10, jsc=1, jec=8, nk=5,
                      cyclic_x=1, cyclic_y=0)
    g = _lib.Grid()
    ctx = C.c_void_p()
    rc = lib.mom_gpu_create(C.byref(dom), C.byref(g), C.byref(ctx))
    assert rc == -1
    assert b"halo" in lib.mom_gpu_last_error(None)


def test_product_has_no_cpu_fallback():
    """The package never imports the oracle."""
    pkg = os.path.join(ROOT, "mom5_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".F90", ".cpp")):
                txt = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in txt.replace("Oracle", ""), fn
