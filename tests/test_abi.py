"""CPU checks of the C ABI: the library loads and exports every function
that include/gdict.h declares (no compute calls: there is no GPU here)."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gdict.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(gd_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_reference_surface():
    names = declared_functions()
    for required in ["gd_lsm_create", "gd_lsm_bulk_build", "gd_lsm_insert",
                     "gd_lsm_erase", "gd_lsm_lookup", "gd_lsm_range_sum",
                     "gd_lsm_successor", "gd_lsm_carry_status",
                     "gd_lsm_resident_bytes", "gd_flix_build",
                     "gd_flix_insert", "gd_flix_remove",
                     "gd_flix_lookup_sorted", "gd_flix_successor_sorted",
                     "gd_flix_rebuild", "gd_flix_total_keys",
                     "gd_last_error_message"]:
        assert required in names


def test_library_exports_every_declared_symbol(gd):
    lib = ctypes.CDLL(gd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header(gd):
    names = set(declared_functions()) - {"gd_last_error_message"}
    assert names == set(gd.SIGNATURES), names ^ set(gd.SIGNATURES)


def test_create_error_without_gpu_is_reported(gd):
    """Creating a handle with an invalid capacity fails before touching the
    GPU and reports std::invalid_argument's message."""
    lib = gd.lib()
    cfg = gd._Config(1 << 33, 16, 0, 0, 0)
    h = ctypes.c_void_p()
    st = lib.gd_lsm_create(ctypes.byref(cfg), ctypes.byref(h))
    assert st == gd.GD_INVALID_ARGUMENT
    assert b"capacity" in lib.gd_last_error_message(None)
