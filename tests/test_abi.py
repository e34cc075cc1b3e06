"""The C-ABI library loads and exports every symbol include/gemini_fk.h declares."""
import ctypes
import os
import re

from gemini_unified_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "gemini_fk.h")).read()
    return sorted(set(re.findall(r"^int (gm_\w+)\(", txt, flags=re.M)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_binding_argument_counts_match_header():
    """Every ctypes prototype has as many arguments as the header declares (a missing or extra
    argument would shift every later one)."""
    txt = open(os.path.join(ROOT, "include", "gemini_fk.h")).read()
    L = _lib.lib()
    for name, params in re.findall(r"^int (gm_\w+)\(([^;]*?)\);", txt, flags=re.M | re.S):
        n = 0 if params.strip() in ("", "void") else params.count(",") + 1
        argtypes = getattr(L, name).argtypes
        if argtypes is None:
            assert n == 0, name
        else:
            assert len(argtypes) == n, (name, len(argtypes), n)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_no_device_is_a_clean_error():
    """gm_ctx_create on a host without a usable device index returns an error code."""
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.gm_ctx_create(10_000, ctypes.byref(h))
    assert rc != 0 and not h.value
    assert len(_lib.last_error(None)) > 0
