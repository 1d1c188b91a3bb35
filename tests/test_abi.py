"""CPU checks of the C-ABI boundary: libitomo.so loads and exports every function that
include/itomo.h declares, and the Python binding covers exactly that set.  No compute
call is made (there is no GPU here)."""
import ctypes
import os
import re

import pytest

from iterative_tomographic_reconstruction_algorithms_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "itomo.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(itomo_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = declared()
    assert "itomo_radon_dd" in names and "itomo_radon_rd" in names
    assert "itomo_sirt_iterate" in names and "itomo_mlem_iterate" in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libitomo.so not built: run __graft_entry__.build()")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert sorted(_lib.EXPORTS) == declared()


def test_version_string_without_gpu():
    L = _lib.lib()
    assert b"gfx950" in L.itomo_version()


def test_plan_rejects_bad_arguments_without_touching_the_gpu():
    # argument validation happens before any HIP call
    L = _lib.lib()
    h = ctypes.c_void_p()
    ang = (ctypes.c_float * 2)(0.1, 0.2)
    rc = L.itomo_plan_create(ctypes.byref(h), 7, 16, ctypes.cast(ang, ctypes.c_void_p), 2, 0)
    assert rc == -1 and b"projector" in L.itomo_last_error()
    rc = L.itomo_plan_create(ctypes.byref(h), 0, 0, ctypes.cast(ang, ctypes.c_void_p), 2, 0)
    assert rc == -1 and b"npix" in L.itomo_last_error()
    nan = (ctypes.c_float * 1)(float("nan"))
    rc = L.itomo_plan_create(ctypes.byref(h), 1, 16, ctypes.cast(nan, ctypes.c_void_p), 1, 0)
    assert rc == -2
    # distance-driven sweep of rows with sin < 0 (theta in (pi, 2pi)) is undefined in
    # the reference C, so the plan refuses it
    bad = (ctypes.c_float * 1)(4.0)
    rc = L.itomo_plan_create(ctypes.byref(h), 0, 16, ctypes.cast(bad, ctypes.c_void_p), 1, 0)
    assert rc == -2 and b"sin <= 0" in L.itomo_last_error()
    # pixel-driven: the bin rule must be 0 (nearest) or 1 (linear)
    rc = L.itomo_plan_create_pd(ctypes.byref(h), 16, ctypes.cast(ang, ctypes.c_void_p), 2, 2, 0)
    assert rc == -1 and b"method" in L.itomo_last_error()
    rc = L.itomo_radon_pd(None, 16, ctypes.cast(ang, ctypes.c_void_p), 2, 2, 0, None, None)
    assert rc == -1
    rc = L.itomo_plan_create_ss(ctypes.byref(h), 16, ctypes.cast(ang, ctypes.c_void_p), 2, -1, 0)
    assert rc == -1 and b"method" in L.itomo_last_error()
