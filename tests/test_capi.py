"""The C-ABI library loads and exports every symbol include/*.h declares.
CPU only: no compute call reaches the device here."""
import ctypes
import os
import re

from tspeed_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("tspeed_amd.h", "tspeed_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ts[ph]_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_the_headers():
    assert declared_symbols() == set(capi.SIGNATURES)


def test_null_and_bad_arguments_are_rejected_without_a_device():
    L = capi.lib()
    assert L.tsp_create(0, 99, 1, ctypes.byref(ctypes.c_void_p())) == capi.TSP_ERR_UNSUPPORTED
    assert L.tsp_create(0, 0, 1, None) == capi.TSP_ERR_ARG
    # split BF16 products exist only for multi-column contexts
    for prec in (capi.TSP_BF16X3, capi.TSP_BF16X5):
        assert L.tsp_create(0, prec, 1, ctypes.byref(ctypes.c_void_p())) == capi.TSP_ERR_UNSUPPORTED
    assert L.tsp_run(None, 1) == capi.TSP_ERR_ARG
    assert L.tsp_set_face_form(None, 1, None) == capi.TSP_ERR_ARG
    assert L.tsp_steps_done(None) == -1
    assert L.tsp_last_error(None) == b"null context"
