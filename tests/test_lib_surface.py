"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
public header declares, and reports errors without a GPU call."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    src = open(os.path.join(ROOT, "include", "fhehip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fhehip_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header_symbols():
    import ctypes
    import fhehip
    L = ctypes.CDLL(fhehip.LIB_PATH)
    syms = _header_symbols()
    assert len(syms) >= 7
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding declares exactly the header surface
    assert sorted(fhehip.SIGNATURES) == syms


def test_version_and_null_checks():
    import ctypes
    import fhehip
    lib = fhehip.lib()
    assert b"gfx950" in lib.fhehip_version()
    h = ctypes.c_void_p()
    rc = lib.fhehip_context_create(None, 0, 16, None, ctypes.byref(h))
    assert rc == fhehip.ERR_INVALID_ARG
    assert lib.fhehip_last_error()
