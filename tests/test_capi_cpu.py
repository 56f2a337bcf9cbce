"""CPU-side checks of the C ABI library: it loads, exports every symbol that
include/gkc_hip.h declares, and rejects bad parameters without a GPU."""
from __future__ import annotations

import ctypes
import os
import re

from gkc_amd import _lib, capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "gkc_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gkc_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.hip_path())
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(capi.SIGNATURES), "ctypes binding out of sync with the header"


def test_invalid_params_rejected():
    L = capi.lib()
    for bad in [capi.Params(0, 11, 0, 64, 1, 0, 1, 0), capi.Params(31, 2, 0, 64, 1, 0, 1, 0),
                capi.Params(31, 33, 0, 64, 1, 0, 1, 0), capi.Params(65, 11, 0, 64, 1, 0, 1, 0),
                capi.Params(31, 11, 0, 0, 1, 0, 1, 0), capi.Params(31, 11, 0, 2048, 1, 0, 1, 0),
                capi.Params(31, 11, 0, 64, 2, 0, 1, 0)]:
        assert not L.gkc_create(0, ctypes.byref(bad))
        assert b"invalid" in L.gkc_last_error()


def test_null_context_is_an_error():
    L = capi.lib()
    assert L.gkc_count(None, 0) == -1
    assert L.gkc_reset(None) == -1
