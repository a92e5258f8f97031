"""The C-ABI library loads on a CPU-only host and exports every entry point
include/ftl_hip.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ftl_hip.h")


def _declared():
    names = set()
    for hdr in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if hdr.endswith(".h"):
            text = open(os.path.join(ROOT, "include", hdr)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(ftl_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    from fast_tensor_leverage_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libftl_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert "ftl_sampler_create" in declared
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from fast_tensor_leverage_amd import _lib
    assert _declared() == set(_lib.SIGNATURES)


def test_no_device_is_reported_not_faked():
    from fast_tensor_leverage_amd import _lib
    L = _lib.load()
    n = L.ftl_device_count()
    assert n >= 0
    if n == 0:
        import numpy as np
        import fast_tensor_leverage_amd as ftl
        with pytest.raises(_lib.FTLError):
            ftl.Sampler([np.zeros((4, 2))], 4, 2)
