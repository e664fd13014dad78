"""The C-ABI library loads without a GPU and exports every symbol the header declares."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(em3d_[a-z0-9_]+)\s*\(", src))
    return names


def test_header_declares_the_binding_table():
    from em3d_amd import _lib
    assert _declared() == set(_lib.SIGNATURES)


def test_library_loads_and_exports_every_symbol():
    from em3d_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libem3d_hip.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    assert _lib.lib().em3d_version().startswith(b"em3d_hip")


def test_product_refuses_non_hip_backend():
    from em3d_amd.operator import require_hip
    from types import SimpleNamespace
    with pytest.raises(TypeError, match="Backend.hip"):
        require_hip(SimpleNamespace(device="cpu"))
