"""The C ABI: every function declared in include/*.h is exported by the
built library, and the ctypes table of the Python package binds exactly
those (no GPU needed: loading the library does not touch a device)."""
from __future__ import annotations

import ctypes as C
import glob
import os
import re

import pytest

from dynamical_model_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(mlc_[a-z0-9_]+)\s*\(", re.M)


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        names |= set(DECL.findall(open(h).read()))
    return names


def test_headers_declare_the_api():
    names = declared()
    for must in ("mlc_create", "mlc_step", "mlc_init", "mlc_lac_create", "mlc_mlp_create"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_headers():
    assert declared() == set(_lib.EXPORTS)
