"""CPU checks of the C ABI library: it loads and exports every symbol the
header declares (no device calls)."""
import ctypes
import os

import pytest

from terrace_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_declares_the_bound_symbols():
    declared = set(_lib.declared_symbols())
    assert declared == set(_lib._SIGS), declared ^ set(_lib._SIGS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libterrace_hip.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in _lib.declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_product_has_no_oracle_or_cpu_fallback():
    pkg = os.path.join(ROOT, "terrace_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text, f
