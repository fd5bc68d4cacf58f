"""The C-ABI library loads and exports every symbol that include/dgami.h declares (CPU only)."""
import ctypes
import os
import re

from dgamore_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dgami.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dgami_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "dgami_ladder_solve_sum" in syms and "dgami_bubble_irrq" in syms
    assert len(syms) >= 15


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_version_and_error_text_without_gpu():
    lib = _lib.load()
    assert lib.dgami_version() >= 1
    assert isinstance(_lib.last_error(), str)
