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


def test_ladder_tune_table_without_gpu():
    """The autotuner's kernel table: candidates by size and precision, set / get / remove, and a
    kernel that cannot run a shape is refused (no GPU call is made)."""
    import pytest

    from dgamore_amd import autotune as A

    A.clear()
    names = lambda n, t, p: [A.kernel_name(k) for k in A.candidates(n, t, p)]  # noqa: E731
    assert names(16, 1, _lib.PLAIN_FP64)[:2] == ["reg0", "reg1"] and "blocked0" not in names(16, 1, 0)
    assert names(120, 1, _lib.PLAIN_FP64) == ["reg3", "col3", "blocked0", "blocked1", "blocked2", "blocked3",
                                              "blocked4"]
    assert names(720, 9, _lib.PLAIN_FP64) == ["blocked0", "blocked1", "blocked2", "blocked3", "blocked4"]
    assert names(1080, 9, _lib.PLAIN_FP64) == ["blocked1", "blocked3", "blocked4"]  # one row per thread: n <= 768
    assert names(720, 9, _lib.MIXED_BF16X3_IR) == ["blocked0", "blocked1", "blocked2", "blocked3", "blocked4",
                                                   "blocked5"]
    assert names(1080, 9, _lib.MIXED_BF16X3_IR) == ["blocked0", "blocked1", "blocked2", "blocked3", "blocked5"]
    assert names(1600, 1, _lib.PLAIN_FP64) == []
    kid = A.candidates(720, 9, 0)[1]
    A.set_choice(720, 9, 0, kid)
    assert A.get_choice(720, 9, 0) == kid and A.get_choice(720, 9, 2) is None
    A.set_choice(720, 9, 0, None)
    assert A.get_choice(720, 9, 0) is None
    with pytest.raises(ValueError):
        A.set_choice(1080, 9, 0, 3 * 256 + 0)
    with pytest.raises(ValueError):
        A.candidates(0, 1, 0)
    A.clear()
