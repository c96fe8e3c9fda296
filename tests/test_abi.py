"""The C-ABI library loads and exports every entry point include/stokes_to.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from stokes_to_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "stokes_to.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(st_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "st_create" in names and "st_solve" in names and len(names) > 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.DEFAULT_LIB):
        pytest.fail("libstokes_hip.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.DEFAULT_LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared()) == set(_lib.SIGNATURES), set(declared()) ^ set(_lib.SIGNATURES)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load(str(tmp_path / "nope.so"))


def test_params_struct_layout_matches_header():
    # field order of st_params / st_stats must follow the header
    src = open(os.path.join(ROOT, "include", "stokes_to.h")).read()
    body = src[src.index("typedef struct {", src.index("ST_INIT_LEARNED")):src.index("} st_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b([a-z_0-9]+)\s*[,;]", body)
    assert names == [f for f, _ in _lib.StParams._fields_]
