"""The C-ABI library loads and exports every entry point the public header
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from aerocfd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "aerocfd_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(acfd_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "acfd_create" in names and "acfd_step" in names and "acfd_inviscid_rhs" in names
    assert set(names) == set(_lib.EXPORTS), set(names) ^ set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_config_struct_layout_matches_header():
    # 8 doubles, 7 ints, 5 reserved ints
    assert ctypes.sizeof(_lib.Config) == 8 * 8 + 12 * 4
    assert ctypes.sizeof(_lib.Subface) == 11 * 4


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_LIB", None)
    try:
        _lib.load()
    except RuntimeError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("load() must fail without the library")
