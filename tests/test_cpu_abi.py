"""The C-ABI library loads and exports every symbol include/gic.h declares
(no GPU calls), and the Python binding declares the same set."""
import ctypes
import os
import re

from genetic_amd import gic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "gic.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gic_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(gic.LIB_PATH)
    names = declared()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert sorted(gic.SIGNATURES) == declared()


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(gic, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(gic, "_LIB", None)
    try:
        gic.lib()
    except RuntimeError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("expected a RuntimeError")
