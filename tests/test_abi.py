"""The C-ABI library and the C++ facade load, and export every entry point
include/lbx.h declares (no GPU call is made)."""
import ctypes
import re
from pathlib import Path

from lambrex_amd import lbx

ROOT = Path(__file__).resolve().parents[1]


def declared_entry_points():
    text = (ROOT / "include" / "lbx.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbx_\w+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_entry_points()
    assert "lbx_iterate" in names and "lbx_ctx_create" in names
    assert len(names) >= 18


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(lbx.LIB_PATH))
    missing = [n for n in declared_entry_points() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(lbx.SIGNATURES) == declared_entry_points()
    lbx.load()


def test_facade_library_loads():
    ctypes.CDLL(str(lbx.LIB_PATH))
    ctypes.CDLL(str(ROOT / "lambrex_amd" / "liblambrex_amd.so"))


def test_last_error_without_context_is_a_string():
    lib = lbx.load()
    assert isinstance(lib.lbx_last_error(None), bytes)
