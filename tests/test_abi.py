"""The C-ABI library: it loads and exports every symbol include/*.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import pathlib
import re

import pytest

from qutip_amd import _lib

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(qa_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "qa_rk_run" in names and "qa_csr_matvec_z" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_load_sets_signatures_without_gpu():
    L = _lib.load()
    assert L.qa_version() == 1
    assert L.qa_last_error() is not None


def test_missing_library_fails_loudly(tmp_path):
    import importlib
    mod = importlib.reload(_lib)
    try:
        with pytest.raises(mod.QutipAmdUnavailable):
            mod.load(tmp_path / "nope.so")
    finally:
        importlib.reload(_lib)
