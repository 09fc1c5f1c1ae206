"""The C-ABI library loads and exports every symbol include/erosion_stokes.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

from erosion_amd import stokes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "erosion_stokes.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)) - {"defined"})


def test_header_declares_the_reference_entry_points():
    syms = declared_symbols()
    assert "stokessolver_" in syms and "computeshearstress_" in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(stokes.LIB_PATH):
        pytest.skip("libstokes_hip.so not built")
    lib = ctypes.CDLL(stokes.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(stokes, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(stokes, "_lib", None)
    with pytest.raises(stokes.StokesError):
        stokes.lib()
