"""The C ABI: the library loads without a GPU and exports every function of include/synthe_hip.h."""

from __future__ import annotations

import ctypes as ct
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "synthe_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(synthe_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "synthe_asynth" in names and "synthe_lte_spectrum" in names and len(names) >= 12


def test_library_exports_every_declared_symbol():
    from kurucz_py_joss_submission_amd import _native

    if not _native.LIB_PATH.exists():
        pytest.skip("libsynthe_hip.so not built")
    lib = ct.CDLL(str(_native.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(_native.SYMBOLS)
    L = _native.load_library()
    assert L.synthe_abi_version() == 2


def test_host_grid_and_error_path():
    """synthe_wavelength_grid is a host routine; bad arguments return a status and a message."""
    from kurucz_py_joss_submission_amd import _native

    if not _native.LIB_PATH.exists():
        pytest.skip("libsynthe_hip.so not built")
    L = _native.load_library()
    assert L.synthe_wavelength_grid(500.0, 510.0, 300000.0, None, 0) == 5941
    assert L.synthe_wavelength_grid(500.0, 510.0, -1.0, None, 0) < 0
    assert b"resolution" in L.synthe_last_error()
    assert L.synthe_continuum(None, 1, 1, None, 2, None, None, None, None, None, 1, None, None, None) < 0
    assert L.synthe_asynth(None, None, None, 0, None, None, None, None, 0.0, None, None, None, None,
                           None, None) < 0
