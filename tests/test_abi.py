"""The C-ABI library loads and exports every symbol include/xca_amd.h declares; host-only entry points
behave (no device compute is called here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from xca_amd import _lib
from xca_amd.dlr import DLRGrid, kernel

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "xca_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xca_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding types every one of them
    assert set(names) <= set(_lib.EXPORTS), set(names) - set(_lib.EXPORTS)


def test_dlr_build_matches_python_grid():
    L = _lib.lib()
    rf = np.zeros(200)
    it = np.zeros(200)
    r = L.xca_dlr_build(100.0, 1e-10, 200, _lib.dptr(rf), _lib.dptr(it))
    g = DLRGrid(10.0, 100.0, 1e-10)
    assert r == g.r
    assert np.array_equal(rf[:r], g.rf) and np.array_equal(it[:r], g.it)
    # too small an output buffer is an error, not an overflow
    assert L.xca_dlr_build(100.0, 1e-10, 4, _lib.dptr(rf), _lib.dptr(it)) < 0


@pytest.mark.parametrize("t,om", [(0.1, 3.0), (-0.2, -5.0), (0.5, 0.0), (0.0, 40.0)])
def test_dlr_kernel(t, om):
    assert _lib.lib().xca_dlr_kernel(t, om) == pytest.approx(kernel(t, om), rel=1e-15)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_LIB_PATH", tmp_path / "absent.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.XcaError, match="missing"):
        _lib.lib()
