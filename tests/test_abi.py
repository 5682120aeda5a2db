"""The C-ABI library loads and exports every entry point include/dfks.h declares.

No compute calls here (no GPU in this environment); only argument-validation
error paths that return before touching the device.
"""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "dfks.h").read_text()
    return sorted(set(re.findall(r"\b(dfks_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_header():
    from dftax_amd import _native
    L = _native.lib()
    names = _declared()
    assert len(names) >= 16
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_native.EXPORTS)


def test_error_paths_without_device():
    from dftax_amd import _native
    L = _native.lib()
    h = ctypes.c_void_p()
    rc = L.dfks_create(None, None, ctypes.byref(h))
    assert rc != 0
    assert b"null" in L.dfks_last_error()
    assert L.dfks_destroy(None) == 0
    assert L.dfks_sizes(None, None) != 0
    assert L.dfks_stage_times(None, None, 0) != 0
    with pytest.raises(_native.DfksError):
        _native.check(L.dfks_scf_step(None, None, None))


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from dftax_amd import _native
    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setattr(_native, "LIB_PATH", tmp_path / "libdfks.so")
    with pytest.raises(_native.DfksError):
        _native.lib()
