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


def _water_desc(nspin, spin):
    """A system descriptor for water/def2-SVP (host arrays only)."""
    import numpy as np
    from dftax_amd import _native
    from dftax_amd.basis import build_basis
    from dftax_amd.system import WATER, Molecule
    mol = Molecule.from_xyz(WATER, "def2-svp", spherical=True)
    orb = _native.BasisArrays(build_basis(mol.symbols, mol.coords, "def2-svp", spherical=True))
    aux = _native.BasisArrays(build_basis(mol.symbols, mol.coords, "even-tempered-jk", spherical=True))
    xyz = np.ascontiguousarray(mol.coords, np.float64)
    z = np.ascontiguousarray(mol.atom_charges(), np.float64)
    g = np.zeros((1, 3))
    w = np.ones(1)
    keep = (orb, aux, xyz, z, g, w)
    desc = _native.SystemDesc(orb.desc, aux.desc, len(z), _native.f64p(xyz), _native.f64p(z), 1,
                              _native.f64p(g), _native.f64p(w), 2, 10, nspin, spin)
    return desc, keep


def test_nspin_in_the_descriptor():
    """P[nspin, nao, nao] (dftax/ks/terms.py:520-527): the descriptor carries
    nspin and spin = n_alpha - n_beta; inconsistent pairs are refused before
    any device work: nspin = 1 needs spin = 0, nspin = 2 needs nelec + spin
    even and 0 <= spin <= nelec (water has 10 electrons)."""
    from dftax_amd import _native
    L = _native.lib()
    h = ctypes.c_void_p()
    for spin in (1, -2, 12):
        desc, _keep = _water_desc(2, spin)
        assert L.dfks_create(ctypes.byref(desc), None, ctypes.byref(h)) == 1
        assert b"nspin = 2 needs" in L.dfks_last_error()
    desc, _keep = _water_desc(1, 2)
    assert L.dfks_create(ctypes.byref(desc), None, ctypes.byref(h)) == 1
    assert b"spin = 0" in L.dfks_last_error()
    desc, _keep = _water_desc(3, 0)
    assert L.dfks_create(ctypes.byref(desc), None, ctypes.byref(h)) == 1
    assert not h.value
    # dfks_energy_fock takes the spin-block count of P explicitly
    assert L.dfks_energy_fock(None, None, 1, None, None, None) == 1
