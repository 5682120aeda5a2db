"""The allocation-free Becke weights entry point against dfks_becke_weights.

dfks_becke_weights_noalloc keeps the partition's pair tables in a fixed device
array of the library instead of allocating; its weights must equal those of
the allocating entry point bit for bit: both run the same weight kernel, and
the pair tables, computed on the device here and on the host there, round every
operation alike (no fused multiply-adds; correctly rounded sqrt and division).
Two calls in a row, the second on a smaller system, check that nothing of the
first call's contents leaks into the second.
"""
import ctypes

import numpy as np
import pytest


def _grid(mol, spec):
    from dftax_amd.grid import atom_bragg, atom_centred_points
    pts, raw, own = atom_centred_points(mol.symbols, mol.coords, spec)
    return pts, raw, own.astype(np.int32), atom_bragg(mol.symbols)


@pytest.mark.gpu
def test_noalloc_becke_weights_match_allocating_entry_point():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dftax_amd import _native as N
    from dftax_amd.grid import becke
    from dftax_amd.system import water_cluster
    L = N.lib()
    stream = N.current_stream()
    cases = [water_cluster((2, 2, 2), seed=0), water_cluster((1, 1, 2), seed=5)]
    for k, mol in enumerate(cases):
        pts, raw, own, bragg = _grid(mol, becke(50, 110))
        w_ref = N.becke_weights(pts, raw, own, mol.coords, bragg)
        dev = {name: torch.tensor(a, device="cuda") for name, a in
               dict(pts=np.ascontiguousarray(pts), raw=raw, own=own, xyz=np.ascontiguousarray(mol.coords),
                    bragg=bragg).items()}
        w = torch.full((len(raw),), float("nan"), dtype=torch.float64, device="cuda")
        print(f"noalloc call {k}: {len(mol.symbols)} atoms, {len(raw)} points", flush=True)
        N.check(L.dfks_becke_weights_noalloc(len(raw), ctypes.c_void_p(dev["pts"].data_ptr()),
                                             ctypes.c_void_p(dev["raw"].data_ptr()),
                                             ctypes.c_void_p(dev["own"].data_ptr()), len(mol.symbols),
                                             ctypes.c_void_p(dev["xyz"].data_ptr()),
                                             ctypes.c_void_p(dev["bragg"].data_ptr()),
                                             ctypes.c_void_p(w.data_ptr()), stream))
        torch.cuda.synchronize()
        got = w.cpu().numpy()
        rel = np.abs(got - w_ref) / np.maximum(np.abs(w_ref), 1e-300)
        print(f"  max rel |dw| {rel.max():.2e}, exact matches {np.mean(got == w_ref):.4f}", flush=True)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, w_ref)


def test_noalloc_rejects_too_many_atoms():
    """Sizes beyond the fixed array are refused before any device work."""
    from dftax_amd import _native as N
    L = N.lib()
    rc = L.dfks_becke_weights_noalloc(10, None, None, None, 1025, None, None, None, None)
    assert rc == 1
    assert b"at most 1024 atoms" in L.dfks_last_error()
