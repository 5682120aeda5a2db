"""Analytic nuclear forces on the GPU against the oracle's central differences.

The golden vectors (tests/golden/make_forces.py) hold, per case, the oracle's
converged occupied orbitals Z and the Richardson-extrapolated central
differences of the frozen-density energy E(R) (dftax/ks/forces.py:128-270:
basis, auxiliary centres and Becke grid follow the atoms; P = 2 Z (Z^T S Z)^-1 Z^T),
term by term. The library's gradient at the same Z is compared term by term
at 2e-7 Ha/Bohr (the differences carry ~1e-9 of truncation and rounding
error; the metric of both cases keeps every eigenvalue, so d(V+) = -V+ dV V+
is the exact derivative). A full forces() call after a GPU SCF is compared
at 1e-6 (the two SCFs' convergence), and the forces sum to zero.
"""
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "forces_oracle.npz"
CASES = {"water_svp_pbe0": (["O", "H", "H"], "def2-svp", "PBE0", (40, 110)),
         "dimer_tzvp_b3lyp": (["O", "H", "H", "O", "H", "H"], "def2-tzvp", "B3LYP", (30, 86))}
PARTS = ("nuclear", "one_electron", "coulomb", "exchange", "xc")


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _build(case, g):
    from dftax_amd import KS, becke, df, hip
    from dftax_amd.system import Molecule
    sym, basis, xc, (nr, na) = CASES[case]
    mol = Molecule(sym, g[f"{case}_coords"], basis, spherical=True)
    ks = KS(mol, xc, grid=becke(nr, na), coulomb=df(chunk=None, engine=hip()))
    return mol, ks, becke(nr, na)


@pytest.mark.parametrize("case", list(CASES))
def test_gradient_terms_match_oracle_differences(case, torch_cuda):
    from dftax_amd.forces import gradient_parts
    g = np.load(GOLDEN)
    print(f"case: forces {case}", flush=True)
    mol, ks, grid = _build(case, g)
    res = SimpleNamespace(mo_coeff=torch_cuda.tensor(g[f"{case}_Z"], device="cuda"))
    parts = gradient_parts(ks, res, grid)
    ref_parts = -g[f"{case}_parts"]          # gradients: minus the oracle's forces
    for k, name in enumerate(PARTS):
        err = np.abs(parts[name] - ref_parts[k]).max()
        print(f"  {name:13s} max |d| {err:.2e}  (max |dE/dR| {np.abs(ref_parts[k]).max():.3e})", flush=True)
    total = sum(parts[n] for n in PARTS)
    err_tot = np.abs(-total - g[f"{case}_forces"]).max()
    print(f"  total forces max |d| {err_tot:.2e}; sum over atoms {np.abs(total.sum(0)).max():.1e}", flush=True)
    for k, name in enumerate(PARTS):
        assert np.abs(parts[name] - ref_parts[k]).max() < 2e-7, name
    assert err_tot < 2e-7
    assert np.abs(total.sum(0)).max() < 1e-8


def test_forces_after_gpu_scf(torch_cuda):
    """forces(mol, xc, result) from the GPU SCF's own orbitals."""
    from dftax_amd import forces, scf
    g = np.load(GOLDEN)
    mol, ks, grid = _build("water_svp_pbe0", g)
    r = scf(ks, e_tol=1e-12, d_tol=1e-10)
    F = forces(mol, "PBE0", r, grid=grid, ks=ks)
    err = np.abs(F - g["water_svp_pbe0_forces"]).max()
    print(f"forces after the GPU SCF (E {r.e_tot:.10f}, oracle {float(g['water_svp_pbe0_e_tot']):.10f}): "
          f"max |d| {err:.2e}", flush=True)
    assert r.converged and err < 1e-6
    assert np.abs(F.sum(0)).max() < 1e-8
