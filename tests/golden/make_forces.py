"""Golden nuclear forces from the CPU oracle: central differences of the
frozen-density energy E(R) (the quantity dftax/ks/forces.py:128-270
differentiates), Richardson-extrapolated from steps h = 4e-3 and 2h Bohr.
The step is set by the Coulomb term: rounding in the fitted E_J (the metric's
condition number ~2e6 times the FP64 epsilon, ~5e-11 Ha) puts ~1e-6 Ha/Bohr of
noise into the net force at h = 1e-4 and ~3e-8 at 4e-3, while the extrapolated
differences at 4e-3 and 8e-3 agree to 3e-8 (Coulomb) and 2e-9 (the other
terms) on the dimer.

Cases (geometries in Angstrom, off equilibrium so every force is sizable):
  water_svp_pbe0   water, def2-SVP, PBE0, becke(40, 110)
  dimer_tzvp_b3lyp water dimer, def2-TZVP (f orbitals, g auxiliaries), B3LYP, becke(30, 86)
Stored per case (forces_oracle.npz, keys prefixed with the case name): the
coordinates (Bohr), the oracle's converged occupied orbitals Z (the frozen
density P = 2 Z (Z^T S Z)^-1 Z^T), the forces and the per-term forces
(nuclear, one-electron, J, K, XC; each term's central difference).

Run: python tests/golden/make_forces.py   (about a minute on 8 cores)
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

OUT = Path(__file__).resolve().parent / "forces_oracle.npz"

CASES = {
    "water_svp_pbe0": ("O 0.0 0.0 0.0; H 0.80 0.62 0.05; H -0.70 0.55 -0.08", "def2-svp", "PBE0", (40, 110)),
    "dimer_tzvp_b3lyp": ("O 0.000 0.000 0.000; H 0.757 0.587 0.000; H -0.757 0.587 0.000;"
                         "O 0.300 -0.400 2.900; H 0.900 -0.100 3.600; H -0.500 0.100 3.100",
                         "def2-tzvp", "B3LYP", (30, 86)),
}


def main() -> None:
    from dftax_amd.grid import becke
    from dftax_amd.system import Molecule
    from oracle import ref
    out = {}
    for name, (geom, basis, xc, (nr, na)) in CASES.items():
        mol = Molecule.from_xyz(geom, basis, spherical=True)
        grid = becke(nr, na)
        oks = ref.OracleKS(mol, xc, grid=grid)
        r = ref.scf(oks, e_tol=1e-12, d_tol=1e-10)
        Z = r["C"][:, :oks.nocc]
        h = 4e-3   # with 2h; see the module docstring
        F1, P1 = ref.fd_forces(mol, xc, Z, grid, h=h, parts=True)
        F2, P2 = ref.fd_forces(mol, xc, Z, grid, h=2 * h, parts=True)
        F = (4 * F1 - F2) / 3
        Fp = (4 * P1 - P2) / 3
        print(f"{name}: SCF {r['e_tot']:.10f} ({r['n_iter']} it); |F|max {np.abs(F).max():.3e}; "
              f"Richardson correction {np.abs(F - F1).max():.1e}", flush=True)
        out.update({f"{name}_coords": mol.coords, f"{name}_Z": Z, f"{name}_forces": F, f"{name}_parts": Fp,
                    f"{name}_e_tot": np.array(r["e_tot"])})
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
