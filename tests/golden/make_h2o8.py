"""Golden vectors of the (H2O)_8 def2-TZVP PBE0 workload from the CPU oracle.

(H2O)_8 from the seeded water-cluster generator (2 x 2 x 2 lattice, 3.1 A,
seed 0), def2-TZVP orbitals (spherical), the even-tempered auxiliary set,
PBE0, the default Becke grid (75, 302, NWChem pruning, r_max 45, cutoff
1e-15). The oracle here is the build's FP64 restatement (oracle/ref.py,
oracle/md_oracle.c); it takes a few minutes on 8 cores, which is why its
outputs are stored instead of recomputed in the GPU test.

Stored (h2o8_pbe0_oracle.npz):
  coords         atom coordinates (Bohr) the generator gave
  int3c_idx/val  20000 random (m, n, P) entries of (mn|P)
  p0_sum, p0_v   checksums of the renormalised minao density P0 (sum, P0 v)
  fock_rows      24 random rows of F(P0) and their indices
  fock_v         F(P0) v for two fixed random vectors v
  e_parts        E_tot, E_one, E_J, E_K, E_xc at P0
  e_scf          SCF energy (DIIS, e_tol 1e-9, d_tol 1e-6) and its iterations

Run: python tests/golden/make_h2o8.py
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

OUT = Path(__file__).resolve().parent / "h2o8_pbe0_oracle.npz"


def main() -> None:
    from dftax_amd.grid import becke
    from dftax_amd.system import water_cluster
    from oracle import ref
    t0 = time.time()
    mol = water_cluster((2, 2, 2), seed=0, basis="def2-tzvp")
    oks = ref.OracleKS(mol, "PBE0", grid=becke())
    nao, naux = oks.int3c.shape[1], oks.int3c.shape[2]
    print(f"nao {nao} naux {naux} ngrid {oks.grid_w.shape[0]} built in {time.time() - t0:.0f} s", flush=True)
    rng = np.random.default_rng(8)
    idx = np.stack([rng.integers(0, nao, 20000), rng.integers(0, nao, 20000), rng.integers(0, naux, 20000)], 1)
    val = oks.int3c[idx[:, 0], idx[:, 1], idx[:, 2]]
    P0 = ref.minao_density(mol, oks.basis)
    P0 *= mol.nelectron / np.sum(P0 * oks.S)
    v = rng.standard_normal((2, nao))
    e, F, parts = oks.energy_fock(P0)
    rows = np.sort(rng.choice(nao, 24, replace=False))
    print(f"E(P0) {e:.12f} in {time.time() - t0:.0f} s", flush=True)
    r = ref.scf(oks, e_tol=1e-9, d_tol=1e-6, P0=P0, verbose=True)
    print(f"SCF {r['e_tot']:.12f} ({r['n_iter']} it, converged {r['converged']}) in {time.time() - t0:.0f} s")
    np.savez_compressed(
        OUT, coords=mol.coords, int3c_idx=idx.astype(np.int32), int3c_val=val,
        p0_sum=np.array(P0.sum()), p0_v=P0 @ v[0], vecs=v, fock_rows=F[rows], fock_row_idx=rows,
        fock_v=np.stack([F @ v[0], F @ v[1]]),
        e_parts=np.array([e, parts["e1"], parts["eJ"], parts["eK"], parts["exc"]]),
        e_scf=np.array([r["e_tot"], r["n_iter"], float(r["converged"])]))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
