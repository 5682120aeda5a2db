"""dftax_amd: dftax's density-fitted Kohn-Sham SCF iteration on MI355X.

Same builder surface as dftax (KS, df, becke, scf, PBE0, B3LYP, CAM-B3LYP, minao,
core, Molecule) and the second-order ``newton`` solver; ``df(..., engine=hip(...))`` runs the DF Coulomb/exchange
and XC terms, DIIS and the Fock eigensolve through libdfks.so, a C-ABI
library of hand-written gfx950 kernels.
"""

from dftax_amd.forces import forces
from dftax_amd.grid import Becke, becke
from dftax_amd.guess import core, minao
from dftax_amd.ks import KS, DFSpec, HipDFKS, HipSpec, df, hip
from dftax_amd.newton import newton
from dftax_amd.scf import KSResult, scf
from dftax_amd.system import WATER, Molecule, serving_stream, tree_molecule, water_cluster
from dftax_amd.xc import B3LYP, CAMB3LYP, PBE, PBE0

__all__ = ["KS", "df", "hip", "becke", "Becke", "scf", "newton", "KSResult", "core", "minao", "Molecule", "PBE", "PBE0",
           "B3LYP", "CAMB3LYP", "DFSpec", "HipSpec", "HipDFKS", "water_cluster", "tree_molecule", "serving_stream", "WATER", "forces"]
