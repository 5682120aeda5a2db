"""Analytic nuclear forces (dftax/ks/forces.py:128-270 surface).

``forces(mol, xc, result)`` returns F = -dE/dR (Ha/Bohr, natom x 3) of the
closed-shell DF-KS energy at the converged density, the quantity dftax
obtains by differentiating the energy rebuilt on the nuclear coordinates:
the basis and auxiliary centres and the Becke grid follow their atoms and
P = 2 Z (Z^T S Z)^-1 Z^T is held at the converged occupied orbitals Z. The
library (``dfks_forces``) contracts the nuclear derivatives of the one-,
two- and three-centre integrals with the densities they meet in dE/dR, and
forms the XC gradient on the atom-centred grid (AO Hessians, the grid
points' motion with their atoms and the Becke weight derivatives).
"""

from __future__ import annotations

import ctypes

import numpy as np

from dftax_amd import _native
from dftax_amd.grid import Becke, atom_bragg, atom_centred_points, becke

PARTS = ("nuclear", "one_electron", "coulomb", "exchange", "xc")


def gradient_parts(ks, result, grid: Becke | None = None) -> dict:
    """dE/dR split into the energy's terms (each with its share of the density
    response to the overlap), natom x 3 each, keyed by PARTS."""
    torch = __import__("torch")
    c = ks.coulomb
    if c.plan.aux_chunk != 0:
        raise ValueError("forces need the materialised DF tensor: df(chunk=None), as dftax's forces (forces.py:163-181)")
    spec = grid if grid is not None else getattr(ks, "grid_spec", None) or becke()
    nocc = ks.nocc[0]
    C = result.mo_coeff[0] if result.mo_coeff.dim() == 3 else result.mo_coeff
    Co = torch.as_tensor(C, dtype=torch.float64, device=c.device)[:, :nocc]
    cm = Co.T.contiguous()                       # column-major nao x nocc
    pts, raw, own = atom_centred_points(ks.symbols, ks.atom_coords, spec)
    pts = np.ascontiguousarray(pts, np.float64)
    raw = np.ascontiguousarray(raw, np.float64)
    own = np.ascontiguousarray(own, np.int32)
    bragg = np.ascontiguousarray(atom_bragg(ks.symbols), np.float64)
    natom = len(ks.symbols)
    out = np.zeros((len(PARTS), natom, 3))
    _native.check(c.L.dfks_forces(c.handle, ctypes.c_void_p(cm.data_ptr()), nocc, pts.shape[0], _native.f64p(pts),
                                  _native.f64p(raw), _native.i32p(own), _native.f64p(bragg), _native.f64p(out),
                                  c._stream()))
    return {k: out[i] for i, k in enumerate(PARTS)}


def forces(mol, xc=None, result=None, *, grid: Becke | None = None, coulomb=None, ks=None) -> np.ndarray:
    """Nuclear forces F = -dE/dR (Ha/Bohr), shape (natom, 3).

    ``forces(mol, xc, result, grid=becke(...), coulomb=df(...))`` rebuilds the
    energy functional for ``mol`` as dftax does (forces.py:128-181; the DF
    backend is materialised); ``ks=`` reuses an already built KS (the one the
    SCF ran on). ``result`` is the converged KSResult.
    """
    from dftax_amd.ks import KS, DFSpec, df
    if result is None:
        raise ValueError("forces need the converged SCF result")
    if ks is None:
        spec = coulomb if coulomb is not None else df()
        if isinstance(spec, DFSpec) and isinstance(spec.chunk, int):
            raise ValueError("forces need the materialised DF backend: df(...) without an explicit chunk")
        spec = DFSpec(auxbasis=spec.auxbasis, chunk=None, screen=spec.screen, spherical=spec.spherical,
                      engine=spec.engine)
        ks = KS(mol, xc, grid=grid, coulomb=spec)
    g = gradient_parts(ks, result, grid)
    return -sum(g[k] for k in PARTS)
