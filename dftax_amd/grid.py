"""Becke/Lebedev molecular grid: the host-side layout.

Radial shells, NWChem angular pruning and the atom-centred point layout follow
dftax/grid/grid.py:104-255 and dftax/grid/becke.py:50-65; the fuzzy-Voronoi
partition (becke.py:68-108) runs on the GPU (``dfks_becke_weights``) because it
is O(ng * natom^2). Points whose final weight is at or below ``cutoff`` are
dropped afterwards (dftax/ks/energy.py:247-254).

The Lebedev-Laikov tables in ``data/lebedev.npz`` hold unit-sphere points and
weights normalised to sum 1 (the surface weight is 4*pi*w).
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache
from pathlib import Path

import numpy as np

from dftax_amd.system import symbol_to_Z

_DATA = Path(__file__).resolve().parent / "data"

# Bragg-Slater radii (Angstrom) for H..Ne, converted with the 0.52917721092
# Bohr length and rounded to 8 decimals as in dftax/grid/becke.py:23-37.
_BRAGG_ANGSTROM = [0.0, 0.35, 1.40, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 1.50]
BRAGG_BOHR = np.round(np.array(_BRAGG_ANGSTROM) / 0.52917721092, 8)

LEBEDEV_LADDER = (6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194,
                  230, 266, 302, 350, 434, 590, 770)

_NWCHEM_ALPHAS = np.array((
    (0.25, 0.5, 1.0, 4.5),
    (0.1667, 0.5, 0.9, 3.5),
    (0.1, 0.4, 0.8, 2.5),
))


def bragg_radius(Z: int) -> float:
    if not 0 < Z < len(BRAGG_BOHR):
        raise ValueError(f"Bragg radius tabulated for H-Ne only, got Z={Z}")
    return float(BRAGG_BOHR[Z])


@dataclass(frozen=True)
class Becke:
    """Grid spec, same fields and defaults as dftax's ``becke()`` (grid.py:31-86)."""

    n_radial: int = 75
    lebedev: int = 302
    chunk: object = "auto"
    prune: str | None = "nwchem"
    r_max: float | None = 45.0
    cutoff: float | None = 1e-15


def becke(n_radial: int = 75, lebedev: int = 302, *, chunk="auto",
          prune: str | None = "nwchem", r_max: float | None = 45.0,
          cutoff: float | None = 1e-15) -> Becke:
    return Becke(n_radial=n_radial, lebedev=lebedev, chunk=chunk, prune=prune,
                 r_max=r_max, cutoff=cutoff)


@lru_cache(maxsize=1)
def _lebedev_tables() -> dict:
    with np.load(_DATA / "lebedev.npz") as npz:
        return {int(k[1:]): np.asarray(npz[k], dtype=np.float64) for k in npz.files}


def lebedev_grid(n: int):
    t = _lebedev_tables()
    if n not in t:
        raise ValueError(f"no Lebedev grid with {n} points; available {sorted(t)}")
    return t[n][:, :3], t[n][:, 3]


def becke_radial(n: int, scale: float):
    """Second-kind Gauss-Chebyshev nodes mapped by r = R(1+x)/(1-x)."""
    i = np.arange(1, n + 1)
    theta = i * np.pi / (n + 1)
    x = np.cos(theta)
    r = scale * (1.0 + x) / (1.0 - x)
    drdx = 2.0 * scale / (1.0 - x) ** 2
    w = (np.pi / (n + 1)) * np.sin(theta) * r ** 2 * drdx
    return r, w


def _region_orders(n_ang: int) -> tuple:
    """Lebedev orders of the five NWChem regions of r / R_bragg for a
    bonding-region order n_ang: 50 and 86 inside, the order one rung below
    n_ang on both sides of the bonding region (50 / 74 / 50 for n_ang = 50)."""
    if n_ang == 50:
        return (50, 74, 74, 74, 50)
    if n_ang not in LEBEDEV_LADDER[4:]:
        raise ValueError(f"lebedev={n_ang} is not a Lebedev order")
    below = LEBEDEV_LADDER[LEBEDEV_LADDER.index(n_ang) - 1]
    return (50, 86, below, n_ang, below)


def nwchem_orders(Z: int, rads: np.ndarray, n_ang: int) -> np.ndarray:
    """Angular order per radial shell by the NWChem region rule (grid.py:116-144):
    the region of a shell is the number of bounds that r / R_bragg exceeds."""
    if n_ang < 50:
        return np.full(rads.shape[0], n_ang, dtype=np.int64)
    bounds = _NWCHEM_ALPHAS[0 if Z <= 2 else (1 if Z <= 10 else 2)]
    region = np.searchsorted(bounds, rads / bragg_radius(Z), side="left")
    return np.asarray(_region_orders(n_ang), dtype=np.int64)[region]


@lru_cache(maxsize=None)
def atom_shell_blocks(Z: int, n_radial: int, n_ang: int, prune, r_max):
    """Per-element radial shells as runs of equal angular order:
    ((radii, radial weights, order), ...) in increasing radius."""
    scale = bragg_radius(Z) if Z == 1 else 0.5 * bragg_radius(Z)
    r, wr = becke_radial(n_radial, scale)
    if r_max is not None:
        inside = r <= r_max
        r, wr = r[inside], wr[inside]
    if prune == "nwchem":
        orders = nwchem_orders(Z, r, n_ang)
    elif prune is None:
        orders = np.full(r.shape[0], n_ang, dtype=np.int64)
    else:
        raise ValueError(f"unknown prune scheme {prune!r}")
    cuts = np.flatnonzero(np.diff(orders)) + 1
    return tuple((rr.copy(), ww.copy(), int(oo[0]))
                 for rr, ww, oo in zip(np.split(r, cuts), np.split(wr, cuts), np.split(orders, cuts))
                 if rr.size)


@lru_cache(maxsize=None)
def _atom_template(Z: int, n_radial: int, n_ang: int, prune, r_max):
    """One atom's points relative to its nucleus and their raw weights, shell
    block by shell block (the same for every atom of element Z)."""
    pts, ws = [], []
    for r, wr, order in atom_shell_blocks(Z, n_radial, n_ang, prune, r_max):
        ang, aw = lebedev_grid(order)
        pts.append((r[:, None, None] * ang[None, :, :]).reshape(-1, 3))
        ws.append((wr[:, None] * (4.0 * np.pi * aw)[None, :]).reshape(-1))
    return np.concatenate(pts), np.concatenate(ws)


def atom_centred_points(symbols, coords_bohr, spec: Becke):
    """Unpartitioned grid: points (ng,3), raw weights (ng,), owner atom (ng,),
    atom by atom; each atom's block is its element's template shifted to it."""
    coords = np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3)
    r_max = None if spec.r_max is None else float(spec.r_max)
    tmpl = [_atom_template(symbol_to_Z(s), spec.n_radial, spec.lebedev, spec.prune, r_max) for s in symbols]
    lens = np.array([t[1].shape[0] for t in tmpl], dtype=np.int64)
    pts = np.repeat(coords, lens, axis=0) + np.concatenate([t[0] for t in tmpl])
    ws = np.concatenate([t[1] for t in tmpl])
    own = np.repeat(np.arange(len(symbols), dtype=np.int32), lens)
    return pts, ws, own


def atom_bragg(symbols) -> np.ndarray:
    return np.array([bragg_radius(symbol_to_Z(s)) for s in symbols])
