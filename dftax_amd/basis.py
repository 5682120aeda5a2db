"""Contracted-GTO basis construction: normalisation, shells, cart->sph.

Follows dftax/basis/loader.py:52-160 and dftax/energy/gto.py:32-125:

* primitive radial norm ``gto_norm(l, a)`` (loader.py:52-64);
* for l <= 1 each cartesian component also gets the contracted norm
  (gto.py:104-125); for l >= 2 only the primitive norm is applied and the
  cartesian->spherical block carries the angular normalisation
  (loader.py:125-131, libcint convention);
* cartesian components in lexicographic order (gto.py:32-66);
* spherical components ordered m = -l..l for l >= 2; l = 0, 1 are identity
  blocks (p stays x, y, z).

The cart->sph blocks are generated here from the real solid-harmonic
expansion rather than read from a table.

The device layout is shell-major: one record per contracted shell
(``l, atom, centre, prim offset, nprim, ao offset``) plus flat primitive
arrays, which is what the integral kernels consume.
"""

from __future__ import annotations

import json
import math
from dataclasses import dataclass
from functools import lru_cache
from pathlib import Path

import numpy as np

from dftax_amd.system import symbol_to_Z

_DATA = Path(__file__).resolve().parent / "data"


def cart_components(l: int) -> list[tuple[int, int, int]]:
    """Cartesian exponents (lx, ly, lz) in lexicographic order (lx desc, then ly desc)."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def nsph(l: int) -> int:
    return 2 * l + 1


def gto_norm(l: int, alpha: float) -> float:
    """Radial normalisation of ``r^l exp(-alpha r^2)`` (loader.py:52-64)."""
    f = (2 ** (2 * l + 3) * math.factorial(l + 1) * (2 * alpha) ** (l + 1.5)
         / (math.factorial(2 * l + 2) * math.sqrt(math.pi)))
    return math.sqrt(f)


def _dfact(n: int) -> float:
    """Double factorial n!! with (-1)!! = 1."""
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def contracted_norm(exps, c_prim, ang) -> float:
    """1/sqrt(self-overlap) of a contracted cartesian GTO (gto.py:104-125)."""
    exps = np.asarray(exps, dtype=np.float64)
    c_prim = np.asarray(c_prim, dtype=np.float64)
    a = exps[:, None] + exps[None, :]
    fac = np.ones_like(a)
    for lk in ang:
        fac = fac * _dfact(2 * lk - 1) / (2.0 * a) ** lk
    s = float(np.sum(c_prim[:, None] * c_prim[None, :] * (np.pi / a) ** 1.5 * fac))
    return 1.0 / math.sqrt(s)


@lru_cache(maxsize=None)
def cart2sph_block(l: int) -> np.ndarray:
    """(ncart(l), nsph(l)) real-solid-harmonic block, columns m = -l..l.

    Coefficient of x^a y^b z^c in sqrt((2l+1)/4pi) * S_lm, with S_lm the real
    solid harmonic in the closed form of Helgaker, Jorgensen & Olsen,
    "Molecular Electronic-Structure Theory" eq. 6.4.47-6.4.50. l = 0, 1 return
    the identity (these shells are normalised per cartesian component).
    """
    if l <= 1:
        return np.eye(ncart(l))
    comps = cart_components(l)
    index = {c: i for i, c in enumerate(comps)}
    out = np.zeros((ncart(l), nsph(l)))
    ang = math.sqrt((2 * l + 1) / (4 * math.pi))
    for col, m in enumerate(range(-l, l + 1)):
        am = abs(m)
        nlm = (1.0 / (2 ** am * math.factorial(l))) * math.sqrt(
            2.0 * math.factorial(l + am) * math.factorial(l - am) / (2.0 if m == 0 else 1.0))
        vm = 0.0 if m >= 0 else 0.5
        for t in range((l - am) // 2 + 1):
            for u in range(t + 1):
                # v runs over vm, vm+1, ... up to floor(am/2 - vm) + vm
                v = vm
                vmax = math.floor(am / 2.0 - vm) + vm
                while v <= vmax + 1e-9:
                    c = ((-1) ** int(round(t + v - vm)) * 0.25 ** t
                         * math.comb(l, t) * math.comb(l - t, am + t)
                         * math.comb(t, u) * math.comb(am, int(round(2 * v))))
                    px = int(round(2 * t + am - 2 * (u + v)))
                    py = int(round(2 * (u + v)))
                    pz = l - 2 * t - am
                    out[index[(px, py, pz)], col] += ang * nlm * c
                    v += 1.0
    return out


@lru_cache(maxsize=None)
def _load(kind: str) -> dict:
    name = "basis_orbital.json" if kind == "orbital" else "basis_aux.json"
    with open(_DATA / name) as fh:
        return json.load(fh)


def available_bases() -> list[str]:
    return [k for k in _load("orbital") if not k.startswith("_")] + \
           [k for k in _load("aux") if not k.startswith("_")]


def element_shells(name: str, symbol: str) -> list[tuple[int, np.ndarray, np.ndarray]]:
    """Raw ``(l, exps, coefs)`` shells of one element for a basis name."""
    key = name.lower()
    orb = _load("orbital")
    aux = _load("aux")
    if key in orb:
        table = orb[key]
        if symbol not in table:
            raise ValueError(f"basis {name!r} has no entry for {symbol}")
        return [(int(s["l"]), np.asarray(s["exps"], float), np.asarray(s["coefs"], float))
                for s in table[symbol]]
    if key in aux:
        table = aux[key]
        if symbol not in table:
            raise ValueError(f"basis {name!r} has no entry for {symbol}")
        out = []
        for s in table[symbol]:
            for k in range(int(s["n"])):
                e = float(s["alpha_max"]) / float(s["beta"]) ** k
                out.append((int(s["l"]), np.array([e]), np.array([1.0])))
        return out
    raise ValueError(f"unknown basis {name!r}; available: {available_bases()}")


@dataclass
class BasisSet:
    """Shell-major basis for a molecule.

    Arrays (one entry per contracted shell): ``l``, ``atom``, ``centre`` (3),
    ``prim0``/``nprim`` into ``exps``/``coefs``, ``ao0`` (first AO in the final
    span) and ``cart0`` (first cartesian AO). ``coefs`` already include the
    primitive norm and, for l <= 1, the contracted norm. ``spherical`` selects
    (2l+1) functions per shell (l >= 2) over ncart(l).
    """

    l: np.ndarray
    atom: np.ndarray
    centre: np.ndarray
    prim0: np.ndarray
    nprim: np.ndarray
    ao0: np.ndarray
    cart0: np.ndarray
    exps: np.ndarray
    coefs: np.ndarray
    spherical: bool
    nao: int
    nao_cart: int

    @property
    def nshell(self) -> int:
        return int(self.l.shape[0])

    @property
    def max_l(self) -> int:
        return int(self.l.max())

    def nfunc(self, ish: int) -> int:
        l = int(self.l[ish])
        return nsph(l) if (self.spherical and l >= 2) else ncart(l)

    def cart2sph(self) -> np.ndarray:
        """Block-diagonal (nao_cart, nao) transform."""
        T = np.zeros((self.nao_cart, self.nao))
        for i in range(self.nshell):
            l = int(self.l[i])
            blk = cart2sph_block(l) if self.spherical else np.eye(ncart(l))
            c0, a0 = int(self.cart0[i]), int(self.ao0[i])
            T[c0:c0 + blk.shape[0], a0:a0 + blk.shape[1]] = blk
        return T

    def atom_ao_slices(self, natm: int) -> list[slice]:
        out = []
        for a in range(natm):
            idx = np.where(self.atom == a)[0]
            if idx.size == 0:
                out.append(slice(0, 0))
                continue
            lo = int(self.ao0[idx[0]])
            hi = int(self.ao0[idx[-1]]) + self.nfunc(int(idx[-1]))
            out.append(slice(lo, hi))
        return out


def build_basis(symbols, coords_bohr, name: str, *, spherical: bool) -> BasisSet:
    """Assemble the shell-major basis of a molecule (atoms in order)."""
    coords = np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3)
    ls, atoms, centres, prim0, nprim, ao0, cart0 = [], [], [], [], [], [], []
    exps_all, coefs_all = [], []
    ao = cart = p = 0
    cache: dict[str, list] = {}
    for a, sym in enumerate(symbols):
        if sym not in cache:
            shells = []
            for l, e, c in element_shells(name, sym):
                cp = c * np.array([gto_norm(l, x) for x in e])
                if l <= 1:
                    # contracted norm is the same for every component of an s/p shell
                    cp = cp * contracted_norm(e, cp, cart_components(l)[0])
                shells.append((l, e, cp))
            cache[sym] = shells
        for l, e, cp in cache[sym]:
            ls.append(l)
            atoms.append(a)
            centres.append(coords[a])
            prim0.append(p)
            nprim.append(len(e))
            ao0.append(ao)
            cart0.append(cart)
            exps_all.extend(e.tolist())
            coefs_all.extend(cp.tolist())
            p += len(e)
            cart += ncart(l)
            ao += nsph(l) if (spherical and l >= 2) else ncart(l)
    return BasisSet(
        l=np.asarray(ls, np.int32), atom=np.asarray(atoms, np.int32),
        centre=np.asarray(centres, np.float64).reshape(-1, 3),
        prim0=np.asarray(prim0, np.int32), nprim=np.asarray(nprim, np.int32),
        ao0=np.asarray(ao0, np.int32), cart0=np.asarray(cart0, np.int32),
        exps=np.asarray(exps_all, np.float64), coefs=np.asarray(coefs_all, np.float64),
        spherical=bool(spherical), nao=ao, nao_cart=cart,
    )
