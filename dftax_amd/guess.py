"""Initial densities: minao (default) and core (dftax/ks/guess.py).

``minao`` projects sto-3g atomic ground-state occupations onto the molecule's
basis through the per-atom cross overlap, ``P_A = C f C^T`` with
``C = S_AA^-1 S_AM`` (guess.py:296-337), assembles the blocks and rescales
Tr(P S) to the electron count (guess.py:602-616). ``core`` fills the
eigenvectors of the core Hamiltonian. Overlaps are computed by the library.
"""

from __future__ import annotations

import threading
from dataclasses import dataclass

import numpy as np

from dftax_amd import _native
from dftax_amd.basis import BasisSet, build_basis, ncart, nsph
from dftax_amd.system import symbol_to_Z


@dataclass(frozen=True)
class CoreSpec:
    pass


@dataclass(frozen=True)
class MinAOSpec:
    basis: str = "sto-3g"


def core() -> CoreSpec:
    return CoreSpec()


def minao(basis: str = "sto-3g") -> MinAOSpec:
    return MinAOSpec(basis=basis)


_MADELUNG = ((1, 0), (2, 0), (2, 1), (3, 0), (3, 1))


def ground_state_config(Z: int):
    """((n, l, count), ...) by Madelung filling (guess.py:190-214, Z <= 18 here)."""
    left, out = Z, []
    for n, l in _MADELUNG:
        c = min(left, 2 * (2 * l + 1))
        if c > 0:
            out.append((n, l, c))
        left -= c
        if left == 0:
            break
    return tuple(out)


def _sub_basis(basis: BasisSet, idx) -> BasisSet:
    ls, prim0, nprim, ao0, cart0, exps, coefs = [], [], [], [], [], [], []
    p = ao = cart = 0
    for i in idx:
        l = int(basis.l[i])
        a, n = int(basis.prim0[i]), int(basis.nprim[i])
        ls.append(l); prim0.append(p); nprim.append(n); ao0.append(ao); cart0.append(cart)
        exps.extend(basis.exps[a:a + n]); coefs.extend(basis.coefs[a:a + n])
        p += n
        ao += nsph(l) if (basis.spherical and l >= 2) else ncart(l)
        cart += ncart(l)
    return BasisSet(l=np.array(ls, np.int32), atom=np.zeros(len(ls), np.int32), centre=basis.centre[list(idx)].copy(),
                    prim0=np.array(prim0, np.int32), nprim=np.array(nprim, np.int32), ao0=np.array(ao0, np.int32),
                    cart0=np.array(cart0, np.int32), exps=np.array(exps), coefs=np.array(coefs),
                    spherical=basis.spherical, nao=ao, nao_cart=cart)


def _union(b1: BasisSet, b2: BasisSet) -> BasisSet:
    off = b1.exps.shape[0]
    return BasisSet(l=np.concatenate([b1.l, b2.l]), atom=np.zeros(b1.nshell + b2.nshell, np.int32),
                    centre=np.concatenate([b1.centre, b2.centre]),
                    prim0=np.concatenate([b1.prim0, b2.prim0 + off]), nprim=np.concatenate([b1.nprim, b2.nprim]),
                    ao0=np.concatenate([b1.ao0, b2.ao0 + b1.nao]),
                    cart0=np.concatenate([b1.cart0, b2.cart0 + b1.nao_cart]),
                    exps=np.concatenate([b1.exps, b2.exps]), coefs=np.concatenate([b1.coefs, b2.coefs]),
                    spherical=b1.spherical, nao=b1.nao + b2.nao, nao_cart=b1.nao_cart + b2.nao_cart)


# Projected atomic blocks by (element, minimal basis, the atom's own shells): a block
# depends on the shells' exponents and coefficients only (both bases sit on the same
# centre), so molecules that share elements and a basis share it; the lock makes the
# cache safe for concurrent builds.
_BLOCKS: dict = {}
_BLOCKS_LOCK = threading.Lock()


def _block_key(sym: str, min_name: str, sub: BasisSet):
    return (sym, min_name, bool(sub.spherical), sub.l.tobytes(), sub.nprim.tobytes(),
            np.ascontiguousarray(sub.exps).tobytes(), np.ascontiguousarray(sub.coefs).tobytes())


def minao_density(symbols, coords, basis: BasisSet, min_name: str = "sto-3g") -> np.ndarray:
    """Block-diagonal projected minimal-basis density (not yet renormalised)."""
    P = np.zeros((basis.nao, basis.nao))
    slices = basis.atom_ao_slices(len(symbols))
    cache = {}
    for a, sym in enumerate(symbols):
        if sym not in cache:
            idx = np.where(basis.atom == a)[0]
            sub = _sub_basis(basis, idx)
            key = _block_key(sym, min_name, sub)
            with _BLOCKS_LOCK:
                hit = _BLOCKS.get(key)
            if hit is not None:
                cache[sym] = hit
                P[slices[a], slices[a]] = hit
                continue
            minb = build_basis([sym], np.asarray(coords)[a:a + 1], min_name, spherical=True)
            if minb.max_l >= 2 and not sub.spherical:
                raise ValueError("minao with a cartesian basis needs an s/p minimal basis")
            per_l: dict[int, list[float]] = {}
            for _, l, c in ground_state_config(symbol_to_Z(sym)):
                per_l.setdefault(l, []).append(float(c))
            f = []
            for ish in range(minb.nshell):
                l = int(minb.l[ish])
                lst = per_l.get(l, [])
                c = lst.pop(0) if lst else 0.0
                f.extend([c / (2 * l + 1)] * (2 * l + 1))
            if any(per_l.values()):
                raise ValueError(f"minimal basis {min_name!r} cannot hold the configuration of {sym}")
            u = _union(sub, minb)
            S = _native.overlap(u)
            n1 = sub.nao
            C = np.linalg.solve(S[:n1, :n1], S[:n1, n1:])
            cache[sym] = (C * np.asarray(f)[None, :]) @ C.T
            with _BLOCKS_LOCK:
                _BLOCKS[key] = cache[sym]
        P[slices[a], slices[a]] = cache[sym]
    return P
