"""Molecular systems and the seeded structure generators used as workloads.

``Molecule`` mirrors dftax's ``Molecule`` (dftax/system/molecule.py:43-106):
element symbols, coordinates in Bohr, charge, spin (= 2S), basis name and the
spherical/cartesian choice. ``water_cluster`` builds the (H2O)_n lattice
clusters and ``tree_molecule`` the saturated acyclic C/N/O molecules of the
serving stream.
"""

from __future__ import annotations

import numpy as np

# CODATA 2018 Bohr radius, 1 Bohr = 0.529177210903 Angstrom
# (same constant as dftax/system/molecule.py:14-15).
ANGSTROM_TO_BOHR = 1.0 / 0.529177210903

_Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8,
      "F": 9, "Ne": 10}


def symbol_to_Z(symbol: str) -> int:
    s = symbol.strip().capitalize()
    if s not in _Z:
        raise ValueError(f"element {symbol!r} is outside this build's range (H-Ne)")
    return _Z[s]


def _parse_atom_string(atom: str):
    symbols, coords = [], []
    for token in atom.replace("\n", ";").split(";"):
        token = token.strip()
        if not token:
            continue
        parts = token.split()
        if len(parts) != 4:
            raise ValueError(f"cannot parse atom line: {token!r}")
        symbols.append(parts[0])
        coords.append([float(x) for x in parts[1:]])
    return symbols, np.asarray(coords, dtype=np.float64)


class Molecule:
    """Atoms, geometry (Bohr), charge/spin and basis name.

    Same constructor surface as dftax/system/molecule.py:43-93
    (``from_xyz`` parses ``"O 0 0 0; H ..."`` strings, Angstrom by default).
    """

    def __init__(self, symbols, coords_bohr, basis: str, charge: int = 0,
                 spin: int = 0, spherical: bool = False):
        self.symbols = [s.strip().capitalize() for s in symbols]
        self.coords = np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3)
        self.basis = basis
        self.charge = int(charge)
        self.spin = int(spin)
        self.spherical = bool(spherical)
        if len(self.symbols) != self.coords.shape[0]:
            raise ValueError("symbols and coords length mismatch")
        nelec = self.nelectron
        if (nelec + self.spin) % 2 != 0:
            raise ValueError(
                f"inconsistent (nelec={nelec}, spin={self.spin}): nelec+spin must be even")
        if nelec - self.spin < 0:
            raise ValueError(f"spin={self.spin} too large for nelec={nelec}")

    @classmethod
    def from_xyz(cls, atom: str, basis: str, *, unit: str = "angstrom",
                 charge: int = 0, spin: int = 0, spherical: bool = False):
        symbols, coords = _parse_atom_string(atom)
        if unit.lower().startswith("ang"):
            coords = coords * ANGSTROM_TO_BOHR
        elif not unit.lower().startswith("b"):
            raise ValueError(f"unit must be 'angstrom' or 'bohr', got {unit!r}")
        return cls(symbols, coords, basis, charge=charge, spin=spin, spherical=spherical)

    def atom_coords(self) -> np.ndarray:
        return self.coords

    def atom_charges(self) -> np.ndarray:
        return np.array([symbol_to_Z(s) for s in self.symbols], dtype=np.float64)

    @property
    def natm(self) -> int:
        return len(self.symbols)

    @property
    def nelectron(self) -> int:
        return int(sum(symbol_to_Z(s) for s in self.symbols) - self.charge)


# ---------------------------------------------------------------------------
# Workload generators
# ---------------------------------------------------------------------------

WATER = "O 0 0 0; H 0.757 0.587 0; H -0.757 0.587 0"


def _random_rotation(rng) -> np.ndarray:
    """Uniform random rotation from a normalised Gaussian quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def _water_monomer() -> np.ndarray:
    """O at the origin, O-H 0.957 Angstrom, H-O-H 104.5 degrees (Angstrom)."""
    r, th = 0.957, np.deg2rad(104.5)
    return np.array([
        [0.0, 0.0, 0.0],
        [r * np.sin(th / 2), r * np.cos(th / 2), 0.0],
        [-r * np.sin(th / 2), r * np.cos(th / 2), 0.0],
    ])


def water_cluster(shape=(2, 2, 2), *, seed: int = 0, spacing: float = 3.1,
                  jitter: float = 0.3, min_dist: float = 1.7,
                  basis: str = "def2-tzvp", spherical: bool = True) -> Molecule:
    """(H2O)_n on a cubic lattice: centres ``spacing`` Angstrom apart, each
    jittered by U(-jitter, jitter) per axis, monomers randomly rotated and
    resampled until no intermolecular atom pair is closer than ``min_dist``.
    """
    rng = np.random.default_rng(seed)
    mono = _water_monomer()
    placed: list[np.ndarray] = []
    for ix in range(shape[0]):
        for iy in range(shape[1]):
            for iz in range(shape[2]):
                base = spacing * np.array([ix, iy, iz], dtype=np.float64)
                for _ in range(10000):
                    centre = base + rng.uniform(-jitter, jitter, size=3)
                    xyz = centre + mono @ _random_rotation(rng).T
                    if not placed:
                        break
                    other = np.concatenate(placed)
                    d = np.linalg.norm(other[:, None, :] - xyz[None, :, :], axis=-1)
                    if d.min() >= min_dist:
                        break
                else:
                    raise RuntimeError("water_cluster: could not place a monomer")
                placed.append(xyz)
    xyz = np.concatenate(placed) * ANGSTROM_TO_BOHR
    symbols = ["O", "H", "H"] * len(placed)
    return Molecule(symbols, xyz, basis, spherical=spherical)


_BOND = {("C", "C"): 1.53, ("C", "N"): 1.47, ("C", "O"): 1.43, ("N", "N"): 1.45,
         ("N", "O"): 1.40, ("O", "O"): 1.48}
_VALENCE = {"C": 4, "N": 3, "O": 2}


def _bond_length(a: str, b: str) -> float:
    if "H" in (a, b):
        return 1.09
    return _BOND.get((a, b), _BOND.get((b, a)))


def tree_molecule(n_heavy: int, rng, *, basis: str = "def2-svp",
                  spherical: bool = True, min_nonbonded: float = 1.0,
                  max_tries: int = 2000) -> Molecule:
    """A saturated acyclic molecule with ``n_heavy`` C/N/O atoms (6:1:1).

    Heavy atoms form a random tree respecting valence; each new atom is placed
    at the bond length from its parent along a tetrahedral direction with a
    random torsion; hydrogens fill the remaining valence. The molecule is
    resampled when any nonbonded pair is closer than ``min_nonbonded`` Angstrom.
    """
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]],
                   dtype=np.float64) / np.sqrt(3.0)
    for _ in range(max_tries):
        elems = list(rng.choice(["C", "N", "O"], size=n_heavy, p=[0.75, 0.125, 0.125]))
        if n_heavy > 1 and all(e == "O" for e in elems):
            continue
        frames, pos, bonds, used = [], [], [], []
        ok = True
        for i, e in enumerate(elems):
            if i == 0:
                pos.append(np.zeros(3))
                frames.append(_random_rotation(rng))
                used.append(0)
                continue
            cands = [j for j in range(i) if used[j] < _VALENCE[elems[j]] - (0 if j == 0 else 1)]
            if not cands:
                ok = False
                break
            j = int(rng.choice(cands))
            slot = used[j] + (0 if j == 0 else 1)
            used[j] += 1
            torsion = _random_rotation(rng) if slot == 0 else np.eye(3)
            d = frames[j] @ tet[slot]
            pos.append(pos[j] + _bond_length(elems[j], e) * d)
            # child frame: first tetrahedral direction points back to the parent
            back = -d
            rot = _align(tet[0], back)
            spin = _axis_rotation(back, rng.uniform(0, 2 * np.pi))
            frames.append(spin @ rot)
            bonds.append((j, i))
            used.append(0)
        if not ok:
            continue
        symbols = list(elems)
        xyz = list(pos)
        for i, e in enumerate(elems):
            first = 0 if i == 0 else 1
            for slot in range(first + used[i], _VALENCE[e]):
                d = frames[i] @ tet[slot]
                xyz.append(pos[i] + 1.09 * d)
                symbols.append("H")
                bonds.append((i, len(xyz) - 1))
        xyz = np.array(xyz)
        bonded = set(bonds) | {(b, a) for a, b in bonds}
        d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
        n = len(xyz)
        bad = False
        for a in range(n):
            for b in range(a + 1, n):
                if (a, b) in bonded:
                    continue
                if d[a, b] < min_nonbonded:
                    bad = True
                    break
            if bad:
                break
        if bad:
            continue
        return Molecule(symbols, xyz * ANGSTROM_TO_BOHR, basis, spherical=spherical)
    raise RuntimeError("tree_molecule: no valid structure found")


def _align(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Rotation taking unit vector u onto unit vector v."""
    u = u / np.linalg.norm(u)
    v = v / np.linalg.norm(v)
    c = float(u @ v)
    w = np.cross(u, v)
    s = np.linalg.norm(w)
    if s < 1e-12:
        if c > 0:
            return np.eye(3)
        perp = np.cross(u, [1.0, 0, 0])
        if np.linalg.norm(perp) < 1e-6:
            perp = np.cross(u, [0, 1.0, 0])
        return _axis_rotation(perp / np.linalg.norm(perp), np.pi)
    k = w / s
    return _axis_rotation(k, np.arctan2(s, c))


def _axis_rotation(k: np.ndarray, angle: float) -> np.ndarray:
    k = np.asarray(k, dtype=np.float64)
    k = k / np.linalg.norm(k)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def serving_stream(n: int = 512, *, seed: int = 0, min_heavy: int = 2,
                   max_heavy: int = 24, basis: str = "def2-svp"):
    """The serving workload: ``n`` tree molecules with 2-24 heavy atoms."""
    rng = np.random.default_rng(seed)
    mols = []
    for _ in range(n):
        k = int(rng.integers(min_heavy, max_heavy + 1))
        mols.append(tree_molecule(k, rng, basis=basis))
    return mols
