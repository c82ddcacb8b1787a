"""Molecules, basis assembly and the seeded synthetic systems.

Semantics follow the reference's chemistry model:
  * coordinates are bohr; Angstrom input is scaled by 1.889726124565
    (src/logger.cpp:31, src/molecule.cpp:20-78);
  * the orbital basis is built atom by atom in geometry order, shells in
    basis-file order (src/molecule.cpp:655-672, src/basisreader.cpp:46-132);
  * E_nuc = sum_{i<j} Z_i Z_j / r_ij (src/molecule.cpp:443-457).

Shells are handed over with RAW contraction coefficients; the native library
normalises them the way libint2::Shell does (the reference passes raw
coefficients to libint2 in src/basis.cpp:33-66).
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

TOBOHR = 1.889726124565  # src/logger.cpp:31
SYMBOL = {1: "H", 2: "He", 6: "C", 7: "N", 8: "O"}
ZNUM = {v: k for k, v in SYMBOL.items()}
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
_BASIS_CACHE: dict = {}


def load_basis(name: str) -> dict:
    key = name.lower()
    if key not in _BASIS_CACHE:
        path = os.path.join(_DATA, key + ".json")
        if not os.path.exists(path):
            # mirrors the reference's BASISLIB error (src/basisreader.cpp:130)
            raise KeyError("BASISLIB: could not read basis " + name)
        with open(path) as fh:
            _BASIS_CACHE[key] = json.load(fh)["elements"]
    return _BASIS_CACHE[key]


@dataclass
class Shell:
    l: int
    exps: list
    coefs: list          # raw (unnormalised) contraction coefficients
    center: tuple        # bohr
    atom: int


@dataclass
class Molecule:
    Z: list                      # atomic numbers
    xyz: np.ndarray              # (natom, 3) bohr
    charge: int = 0
    multiplicity: int = 1
    name: str = "molecule"

    @property
    def natom(self) -> int:
        return len(self.Z)

    @property
    def nel(self) -> int:
        return int(sum(self.Z)) - self.charge

    def enuc(self) -> float:
        e = 0.0
        for i in range(self.natom):
            for j in range(i + 1, self.natom):
                e += self.Z[i] * self.Z[j] / float(np.linalg.norm(self.xyz[i] - self.xyz[j]))
        return e

    def shells(self, basis: str = "cc-pvdz") -> list:
        data = load_basis(basis)
        out = []
        for a, (z, r) in enumerate(zip(self.Z, self.xyz)):
            sym = SYMBOL[z]
            if sym not in data:
                raise KeyError("BASISLIB: could not read basis %s for atom %s" % (basis, sym))
            for sh in data[sym]:
                out.append(Shell(sh["l"], list(sh["exps"]), list(sh["coefs"]),
                                 (float(r[0]), float(r[1]), float(r[2])), a))
        return out

    def translated(self, d) -> "Molecule":
        return Molecule(list(self.Z), self.xyz + np.asarray(d, float)[None, :],
                        self.charge, self.multiplicity, self.name)

    def displaced(self, atom: int, axis: int, h: float) -> "Molecule":
        xyz = self.xyz.copy()
        xyz[atom, axis] += h
        return Molecule(list(self.Z), xyz, self.charge, self.multiplicity, self.name)


def nbf_of(shells, pure: bool = True) -> int:
    return sum((2 * s.l + 1) if pure else (s.l + 1) * (s.l + 2) // 2 for s in shells)


def from_angstrom(symbols, coords, name="molecule") -> Molecule:
    return Molecule([ZNUM[s] for s in symbols], np.asarray(coords, float) * TOBOHR, name=name)


# ----------------------------------------------------------------------------
# Fixed geometries
# ----------------------------------------------------------------------------

def water() -> Molecule:
    """O at the origin, H at (+-0.757, 0.586, 0) Angstrom."""
    return from_angstrom(["O", "H", "H"], [[0, 0, 0], [0.757, 0.586, 0], [-0.757, 0.586, 0]], "water")


def water_h2otest() -> Molecule:
    """Geometry of the reference's test/h2otest.mol:7-11 (bohr)."""
    return Molecule([8, 1, 1], np.array([[0.0, -0.143226, 0.0],
                                         [1.63803684, 1.1365488, 0.0],
                                         [-1.63803684, 1.1365488, 0.0]]), name="h2otest")


def benzene() -> Molecule:
    """Planar D6h benzene: C on a 1.39 A radius hexagon, H radially at 2.48 A."""
    sym, xyz = [], []
    for k in range(6):
        t = math.pi / 3 * k
        sym.append("C")
        xyz.append([1.39 * math.cos(t), 1.39 * math.sin(t), 0.0])
    for k in range(6):
        t = math.pi / 3 * k
        sym.append("H")
        xyz.append([2.48 * math.cos(t), 2.48 * math.sin(t), 0.0])
    return from_angstrom(sym, xyz, "benzene")


def benzene_dz_ref() -> Molecule:
    """Geometry of the reference's test/benzene-dz.mol:6-19 (Angstrom)."""
    c = [[0.000, 1.387, 0.0], [1.201, 0.693, 0.0], [1.201, -0.693, 0.0],
         [0.000, -1.387, 0.0], [-1.201, -0.693, 0.0], [-1.201, 0.693, 0.0]]
    h = [[0.000, 2.469, 0.0], [2.139, 1.235, 0.0], [2.139, -1.235, 0.0],
         [0.000, -2.469, 0.0], [-2.139, -1.235, 0.0], [-2.139, 1.235, 0.0]]
    return from_angstrom(["C"] * 6 + ["H"] * 6, c + h, "benzene-dz")


# ----------------------------------------------------------------------------
# Seeded generators
# ----------------------------------------------------------------------------

def _haar_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(nwater: int = 20, seed: int = 20, spacing: float = 3.0,
                  jitter: float = 0.3, min_dist: float = 1.0) -> Molecule:
    """(H2O)_n: monomers (water() geometry, centred on O) placed at the first n
    points of a 3x3x3 cubic lattice (lexicographic order, `spacing` Angstrom),
    each jittered by U(-jitter, jitter) per axis and rotated by a Haar-random
    rotation drawn from numpy.random.default_rng(seed).  If any intermolecular
    H..H or O..H distance is below `min_dist` Angstrom the seed is advanced by
    one and the draw repeated (the seed actually used is kept in .name)."""
    mono = np.array([[0, 0, 0], [0.757, 0.586, 0], [-0.757, 0.586, 0]], float)
    pts = [(i, j, k) for i in range(3) for j in range(3) for k in range(3)][:nwater]
    s = seed
    while True:
        rng = np.random.default_rng(s)
        sym, xyz, mol_of = [], [], []
        for m, (i, j, k) in enumerate(pts):
            c = np.array([i, j, k], float) * spacing + rng.uniform(-jitter, jitter, 3)
            R = _haar_rotation(rng)
            for a, p in zip("OHH", mono):
                sym.append(a)
                xyz.append(c + R @ p)
                mol_of.append(m)
        xyz = np.array(xyz)
        ok = True
        for a in range(len(sym)):
            for b in range(a + 1, len(sym)):
                if mol_of[a] == mol_of[b] or (sym[a] == "O" and sym[b] == "O"):
                    continue
                if np.linalg.norm(xyz[a] - xyz[b]) < min_dist:
                    ok = False
                    break
            if not ok:
                break
        if ok:
            return from_angstrom(sym, xyz, "(H2O)%d seed %d" % (nwater, s))
        s += 1


ATOMIC_MASS = {1: 1.00794, 6: 12.0107, 7: 14.0067, 8: 15.9994}


def fragment_info(mol: Molecule, groups, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=True, occ=None,
                  radius=None) -> list:
    """FragmentInfo (inc/fock.hpp:74-95) of each atom group (lists of atom
    indices, consecutive and in order, so each fragment's shells are one
    contiguous run of the basis): functions, shells, auxiliary functions and
    shells, occupied count (default nel/2 of the neutral group), centre of
    mass (bohr) and radius (default: the largest atom distance from it plus
    2 bohr)."""
    obs, abs_ = mol.shells(basis), mol.shells(aux)
    out, start, astart = [], 0, 0
    for g, atoms in enumerate(groups):
        atoms = list(atoms)
        sh = [s for s in obs if s.atom in atoms]
        ash = [s for s in abs_ if s.atom in atoms]
        m = np.array([ATOMIC_MASS[mol.Z[a]] for a in atoms])
        com = (m[:, None] * mol.xyz[atoms]).sum(axis=0) / m.sum()
        rad = max(float(np.linalg.norm(mol.xyz[a] - com)) for a in atoms) + 2.0 if radius is None else radius
        nb, na = nbf_of(sh, pure), nbf_of(ash, pure)
        nel = int(sum(mol.Z[a] for a in atoms))
        out.append(dict(occ=(nel // 2 if occ is None else occ[g]), nbfs=nb, naux=na, start=start, auxstart=astart,
                        nshells=len(sh), ndfshells=len(ash), radius=rad, com=com))
        start += nb
        astart += na
    return out


def random_chain(natom: int = 12, seed: int = 12) -> Molecule:
    """Bonded random C/H/N/O chain: bond lengths U(1.0, 1.5) A, bond angles
    U(100, 125) deg and torsions U(-180, 180) deg from default_rng(seed);
    rejected (seed advanced) if any non-bonded pair is closer than 0.9 A; the
    last atom is redrawn until the electron count is even."""
    s = seed
    while True:
        rng = np.random.default_rng(s)
        syms = list(rng.choice(["H", "C", "N", "O"], size=natom))
        tries = 0
        while sum(ZNUM[x] for x in syms) % 2 and tries < 50:
            syms[-1] = str(rng.choice(["H", "C", "N", "O"]))
            tries += 1
        xyz = [np.zeros(3)]
        d = rng.uniform(1.0, 1.5)
        xyz.append(np.array([d, 0, 0]))
        for a in range(2, natom):
            d = rng.uniform(1.0, 1.5)
            ang = math.radians(rng.uniform(100, 125))
            tor = math.radians(rng.uniform(-180, 180))
            p1 = xyz[a - 1]
            p2 = xyz[a - 2]
            p3 = xyz[a - 3] if a >= 3 else p2 + np.array([0, 1.0, 0])
            b1 = p1 - p2
            b1 /= np.linalg.norm(b1)
            n = np.cross(p2 - p3, b1)
            if np.linalg.norm(n) < 1e-8:
                n = np.cross(b1, [0, 0, 1.0])
            n /= np.linalg.norm(n)
            m = np.cross(n, b1)
            dvec = np.array([-d * math.cos(ang), d * math.sin(ang) * math.cos(tor),
                             d * math.sin(ang) * math.sin(tor)])
            xyz.append(p1 + dvec[0] * b1 + dvec[1] * m + dvec[2] * n)
        xyz = np.array(xyz)
        ok = sum(ZNUM[x] for x in syms) % 2 == 0
        for a in range(natom):
            for b in range(a + 2, natom):
                if np.linalg.norm(xyz[a] - xyz[b]) < 0.9:
                    ok = False
        if ok:
            return from_angstrom(syms, xyz, "chain%d seed %d" % (natom, s))
        s += 1


def random_molecule(rng, natom: int) -> Molecule:
    """Random H/C/N/O cluster of `natom` atoms, positions rejection-sampled in
    a cube sized for ~1.3 A neighbour spacing with a 0.9 A minimum distance;
    the last atom is resampled until the electron count is even."""
    syms = list(rng.choice(["H", "C", "N", "O"], size=natom))
    while sum(ZNUM[x] for x in syms) % 2:
        syms[-1] = str(rng.choice(["H", "C", "N", "O"]))
    box = 1.3 * max(natom, 2) ** (1.0 / 3.0) + 0.5
    pts = []
    while len(pts) < natom:
        p = rng.uniform(0, box, 3)
        if all(np.linalg.norm(p - q) >= 0.9 for q in pts):
            pts.append(p)
    return from_angstrom(syms, np.array(pts), "rand%d" % natom)


def random_batch(nmol: int = 256, seed: int = 256, lo: int = 3, hi: int = 30) -> list:
    rng = np.random.default_rng(seed)
    return [random_molecule(rng, int(rng.integers(lo, hi + 1))) for _ in range(nmol)]
