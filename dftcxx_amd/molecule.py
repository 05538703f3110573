"""Molecules, basis sets and the synthetic workloads.

A :class:`Molecule` holds atoms (Z, position in bohr) and a basis-set name and
flattens the basis into the shell arrays the C ABI takes (``dg_basis`` in
include/dftgrid.h).  The CGF order is the reference's (src/molecule.cpp:135-238):
elements in ascending Z (the order of the basis file), then atoms of that
element in input order, then the shells of the element in file order; a P
shell expands to px, py, pz and a D shell to xx, xy, xz, yy, yz, zz
(src/molecule.cpp:176-218, src/cgf.h:208-221).

Geometries given in angstrom are converted with the reference's constant
1.889725989 (src/molecule.cpp:65).
"""
import json
import math
import os

import numpy as np

ANGSTROM_TO_BOHR = 1.889725989
ELEMENTS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg",
            "Al", "Si", "P", "S", "Cl", "Ar"]
SHELL_L = {"S": 0, "P": 1, "D": 2}
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis")
_BASIS_CACHE = {}


def load_basis(name):
    """Return {Z: [(L, alphas, coefs), ...]} for a basis set shipped in data/basis."""
    if name not in _BASIS_CACHE:
        path = os.path.join(_DATA, name + ".json")
        if not os.path.exists(path):
            raise FileNotFoundError(f"unknown basis set '{name}' (no {path})")
        with open(path) as f:
            doc = json.load(f)
        table = {}
        for z, shells in doc["elements"].items():
            table[int(z)] = [(SHELL_L[k], [p[0] for p in prims], [p[1] for p in prims])
                             for k, prims in shells]
        _BASIS_CACHE[name] = table
    return _BASIS_CACHE[name]


def element_number(symbol):
    """Atomic number of an element symbol, H..Ar (src/molecule.cpp:277-288)."""
    try:
        return ELEMENTS.index(symbol) + 1
    except ValueError:
        raise ValueError(f"Unknown element: {symbol}") from None


class Molecule:
    """Atoms in bohr plus a basis-set name."""

    def __init__(self, Z, xyz_bohr, basis, name=""):
        self.Z = np.asarray(Z, dtype=np.int32)
        self.xyz = np.ascontiguousarray(np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3))
        if len(self.Z) != len(self.xyz) or len(self.Z) == 0:
            raise ValueError("Z and xyz must describe the same, non-empty set of atoms")
        if self.Z.min() < 1 or self.Z.max() > 18:
            raise ValueError("only elements H..Ar are supported")
        self.basis = basis
        self.name = name

    @classmethod
    def from_angstrom(cls, symbols_or_Z, xyz_ang, basis, name=""):
        Z = [s if isinstance(s, (int, np.integer)) else element_number(s) for s in symbols_or_Z]
        return cls(Z, np.asarray(xyz_ang, dtype=np.float64) * ANGSTROM_TO_BOHR, basis, name)

    @property
    def natoms(self):
        return len(self.Z)

    @property
    def nelec(self):
        return int(self.Z.sum())

    def shells(self):
        """Flatten the basis into shell arrays in the reference CGF order.

        ``basis`` is a shipped basis-set name or a table {Z: [(L, alphas,
        coefs), ...]} in the same form as :func:`load_basis` returns.

        Returns a dict with ``atom`` (int32), ``L`` (int32), ``nprim`` (int32),
        ``offset`` (int32, into the primitive arrays), ``alpha`` and ``coef``
        (float64), and ``nbf``.
        """
        table = self.basis if isinstance(self.basis, dict) else load_basis(self.basis)
        atom, L, nprim, offset, alpha, coef = [], [], [], [], [], []
        for z in sorted(set(int(v) for v in self.Z)):
            if z not in table:
                raise ValueError(f"basis '{self.basis}' has no entry for Z={z}")
            for ia in range(self.natoms):
                if int(self.Z[ia]) != z:
                    continue
                for l, a, c in table[z]:
                    atom.append(ia)
                    L.append(l)
                    nprim.append(len(a))
                    offset.append(len(alpha))
                    alpha.extend(a)
                    coef.extend(c)
        nbf = sum((l + 1) * (l + 2) // 2 for l in L)
        return {"atom": np.asarray(atom, np.int32), "L": np.asarray(L, np.int32),
                "nprim": np.asarray(nprim, np.int32), "offset": np.asarray(offset, np.int32),
                "alpha": np.asarray(alpha, np.float64), "coef": np.asarray(coef, np.float64),
                "nbf": nbf}

    def nbf(self):
        return self.shells()["nbf"]

    def __repr__(self):
        return f"Molecule({self.name!r}, natoms={self.natoms}, basis={self.basis!r})"


# ----------------------------------------------------------------------------
# Synthetic workloads (geometries in angstrom, converted to bohr)
# ----------------------------------------------------------------------------

def h2o(basis="sto3g"):
    """H2O: O at the origin, O-H 0.9572 A, H-O-H 104.52 deg, in the xy plane."""
    half = math.radians(104.52 / 2.0)
    d = 0.9572
    xyz = [[0.0, 0.0, 0.0],
           [d * math.sin(half), d * math.cos(half), 0.0],
           [-d * math.sin(half), d * math.cos(half), 0.0]]
    return Molecule.from_angstrom([8, 1, 1], xyz, basis, "h2o")


def benzene(basis="p631"):
    """Benzene D6h: C at radius 1.39 A, H at 2.48 A, every 60 deg, plane z=0.

    Atom order: C0..C5 then H0..H5.
    """
    xyz, Z = [], []
    for rad, z in ((1.39, 6), (2.48, 1)):
        for k in range(6):
            t = math.radians(60.0 * k)
            xyz.append([rad * math.cos(t), rad * math.sin(t), 0.0])
            Z.append(z)
    return Molecule.from_angstrom(Z, xyz, basis, "benzene")


def alkane(n, basis="p631"):
    """All-anti linear alkane C_nH_{2n+2} (C-C 1.54 A, C-H 1.09 A, tetrahedral).

    Backbone C_i = (i L sin(t/2), (i mod 2) L cos(t/2), 0) with t = acos(-1/3).
    Each carbon carries two hydrogens out of the backbone plane along
    cos(t/2) u_out +/- sin(t/2) z, u_out being the in-plane unit vector opposite
    the bisector of its two C-C bonds (a virtual neighbour C_{-1} or C_n stands
    in at the chain ends); each terminal carbon carries one more hydrogen in
    the plane, along its virtual bond.  Atom order: C_0..C_{n-1}, then for each
    carbon its +z and -z hydrogens, then the terminal hydrogens of C_0 and
    C_{n-1}.
    """
    if n < 1:
        raise ValueError("alkane needs n >= 1")
    theta = math.acos(-1.0 / 3.0)
    L, dch = 1.54, 1.09
    s, c = math.sin(theta / 2.0), math.cos(theta / 2.0)

    def cpos(i):
        return np.array([i * L * s, (i % 2) * L * c, 0.0])

    carbons = [cpos(i) for i in range(n)]
    hyd = []
    term = []
    for i in range(n):
        prev_c, next_c = cpos(i - 1), cpos(i + 1)
        b1 = (prev_c - carbons[i]) / L
        b2 = (next_c - carbons[i]) / L
        bis = b1 + b2
        u_out = -bis / np.linalg.norm(bis)
        for sign in (1.0, -1.0):
            hyd.append(carbons[i] + dch * (c * u_out + sign * s * np.array([0.0, 0.0, 1.0])))
        if i == 0:
            term.append(carbons[i] + dch * b1)
        if i == n - 1:
            term.append(carbons[i] + dch * b2)
    xyz = carbons + hyd + term
    Z = [6] * n + [1] * (len(hyd) + len(term))
    return Molecule.from_angstrom(Z, np.array(xyz), basis, f"C{n}H{2 * n + 2}")


def ch4(basis="sto3g"):
    """Methane: C at the origin, C-H 1.09 A along (1,1,1), (1,-1,-1), (-1,1,-1), (-1,-1,1)."""
    d = 1.09 / math.sqrt(3.0)
    xyz = [[0, 0, 0], [d, d, d], [d, -d, -d], [-d, d, -d], [-d, -d, d]]
    return Molecule.from_angstrom([6, 1, 1, 1, 1], xyz, basis, "ch4")


def n2(basis="sto3g"):
    return Molecule.from_angstrom([7, 7], [[0, 0, -0.55], [0, 0, 0.55]], basis, "n2")


def co(basis="sto3g"):
    return Molecule.from_angstrom([6, 8], [[0, 0, 0], [0, 0, 1.13]], basis, "co")


def ethane(basis="sto3g"):
    """Staggered ethane: C-C 1.54 A along z, C-H 1.09 A, tetrahedral angles."""
    zc = 0.77
    rho = 1.09 * 2.0 * math.sqrt(2.0) / 3.0
    dz = 1.09 / 3.0
    xyz = [[0, 0, zc], [0, 0, -zc]]
    for k in range(3):
        t = math.radians(120.0 * k)
        xyz.append([rho * math.cos(t), rho * math.sin(t), zc + dz])
    for k in range(3):
        t = math.radians(120.0 * k + 60.0)
        xyz.append([rho * math.cos(t), rho * math.sin(t), -zc - dz])
    return Molecule.from_angstrom([6, 6, 1, 1, 1, 1, 1, 1], xyz, basis, "ethane")


class SplitMix64:
    """SplitMix64 generator (Steele, Lea & Flood 2014) with Box-Muller normals."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.state = seed & self.MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self):
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self):
        """One standard normal: Box-Muller with u1 in (0, 1], cosine branch."""
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


BATCH_SPECIES = (h2o, ch4, n2, co, ethane)


def batch(count=256, seed=20261202, sigma=0.05):
    """Seeded batch of perturbed small molecules.

    Per molecule, in this draw order: species ~ U{H2O, CH4, N2, CO, ethane};
    basis = STO-3G if u < 1/2 else 6-31G; then for each atom x, y, z each
    displaced by N(0, sigma) angstrom.
    """
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        make = BATCH_SPECIES[min(int(rng.uniform() * len(BATCH_SPECIES)), len(BATCH_SPECIES) - 1)]
        basis = "sto3g" if rng.uniform() < 0.5 else "p631"
        m = make(basis)
        disp = np.array([[rng.normal() for _ in range(3)] for _ in range(m.natoms)])
        m.xyz = m.xyz + sigma * ANGSTROM_TO_BOHR * disp
        out.append(m)
    return out


WORKLOADS = {
    "h2o_sto3g": (lambda: h2o("sto3g"), 50, 110),
    "benzene_631g": (lambda: benzene("p631"), 75, 302),
    "c10h22_631g": (lambda: alkane(10, "p631"), 75, 302),
    "c30h62_631g": (lambda: alkane(30, "p631"), 75, 302),
}
