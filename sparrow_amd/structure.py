"""Atom collections: element symbols, XYZ input, Angstrom -> bohr.

The reference reads XYZ files through Utils::XyzStreamHandler (Angstrom in the file, bohr
inside), with CODATA 2014 bohr_per_angstrom = 1/0.52917721067 (pinned by
Tests/AppTests/molden_input_reference.molden:4-31 against Tests/AppTests/sft_1.xyz).
"""
import numpy as np

ANGSTROM_PER_BOHR = 0.52917721067
BOHR_PER_ANGSTROM = 1.0 / ANGSTROM_PER_BOHR

SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn "
           "Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce "
           "Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn").split()
Z_OF = {s.lower(): i + 1 for i, s in enumerate(SYMBOLS)}


def atomic_number(sym):
    return Z_OF[sym.strip().lower()]


class AtomCollection:
    """Elements (atomic numbers) and positions in bohr, as Utils::AtomCollection."""

    def __init__(self, Z, positions_bohr):
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        self.positions = np.ascontiguousarray(positions_bohr, dtype=np.float64).reshape(-1, 3)
        if len(self.Z) != len(self.positions):
            raise ValueError("element and position counts differ")

    def __len__(self):
        return len(self.Z)

    @classmethod
    def from_angstrom(cls, symbols_or_Z, xyz_angstrom):
        Z = [s if isinstance(s, (int, np.integer)) else atomic_number(s) for s in symbols_or_Z]
        return cls(Z, np.asarray(xyz_angstrom, dtype=np.float64) * BOHR_PER_ANGSTROM)

    @classmethod
    def from_xyz_string(cls, text):
        lines = text.strip("\n").splitlines()
        n = int(lines[0].split()[0])
        syms, xyz = [], []
        for line in lines[2:2 + n]:
            t = line.split()
            syms.append(t[0])
            xyz.append([float(v) for v in t[1:4]])
        return cls.from_angstrom(syms, xyz)

    @classmethod
    def from_xyz_file(cls, path):
        with open(path) as f:
            return cls.from_xyz_string(f.read())

    def copy(self):
        return AtomCollection(self.Z.copy(), self.positions.copy())
