"""Molden files of a converged restricted calculation, in the layout the reference writes
(src/Sparrow/Sparrow/Implementations/MoldenFileGenerator.cpp:44-165): atoms in bohr, the
method's STO-6G expansions ([GTO]; no d shells for MNDO, AM1, PM3 and RM1), spherical d
([5D]) and one [MO] block per orbital, ascending in energy, with its coefficients in the
Molden order of the atom's functions (s, px, py, pz, dz2, dxz, dyz, dx2-y2, dxy).

Usage: write(path, method, Z, xyz_bohr, eps, C, nocc) or text(...) for the string; eps and C
come from Context.scf(..., want_orbitals=True) (C column k = orbital k, AO order per atom
s, px, py, pz, dx2-y2, dxz, dz2, dyz, dxy).  An unrestricted calculation passes eps and C as
{"alpha": ..., "beta": ...} and nocc as (n_alpha, n_beta): an Alpha block, then a Beta block,
occupations 1.0 (MoldenFileGenerator.cpp:133-148).
"""
import os

import numpy as np

from .params import PARAM_DIR
from .structure import SYMBOLS

# Molden position of each of an atom's functions -> its index in the AO order of the library
# (MoldenFileGenerator.cpp:37-42)
MOLDEN_TO_AO = (0, 1, 2, 3, 6, 5, 7, 4, 8)
SMALL_BASIS = ("MNDO", "AM1", "PM3", "RM1")  # s/p valence only: d expansions are not written


def _gto_records(method):
    """{Z: [(l, [(exponent, coefficient)] * 6)]} from the method's parameter table."""
    out = {}
    with open(os.path.join(PARAM_DIR, method.lower() + ".txt")) as f:
        for line in f:
            t = line.split()
            if not t or t[0] != "gto":
                continue
            Z, l, npr = int(t[1]), int(t[2]), int(t[3])
            v = [float(x) for x in t[4:4 + 2 * npr]]
            out.setdefault(Z, []).append((l, list(zip(v[0::2], v[1::2]))))
    return out


def _f(x, w=18):
    return f"{x:{w}.10f}"


def text(method, Z, xyz_bohr, eps, C, nocc):
    method = method.upper()
    Z = [int(z) for z in Z]
    xyz = np.asarray(xyz_bohr, dtype=float).reshape(-1, 3)
    gtos = _gto_records(method)
    small = method in SMALL_BASIS
    lines = ["[Molden Format]", "Written by Sparrow, the semiempirical library of the SCINE software package",
             "[Atoms] (AU)"]
    for a, z in enumerate(Z):
        sym = SYMBOLS[z - 1]
        lines.append(f"{sym:<5}{a + 1:<10}{z:<10} {_f(xyz[a, 0])} {_f(xyz[a, 1])} {_f(xyz[a, 2])}")
    lines.append("[GTO]")
    naos = []
    for a, z in enumerate(Z):
        lines.append(f"{a + 1} 0")
        nao = 0
        for l, prims in gtos[z]:
            if l == 2 and small:
                continue
            lines.append(f"{'spd'[l]:<5}{len(prims):<5}{'1.00':<5}")
            for e, c in prims:
                lines.append(f"{_f(e)} {_f(c)}")
            nao += 2 * l + 1
        naos.append(nao)
        lines.append("")
    n = sum(naos)
    if isinstance(eps, dict):  # unrestricted: Alpha then Beta, singly occupied
        blocks = [(eps["alpha"], C["alpha"], nocc[0], "Alpha", "1.0"), (eps["beta"], C["beta"], nocc[1], "Beta", "1.0")]
    else:
        blocks = [(eps, C, nocc, "Alpha", "2.0")]
    lines += ["[5D]", "[7F]", "[9G]", "[MO]"]
    first = np.concatenate([[0], np.cumsum(naos)[:-1]]).astype(int)
    for e, c, no, spin, occ in blocks:
        e = np.asarray(e, dtype=float)
        c = np.asarray(c, dtype=float)
        if c.shape != (n, n) or e.shape != (n,):
            raise ValueError(f"orbitals must be {n} x {n} with {n} energies for this structure and method")
        for k in range(n):
            lines += [f"{'Sym=':<10}{'A1':<10}", f"{'Ene=':<10}{e[k]:<10.10f}", f"{'Spin=':<10}{spin:<10}",
                      f"{'Occup=':<10}{(occ if k < no else '0.0'):<10}"]
            for a in range(len(Z)):
                for m in range(naos[a]):
                    lines.append(f"{first[a] + m + 1:>4}{c[first[a] + MOLDEN_TO_AO[m], k]:18.10f}")
    return "\n".join(lines) + "\n"


def write(path, method, Z, xyz_bohr, eps, C, nocc):
    with open(path, "w") as f:
        f.write(text(method, Z, xyz_bohr, eps, C, nocc))
