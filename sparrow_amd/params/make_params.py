#!/usr/bin/env python3
"""Convert the reference's NDDO parameter sets into this package's flat text tables.

Reads, for each of MNDO, AM1, PM3, RM1 and PM6:
  - src/Sparrow/Sparrow/Resources/<Method>/parameters.json   (schema: Parameters.cpp:22-62,76-86)
  - src/Sparrow/Sparrow/Resources/<Method>/<METHOD>-STO-6G.basis (Turbomole format; the
    coefficients are the primitive-normalised contraction coefficients, i.e. the
    `normalizedCoefficient` field that GTOOverlapMatrixBlock.cpp:69-120 consumes, printed
    at 15 significant digits by Tests/Nddo/Sto6gBasisTest.cpp:33-48)
and writes <method>.txt next to this script.  The numbers are the published parameter values
in their original units (eV, Angstrom, Angstrom^-1); all unit conversion happens in the
library when the table is loaded.

Output format (whitespace separated, one record per line, '#' starts a comment):
  method <NAME> basis <sp|spd> diatomic <0|1>
  atom Z Uss Upp Udd betaS betaP betaD zetaS zetaP zetaD zetaSint zetaPint zetaDint
       gss gpp gsp gp2 hsp pcore f0sd g2sd alpha nGauss [a b c]*nGauss
  pair Z1 Z2 exponent factor
  gto Z l nPrim exp1 coef1 ... expN coefN
This script needs the reference checkout and is only run by hand when the tables are
regenerated; the library never reads the reference.
"""
import json
import os
import sys

REF = "/root/reference/src/Sparrow/Sparrow/Resources"
METHODS = [
    # name, resource dir, basis file stem, basis functions, diatomic parameters
    ("MNDO", "Mndo", "MNDO", "sp", 0),
    ("AM1", "Am1", "AM1", "sp", 0),
    ("PM3", "Pm3", "PM3", "sp", 0),
    ("RM1", "Rm1", "RM1", "sp", 0),
    ("PM6", "Pm6", "PM6", "spd", 1),
]

SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn "
           "Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce "
           "Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn").split()
Z_OF = {s.lower(): i + 1 for i, s in enumerate(SYMBOLS)}


def read_basis(path):
    """Parse a Turbomole $basis block into {Z: [(l, [(exp, coef), ...]), ...]}."""
    out = {}
    lines = open(path).read().splitlines()
    i = 0
    cur = None
    while i < len(lines):
        tok = lines[i].split()
        if not tok or tok[0] in ("$basis", "$end"):
            i += 1
            continue
        if tok[0] == "*":
            i += 1
            continue
        if tok[0].lower() in Z_OF and len(tok) == 2:
            cur = Z_OF[tok[0].lower()]
            out[cur] = []
            i += 1
            continue
        if len(tok) == 2 and tok[1] in ("s", "p", "d"):
            n = int(tok[0])
            l = "spd".index(tok[1])
            prims = []
            for k in range(n):
                e, c = lines[i + 1 + k].split()
                prims.append((float(e), float(c)))
            out[cur].append((l, prims))
            i += 1 + n
            continue
        raise ValueError(f"unexpected line in {path}: {lines[i]!r}")
    return out


def main(outdir):
    for name, sub, stem, basis, diat in METHODS:
        par = json.load(open(os.path.join(REF, sub, "parameters.json")))
        gto = read_basis(os.path.join(REF, sub, f"{stem}-STO-6G.basis"))
        rows = [f"# {name} parameters converted by make_params.py (eV / Angstrom units)",
                f"method {name} basis {basis} diatomic {diat}"]
        for a in sorted(par["atomic"], key=lambda a: a["key"]):
            z = a["key"]
            p = a["value"]["pack"]
            g = a["value"].get("gaussianRepulsion", [])
            vals = [p["oneCenterEnergy"]["s"], p["oneCenterEnergy"]["p"], p["oneCenterEnergy"]["d"],
                    p["beta"]["s"], p["beta"]["p"], p["beta"]["d"],
                    p["orbitalExponent"]["s"], p["orbitalExponent"]["p"], p["orbitalExponent"]["d"],
                    p["internalExponent"]["s"], p["internalExponent"]["p"], p["internalExponent"]["d"],
                    p["gss"], p["gpp"], p["gsp"], p["gp2"], p["hsp"], p["pcore"], p["f0sd"], p["g2sd"],
                    p["alpha"]]
            gv = []
            for t in g:
                gv += [t["a"], t["b"], t["c"]]
            rows.append("atom %d %s %d %s" % (z, " ".join(repr(float(v)) for v in vals), len(g),
                                              " ".join(repr(float(v)) for v in gv)))
        for d in par.get("diatomic", []):
            k, v = d["key"], d["value"]
            rows.append("pair %d %d %r %r" % (k["first"], k["second"], float(v["exponent"]), float(v["factor"])))
        for z in sorted(gto):
            for l, prims in gto[z]:
                rows.append("gto %d %d %d %s" % (z, l, len(prims),
                                                 " ".join("%r %r" % pc for pc in prims)))
        with open(os.path.join(outdir, f"{name.lower()}.txt"), "w") as f:
            f.write("\n".join(rows) + "\n")
        print(f"{name}: {len(par['atomic'])} elements, {len(par.get('diatomic', []))} pairs, "
              f"{len(gto)} STO-6G sets")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.dirname(os.path.abspath(__file__)))
