#!/usr/bin/env python3
"""Extract the reference's known-answer values into tests/golden/reference_values.json.

Sources (read as text from the reference checkout; this script is run by hand, the JSON it
writes is the committed fixture):
  src/Sparrow/Tests/MethodsTests/{Pm6/PM6Test.cpp, Mndo/MNDOTest.cpp, Am1/AM1Test.cpp}
      restricted tests: XYZ input, expected one-electron matrices (isApprox tolerance),
      electronic / repulsion / total energies and eigenvalues (DoubleNear tolerance), eV units
  src/Sparrow/Tests/AppTests/test_sparrow.py
      CLI energies of sft_1 per method (assertAlmostEqual places), MOPAC energies of sft_5a/b
  src/Sparrow/Tests/AppTests/molden_input_reference.molden
      PM6 orbital energies of sft_1 (Hartree, 10 decimals); into molden_sft1_pm6.json: the file's
      header, [Atoms] and [GTO] blocks and the MO blocks of a subset of orbitals (text lines)
  src/Sparrow/Tests/AppTests/{sft_1,sft_5a,sft_5b}.xyz   input structures (data files)
  src/Sparrow/Tests/TimeDependent/CISTestCalculation.cpp
      into cis_pm3.json: the PM3 test structures (XYZ, Angstrom) and, per dense-CIS test, the
      spin block, reference (restricted / unrestricted), roots compared, expected excitation
      energies (eV) and DoubleNear tolerance
  src/Sparrow/Tests/Nddo/{MultipoleParameters,oneCenterTwoElectronIntegrals,LocalIntegral,
      VuvBMatrix,Sto6gBasis}Test.cpp  into nddo_unit.json: the literal inputs and expected values of
      the parameter-level tests (charge separations, Klopman-Ohno terms, one-centre integrals, the
      C-O and H-V local two-centre integrals, the VuvB set-up); with the PM6 and MNDO STO-6G
      records of H, C, N, O, S from Resources/*/*-STO-6G.basis, the files Sto6gBasisTest compares
      the runtime Slater-to-Gaussian expansion with
"""
import json
import os
import re
import sys

REF = "/root/reference/src/Sparrow/Tests"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_values.json")
OUT_MOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "molden_sft1_pm6.json")
OUT_CIS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cis_pm3.json")
OUT_UNIT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nddo_unit.json")
IE = "/root/reference/src/Sparrow/Sparrow/Implementations/Nddo/Utils/IntegralsEvaluationUtils"
RES = "/root/reference/src/Sparrow/Sparrow/Resources"
MOLDEN_MOS = (0, 1, 10, 20, 30, 37, 38, 39, 40, 50, 60, 69)  # orbitals kept (HOMO 38, LUMO 39)


def molden_fixture(text):
    lines = text.split("\n")
    i_atoms, i_gto, i_5d, i_mo = (lines.index(t) for t in ("[Atoms] (AU)", "[GTO]", "[5D]", "[MO]"))
    blocks, cur = [], None
    for ln in lines[i_mo + 1:]:
        if ln.startswith("Sym="):
            cur = [ln]
            blocks.append(cur)
        elif cur is not None and ln:
            cur.append(ln)
    mos = []
    for k in MOLDEN_MOS:
        b = blocks[k]
        mos.append({"index": k, "lines": b})
    return {"source": "src/Sparrow/Tests/AppTests/molden_input_reference.molden (sft_1, PM6)",
            "header": lines[:i_atoms], "atoms": lines[i_atoms:i_gto], "gto": lines[i_gto:i_5d],
            "mo_header": lines[i_5d:i_mo + 1], "n_mo": len(blocks), "mos": mos}


def literal_xyz(block):
    m = re.search(r'std::stringstream\s+\w+\s*\(((?:\s*"[^"]*")+)\s*\)', block)
    if not m:
        m = re.search(r'std::stringstream\s*\(((?:\s*"[^"]*")+)\s*\)', block)
    if not m:
        return None
    parts = re.findall(r'"([^"]*)"', m.group(1))
    return "".join(parts).replace("\\n", "\n")


def numbers(text):
    return [float(x) for x in re.findall(r"-?\d+\.?\d*(?:e[-+]?\d+)?", text)]


def parse_method_tests(path, method):
    src = open(path).read()
    out = []
    starts = [m.start() for m in re.finditer(r"TEST_F\(", src)] + [len(src)]
    for a, b in zip(starts[:-1], starts[1:]):
        block = src[a:b]
        name = re.match(r"TEST_F\(\w+,\s*(\w+)\)", block).group(1)
        line = src[:a].count("\n") + 1
        if name.startswith("DISABLED") or "Unrestricted" in block or "setDensityMatrix" in block or "setSpinMultiplicity(2)" in block:
            continue
        xyz = literal_xyz(block)
        if xyz is None:
            continue
        rec = {"name": name, "method": method, "source": f"{os.path.relpath(path, REF)}:{line}", "xyz": xyz}
        mc = re.search(r"setMolecularCharge\((-?\d+)\)", block)
        rec["charge"] = int(mc.group(1)) if mc else 0
        mm = re.search(r"setScfMixer\(Utils::scf_mixer_t::(\w+)\)", block)
        rec["mixer"] = mm.group(1) if mm else "default"
        for kind in ("ElectronicEnergy", "RepulsionEnergy"):
            m = re.search(r"get%s\(\)\s*\*\s*(?:Utils::Constants::)?ev_per_hartree,\s*DoubleNear\(([-\d.e]+),\s*([\d.e-]+)\)" % kind,
                          block)
            if m:
                rec[kind] = [float(m.group(1)), float(m.group(2))]
        m = re.search(r"getEnergy\(\)\s*\*\s*(?:Utils::Constants::)?ev_per_hartree,\s*DoubleNear\(([-\d.e]+),\s*([\d.e-]+)\)", block)
        if m:
            rec["Energy"] = [float(m.group(1)), float(m.group(2))]
        m = re.search(r"std::vector<double> expected = \{([^}]*)\};.*?DoubleNear\(expected\[i\],\s*([\d.e-]+)\)", block, re.S)
        if m:
            rec["Eigenvalues"] = [numbers(m.group(1)), float(m.group(2))]
        m = re.search(r"Eigen::Matrix<double,\s*(\d+),\s*(\d+)>\s*expected;\s*expected\s*<<(.*?);.*?isApprox\(expected\s*/\s*ev_per_hartree,\s*([\d.e-]+)\)",
                      block, re.S)
        if not m:
            m = re.search(r"Eigen::Matrix<double,\s*(\d+),\s*(\d+)>\s*expected;\s*expected\s*<<(.*?);.*?isApprox\(expected\s*/\s*Utils::Constants::ev_per_hartree,\s*([\d.e-]+)\)",
                          block, re.S)
        if m:
            r, c = int(m.group(1)), int(m.group(2))
            vals = numbers(m.group(3))
            assert len(vals) == r * c, (name, len(vals), r, c)
            rec["OneElectronMatrix"] = [vals, r, c, float(m.group(4))]
        if len(rec) > 6:
            out.append(rec)
    return out


def cis_fixture(path):
    src = open(path).read()
    setup = src[src.index("void SetUp()"):]
    streams = {}
    for m in re.finditer(r'std::stringstream\s+(\w+)\s*\(((?:\s*"[^"]*")+)\s*\)', setup):
        streams[m.group(1)] = "".join(re.findall(r'"([^"]*)"', m.group(2))).replace("\\n", "\n")
    calc_stream = {"calculator": "ss"}
    for m in re.finditer(r"(calculator\w+)->setStructure\(Utils::XyzStreamHandler::read\((\w+)\)\)", setup):
        calc_stream[m.group(1)] = m.group(2)
    for m in re.finditer(r"auto (\w+) = Utils::XyzStreamHandler::read\((\w+)\);\s*(calculator\w+)->setStructure\(\1\)", setup):
        calc_stream[m.group(3)] = m.group(2)
    tests = []
    for m in re.finditer(r"TEST_F\(ACISTestCalculation, (CanPerformDenseCalculation\w*)\) \{(.*?)\n\}", src, re.S):
        name, body = m.group(1), m.group(2)
        calc = re.search(r"setReferenceCalculator\((calculator\w*)->clone\(\)\)", body).group(1)
        unrestricted = "SpinMode::Unrestricted" in body
        block = re.search(r'spinBlock, (?:"(\w+)"|Utils::SettingsNames::SpinBlocks::(\w+))\)', body)
        spin = "unrestricted" if unrestricted else (block.group(1) or block.group(2))
        nroots = int(re.search(r"numberOfEigenstates, (\d+)\)", body).group(1))
        vals = numbers(re.search(r"Energies << ([^;]*);", body).group(1))
        tol = float(re.search(r"DoubleNear\(\w+\(i\), ([\d.e-]+)\)", body).group(1))
        bound = re.search(r"for \(int i = 0; i < (\d+); \+\+i\)", body)
        ncmp = int(bound.group(1)) if bound else nroots
        line = src[:m.start()].count("\n") + 1
        tests.append({"name": name, "structure": calc_stream[calc], "spin": spin, "roots": ncmp,
                      "energies_ev": vals[:ncmp], "tolerance_ev": tol,
                      "source": f"TimeDependent/CISTestCalculation.cpp:{line}"})
    return {"method": "PM3", "source": "src/Sparrow/Tests/TimeDependent/CISTestCalculation.cpp (ORCA 4.1.0 CIS values)",
            "structures": streams, "tests": tests}


def arith(expr):
    """Value of a literal arithmetic expression of the test files (numbers, + - * /, parentheses,
    sqrt(), SQRT2); anything else is rejected."""
    e = expr.replace("Utils::Constants::", "").replace("SQRT2", "sqrt(2)").strip()
    if not re.fullmatch(r"[0-9eE.+\-*/() ]*(sqrt\([0-9.]+\)[0-9eE.+\-*/() ]*)*", e):
        raise ValueError(f"not a literal expression: {expr!r}")
    e = e.replace("sqrt(", "(lambda v: v ** 0.5)(")
    return float(eval(e, {"__builtins__": {}}, {}))


def enum_order(text, name):
    body = re.search(r"enum class %s[^{]*\{(.*?)\};" % name, text, re.S).group(1)
    body = re.sub(r"//[^\n]*", "", body)
    return [t.strip() for t in body.split(",") if t.strip()]


def unit_fixture():
    src = os.path.join(REF, "Nddo")
    pairs = enum_order(open(os.path.join(IE, "GeneralTypes.h")).read(), "twoElIntegral_t")
    mp = enum_order(open(os.path.join(IE, "multipoleTypes.h")).read(), "MultipolePair")
    out = {"twoElIntegral_t": pairs, "MultipolePair": mp}
    # --- MultipoleParametersTest: charge separations and Klopman-Ohno terms
    t = open(os.path.join(src, "MultipoleParametersTest.cpp")).read()
    cs, kl = [], []
    for m in re.finditer(r"TEST_F\(\w+, (\w+)\) \{(.*?)\n\}", t, re.S):
        name, body = m.group(1), m.group(2)
        line = t[:m.start()].count("\n") + 1
        c = re.search(r"d\.computeFromExponents\(([^)]*)\)", body)
        if c:
            args = [arith(a) for a in c.group(1).split(",")]
            ns, np_ = int(args[0]), int(args[1])
            nd, z = (int(args[2]), args[3:]) if len(args) == 6 else (0, args[2:] + [0.0])
            exp = {q: [arith(v), float(tol)] for q, v, tol in
                   re.findall(r"d\.get\(MultipolePair::(\w+)\), DoubleNear\(([^,]+), ([\de.-]+)\)", body)}
            cs.append({"name": name, "source": f"Tests/Nddo/MultipoleParametersTest.cpp:{line}", "ns": ns, "np": np_,
                       "nd": nd, "zeta": z, "expected": exp})
        g = re.search(r"k\.generateUpTo(S|P|D)\((.*?)\);", body, re.S)
        if g:
            vals = {v: arith(x) for v, x in re.findall(r"double (\w+) = ([^;]+);", body)}
            order = "SPD".index(g.group(1))
            args = [a.strip() for a in g.group(2).replace("\n", " ").split(",")]
            names = []
            for a in args:  # "gss * Utils::Constants::hartree_per_ev" -> ("gss", eV) or ("D1sp", bohr)
                v = a.split("*")[0].strip()
                names.append([v, "hartree_per_ev" in a])
            exp = {q: [arith(v), float(tol)] for q, v, tol in
                   re.findall(r"k\.get\(MultipolePair::(\w+)\), DoubleNear\(([^,]+), ([\de.-]+)\)", body)}
            nan = re.findall(r"isnan\(k\.get\(MultipolePair::(\w+)\)\)", body)
            kl.append({"name": name, "source": f"Tests/Nddo/MultipoleParametersTest.cpp:{line}", "order": order,
                       "arguments": [[nm, vals[nm], ev] for nm, ev in names], "expected": exp, "nan_or_inf": nan})
    out["charge_separations"], out["klopman"] = cs, kl
    # --- oneCenterTwoElectronIntegralsTest
    t = open(os.path.join(src, "oneCenterTwoElectronIntegralsTest.cpp")).read()
    oc = []
    for m in re.finditer(r"TEST_F\(\w+, (\w+)\) \{(.*?)\n\}", t, re.S):
        name, body = m.group(1), m.group(2)
        line = t[:m.start()].count("\n") + 1
        el = re.search(r"integral\.setElement\(Utils::ElementType::(\w+)\)", body).group(1)
        given = [[int(a), int(b), int(c), int(d), float(v)] for a, b, c, d, v in
                 re.findall(r"integral\.set\((\d), (\d), (\d), (\d), ([\d.]+) / Utils::Constants::ev_per_hartree\)", body)]
        rec = {"name": name, "source": f"Tests/Nddo/oneCenterTwoElectronIntegralsTest.cpp:{line}", "element": el,
               "given_ev": given, "count": int(re.search(r"getNumberIntegrals\(\), Eq\((\d+)\)", body).group(1))}
        se = re.search(r"sc\.setExponents\(([^)]*)\)", body)
        if se:
            rec["slater_condon"] = {"zeta": [float(x) for x in se.group(1).split(",")],
                                    "set_ev": {k: float(v) for k, v in re.findall(
                                        r"sc\.set\((\w+), ([\d.]+) / Utils::Constants::ev_per_hartree\)", body)}}
        exp = [[int(i), float(v), float(tol)] for i, v, tol in re.findall(
            r"integral\.get\((\d+)\), DoubleNear\(([-\d.]+) / Utils::Constants::ev_per_hartree, ([\de.-]+)\)", body)]
        exp += [[0, float(v), float(tol)] for v, tol in re.findall(
            r"integral\.get\(0, 0, 0, 0\), DoubleNear\(([-\d.]+) / Utils::Constants::ev_per_hartree, ([\de.-]+)\)", body)]
        rec["expected_ev"] = exp
        rec["tolerance_unit"] = "Hartree (the test compares values divided by ev_per_hartree)"
        oc.append(rec)
    out["one_center"] = oc
    # --- LocalIntegralTest: literal rho sets, D from exponents, C-O matrix, H-V entries
    t = open(os.path.join(src, "LocalIntegralTest.cpp")).read()
    setup = t[t.index("void SetUp()"):t.index("TEST_F")]
    rho = {}
    for who, q, v in re.findall(r"K(\w)\.set\(MultipolePair::(\w+), ([\d.]+)\)", setup):
        rho.setdefault(who, {})[q] = float(v)
    dspec = {}
    for who, a in re.findall(r"D(\w)\.computeFromExponents\(([^)]*)\)", setup):
        args = [float(x) for x in a.split(",")]
        dspec[who] = {"ns": int(args[0]), "np": int(args[1]), "nd": int(args[2]) if len(args) == 6 else 0,
                      "zeta": args[3:] if len(args) == 6 else args[2:] + [0.0]}
    loc = {"rho": rho, "charge_separations": dspec, "note": "atoms without a computeFromExponents call keep D = 0"}
    m = re.search(r"TEST_F\(\w+, GivesCorrectResultsForCOPair\) \{(.*?)\n\}", t, re.S)
    body = m.group(1)
    co = {"source": f"Tests/Nddo/LocalIntegralTest.cpp:{t[:m.start()].count(chr(10)) + 1}",
          "R": float(re.search(r"double R = ([\d.]+);", body).group(1)), "atoms": ["C", "O"], "entries": []}
    for i, j, t1, t2 in re.findall(r"mat\((\d), (\d)\) =\s*\(Calc::getIntegral<[^>]*>\(twoElIntegral_t::(\w+), "
                                   r"twoElIntegral_t::(\w+), R, DC, DO, KC, KO\)", body):
        co["entries"].append([int(i), int(j), pairs.index(t1), pairs.index(t2)])
    rule = re.search(r"mat\((\d), (\d)\) = 0\.5 \* \(mat\((\d), (\d)\) - mat\((\d), (\d)\)\);", body)
    co["entry_rule"] = [int(x) for x in rule.groups()]  # mat(a,b) = 0.5 (mat(c,d) - mat(e,f))
    vals = numbers(re.search(r"expectedMatrix <<(.*?);", body, re.S).group(1))
    co["expected_ev_rowmajor"] = vals
    co["tolerance_ev"] = float(re.search(r"DoubleNear\(expectedMatrix\(i, j\), ([\de.-]+)\)", body).group(1))
    loc["CO"] = co
    m = re.search(r"TEST_F\(\w+, GivesCorrectResultsForHVPair\) \{(.*?)\n\}", t, re.S)
    body = m.group(1)
    hv = {"source": f"Tests/Nddo/LocalIntegralTest.cpp:{t[:m.start()].count(chr(10)) + 1}",
          "R": float(re.search(r"double R = ([\d.]+);", body).group(1)), "atoms": ["H", "V"], "entries": []}
    for t1, t2, v, tol in re.findall(r"getIntegral<[^>]*>\(twoElIntegral_t::(\w+), twoElIntegral_t::(\w+), R, DH, DV, "
                                     r"KH, KV\)\)\s*\.value\(\) \*\s*Utils::Constants::ev_per_hartree,\s*"
                                     r"DoubleNear\(([-\d.]+), ([\de.-]+)\)", body):
        hv["entries"].append([pairs.index(t1), pairs.index(t2), float(v), float(tol)])
    loc["HV"] = hv
    out["local_integrals"] = loc
    # --- VuvBMatrixTest: VuvB of a V atom from a point core (charge zb, Klopman term pcore) vs the
    # l=2 x l=0 two-centre block
    t = open(os.path.join(src, "VuvBMatrixTest.cpp")).read()
    vs = {q: float(v) for q, v in re.findall(r"KV\.set\(MultipolePair::(\w+), ([\d.]+)\)", t)}
    a = [float(x) for x in re.search(r"DV\.computeFromExponents\(([^)]*)\)", t).group(1).split(",")]
    out["vuvb"] = {"source": "Tests/Nddo/VuvBMatrixTest.cpp:43-55", "rho_V": vs,
                   "D_V": {"ns": int(a[0]), "np": int(a[1]), "nd": int(a[2]), "zeta": a[3:]},
                   "pcore": float(re.search(r"pcore\{([\d.]+)\}", t).group(1)),
                   "zb": float(re.search(r"zb\{([\d.]+)\}", t).group(1)),
                   "Rab": [float(x) for x in re.search(r"Rab\(([^)]*)\)", t).group(1).split(",")]}
    # --- STO-6G basis records (the files Sto6gBasisTest compares with the runtime expansion)
    sto = {}
    for meth in ("PM6", "MNDO"):
        text = open(os.path.join(RES, meth.capitalize() if meth != "MNDO" else "Mndo", f"{meth}-STO-6G.basis")).read()
        recs = {}
        for blk in text.split("*\n"):
            lines = [ln for ln in blk.split("\n") if ln.strip()]
            if len(lines) == 1 and "STO-6G" in lines[0]:
                cur = lines[0].split()[0]
                continue
            if lines and re.match(r"\s+6\s+[spd]", lines[0]) and cur in ("h", "c", "n", "o", "s"):
                shells, k = [], 0
                while k < len(lines):
                    l = lines[k].split()[1]
                    prims = [[float(x) for x in lines[k + 1 + q].split()] for q in range(6)]
                    shells.append({"l": l, "primitives": prims})
                    k += 7
                recs[cur] = shells
        sto[meth] = recs
    out["sto6g"] = {"source": "Resources/{Pm6,Mndo}/*-STO-6G.basis (Tests/Nddo/Sto6gBasisTest.cpp:127-160)",
                    "records": sto}
    return out


def main():
    data = {"units": "eV for MethodsTests values, Hartree for CLI/molden values", "method_tests": [], "cli": {}}
    for rel, method in (("MethodsTests/Pm6/PM6Test.cpp", "PM6"), ("MethodsTests/Mndo/MNDOTest.cpp", "MNDO"),
                        ("MethodsTests/Am1/AM1Test.cpp", "AM1")):
        data["method_tests"] += parse_method_tests(os.path.join(REF, rel), method)
    app = open(os.path.join(REF, "AppTests/test_sparrow.py")).read()
    cli = []
    for m in re.finditer(r"def (test_energy_sft_1_(\w+?)_charge_0_multiplicity_1_restricted)\(self\):(.*?)assertAlmostEqual\(calculation.get_energy\(\),\s*([-\d.]+),\s*(\d+)\)",
                         app, re.S):
        method = m.group(2).upper()
        if method in ("MNDO", "AM1", "PM3", "RM1", "PM6"):
            line = app[:m.start()].count("\n") + 1
            cli.append({"method": method, "energy": float(m.group(4)), "places": int(m.group(5)),
                        "source": f"AppTests/test_sparrow.py:{line}"})
    data["cli"]["sft_1_energies"] = cli
    m = re.search(r"mopac_energy_sft_5a = ([-\d.]+)\s*mopac_energy_sft_5b = ([-\d.]+)", app)
    data["cli"]["sft_5_mopac"] = {"sft_5a": float(m.group(1)), "sft_5b": float(m.group(2)), "places": 3,
                                 "difference_places": 4, "source": "AppTests/test_sparrow.py:426-451"}
    mo = open(os.path.join(REF, "AppTests/molden_input_reference.molden")).read()
    data["cli"]["sft_1_pm6_molden_orbital_energies"] = [float(x) for x in re.findall(r"Ene=\s*([-\d.]+)", mo)]
    data["structures"] = {nm: open(os.path.join(REF, "AppTests", nm + ".xyz")).read() for nm in ("sft_1", "sft_5a", "sft_5b")}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    with open(OUT_MOLDEN, "w") as f:
        json.dump(molden_fixture(mo), f, indent=0)
    with open(OUT_UNIT, "w") as f:
        json.dump(unit_fixture(), f, indent=1)
    with open(OUT_CIS, "w") as f:
        json.dump(cis_fixture(os.path.join(REF, "TimeDependent/CISTestCalculation.cpp")), f, indent=1)
    print(f"{len(data['method_tests'])} method tests, {len(cli)} CLI energies, "
          f"{len(data['cli']['sft_1_pm6_molden_orbital_energies'])} molden orbital energies -> {OUT}")


if __name__ == "__main__":
    sys.exit(main())
