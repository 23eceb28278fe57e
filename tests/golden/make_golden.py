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
"""
import json
import os
import re
import sys

REF = "/root/reference/src/Sparrow/Tests"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_values.json")
OUT_MOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "molden_sft1_pm6.json")
OUT_CIS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cis_pm3.json")
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
    with open(OUT_CIS, "w") as f:
        json.dump(cis_fixture(os.path.join(REF, "TimeDependent/CISTestCalculation.cpp")), f, indent=1)
    print(f"{len(data['method_tests'])} method tests, {len(cli)} CLI energies, "
          f"{len(data['cli']['sft_1_pm6_molden_orbital_energies'])} molden orbital energies -> {OUT}")


if __name__ == "__main__":
    sys.exit(main())
