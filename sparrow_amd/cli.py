"""Command-line front end with the reference's `sparrow` options and output layout, over the GPU
calculator (NDDO methods MNDO, AM1, PM3, RM1, PM6).

Reference: src/Sparrow/App/CommandLineOptions.cpp:60-136 (option names and short forms) and
src/Sparrow/App/CalculationHandler.cpp (printHeader :437-448, printCalculationConverged :237-244,
printPrettyResults :373-390, printFooter :450-470, printResultsToFile energy.dat :246-260,
printWavefunction :510-530).  The lines a caller parses are the same: "Energy [hartree]:" followed
by the energy (10 significant digits), "Gradients [hartree/bohr]:" followed by a column header
and one "index x y z" row per atom, ended by a blank line.

Usage: python -m sparrow_amd.cli -x structure.xyz -M PM6 [-c 0] [-s 1] [-u] [-m diis] [-I 100]
       [-t 1e-7] [--density_rmsd_criterion 1e-5] [-G] [-W] [-o] [-D description]
       (or the `sparrow` script beside this module).
Options of the reference that concern other methods or properties (Hessians, thermochemistry,
excited states, orbital steering, bond orders, custom parameter files) are rejected with an error;
--temperature, --symmetry_number, --spin_block, --number_orbital_mixes and --log are accepted and
have no effect on a ground-state energy or gradient.
"""
import argparse
import datetime
import os
import socket
import sys
import time

import numpy as np

UNSUPPORTED = (("hessian", "H"), ("atomic_hessians", "A"), ("thermochemistry", "C"), ("excited_states", "E"),
               ("bond_orders", "B"), ("suppress_normal_modes", "N"))


def parser():
    ap = argparse.ArgumentParser(prog="sparrow", description="NDDO ground-state calculations on the GPU")
    ap.add_argument("-x", "--structure", required=False, help="sets the path to the coordinates file")
    ap.add_argument("-M", "--method", default="PM6", help="sets the method used to perform the calculation")
    ap.add_argument("-c", "--molecular_charge", type=int, default=0)
    ap.add_argument("-s", "--spin_multiplicity", type=int, default=1)
    ap.add_argument("-u", "--unrestricted", action="store_true", help="runs unrestricted calculation (UHF)")
    ap.add_argument("-m", "--scf_mixer", default="diis")
    ap.add_argument("-I", "--max_scf_iterations", type=int, default=100)
    ap.add_argument("-t", "--self_consistence_criterion", type=float, default=1e-7)
    ap.add_argument("--density_rmsd_criterion", type=float, default=1e-5)
    ap.add_argument("-p", "--method_parameters", default="")
    ap.add_argument("-G", "--gradients", action="store_true")
    ap.add_argument("-W", "--wavefunction", action="store_true", help="outputs a molden file of the orbitals")
    ap.add_argument("-o", "--output_to_file", action="store_true")
    ap.add_argument("-D", "--description", default="")
    ap.add_argument("-l", "--log", default="warning")
    ap.add_argument("-T", "--temperature", type=float, default=298.15)
    ap.add_argument("-S", "--symmetry_number", type=int, default=1)
    ap.add_argument("-b", "--spin_block", default="singlet")
    ap.add_argument("-O", "--number_orbital_steers", type=int, default=0)
    ap.add_argument("--number_orbital_mixes", type=int, default=0)
    for long, short in UNSUPPORTED:
        ap.add_argument(f"-{short}", f"--{long}", action="store_true")
    return ap


def run(argv, out=sys.stdout):
    from .calculator import Calculator
    from .structure import AtomCollection
    a = parser().parse_args(argv)
    if not a.structure:
        raise SystemExit("No coordinate file given.")
    asked = [f"--{long}" for long, _ in UNSUPPORTED if getattr(a, long)]
    if asked or a.number_orbital_steers:
        raise SystemExit(f"not available in this build (NDDO ground state only): {asked or ['--number_orbital_steers']}")
    calc = Calculator(a.method)
    s = calc.settings
    s["molecular_charge"] = a.molecular_charge
    s["spin_multiplicity"] = a.spin_multiplicity
    s["spin_mode"] = "unrestricted" if a.unrestricted else ("restricted" if a.spin_multiplicity == 1 else "any")
    s["scf_mixer"] = a.scf_mixer
    s["max_scf_iterations"] = a.max_scf_iterations
    s["self_consistence_criterion"] = a.self_consistence_criterion
    s["density_rmsd_criterion"] = a.density_rmsd_criterion
    s["method_parameters"] = a.method_parameters
    with open(a.structure) as f:
        calc.set_structure(AtomCollection.from_xyz_string(f.read()))
    props = ["energy"]
    if a.gradients:
        props.append("gradients")
    if a.wavefunction:
        props += ["orbital_energies", "coefficient_matrix"]
    calc.set_required_properties(props)
    call = "sparrow " + " ".join(argv)
    out.write("SPARROW, command-line program\n")
    out.write(f"Host: {socket.gethostname()}\n")
    out.write(f"Start: {datetime.datetime.now().strftime('%c')}\n\n")
    out.write(f"Program call: {call}\n")
    out.write(f"{'Method: ':<25}{calc.name()}\n")
    for k in ("molecular_charge", "spin_multiplicity", "spin_mode", "self_consistence_criterion",
              "density_rmsd_criterion", "max_scf_iterations", "scf_mixer"):
        out.write(f"{k + ': ':<40}{s[k]}\n")
    out.write("\n")
    t0 = time.perf_counter()
    r = calc.calculate(a.description)
    ms = (time.perf_counter() - t0) * 1e3
    out.write("SCF converged!\n" if r["successful_calculation"] else "SCF not converged!\n")
    out.write("=" * 80 + "\n")
    if not r["successful_calculation"]:
        out.write("Be warned: the scf method has not converged.\n")
    out.write(f"Calculation: {a.description}\n\n")
    out.write("Energy [hartree]:\n")
    out.write(f"{r['energy']:.10g}\n\n")
    if a.gradients:
        out.write("Gradients [hartree/bohr]:\n")
        out.write(f"{'':>6}{'X':>20}{'Y':>20}{'Z':>20}\n")
        for i, g in enumerate(np.asarray(r["gradients"])):
            out.write(f"{i:>6}{g[0]:>20.6e}{g[1]:>20.6e}{g[2]:>20.6e}\n")
        out.write("\n")
    out.write(f"\n{'Ground state calculation time: ':<40}{ms:>5.0f} ms.\n")
    under = "_" if a.description else ""
    if a.output_to_file:
        with open(f"{a.description}{under}energy.dat", "w") as f:
            f.write(f"# SPARROW, command-line program\n# Program call: {call}\n# Energy [hartree]:\n{r['energy']:.10g}\n")
    if a.wavefunction:
        if not r["successful_calculation"]:
            raise SystemExit("Wavefunction requested, but ground state calculation was not successful.")
        calc.write_molden(f"{a.description}{under}wavefunction.molden.input")
    return r


def main():
    run(sys.argv[1:])


if __name__ == "__main__":
    main()
