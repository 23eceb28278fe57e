"""Unrestricted SCF (sparrow_amd.uhf: GPU Fock builds, eigensolves and products) through the
calculator front end: the reference's doublet boron (PM6Test.cpp:211-233, MOPAC values), the
restricted limit of a closed-shell molecule, and a radical against the oracle's spin-polarised
G-build iterated to self-consistency on the CPU.
"""
import numpy as np
import pytest

import sparrow_amd
from sparrow_amd.structure import AtomCollection

pytestmark = pytest.mark.gpu
EV = 27.21138602


def _calc(method, text, mult, mode="unrestricted"):
    c = sparrow_amd.get_calculator(method)
    c.set_structure(AtomCollection.from_xyz_string(text))
    c.settings["spin_multiplicity"] = mult
    c.settings["spin_mode"] = mode
    c.settings["self_consistence_criterion"] = 1e-10
    c.settings["density_rmsd_criterion"] = 1e-8
    c.settings["max_scf_iterations"] = 300
    c.set_required_properties(["energy", "orbital_energies", "density_matrix"])
    return c


def test_unrestricted_boron_doublet_matches_reference():
    c = _calc("PM6", "1\n\nB -0.1058354421 0.1058354421 -0.1587531632\n", 2)
    r = c.calculate()
    assert r.successful_calculation
    assert abs(r.energy * EV - (-49.53668)) <= 1e-5
    wa, wb = r["orbital_energies"]["alpha"], r["orbital_energies"]["beta"]
    assert np.max(np.abs(wa * EV - np.array([-11.74716, -5.78067, -0.09376, -0.09376]))) <= 1e-5
    assert np.max(np.abs(wb * EV - np.array([-10.49432, 0.62041, 0.62041, 2.04873]))) <= 1e-5


def test_unrestricted_closed_shell_is_restricted():
    text = "4\n\nC 0 0 0\nO 0 0 1.2122\nH 0.937197 0 -0.584262\nH -0.937197 0 -0.584262\n"
    u = _calc("AM1", text, 1).calculate()
    c = _calc("AM1", text, 1, mode="restricted")
    c.settings["density_builder"] = "diagonalization"
    r = c.calculate()
    assert u.successful_calculation and r.successful_calculation
    assert abs(u.energy - r.energy) <= 1e-8
    assert np.max(np.abs(u["density_matrix"] - r["density_matrix"])) <= 1e-5


@pytest.mark.parametrize("method", ["PM6", "MNDO"])
def test_unrestricted_radical_matches_oracle(oracles, method):
    """Methyl radical (doublet): the GPU-driven UHF energy against the oracle's spin-polarised
    G-build iterated on the CPU from the GPU's converged densities (a fixed point: one oracle
    Fock build, aufbau per spin, repeated until stationary)."""
    text = "4\n\nC 0 0 0\nH 1.08 0 0\nH -0.54 0.935 0\nH -0.54 -0.935 0.05\n"
    c = _calc(method, text, 2)
    c.set_required_properties(["energy", "density_matrix"])
    r = c.calculate()
    assert r.successful_calculation
    from sparrow_amd.uhf import unrestricted_scf
    from sparrow_amd import _native
    m = AtomCollection.from_xyz_string(text)
    ctx = _native.Context(0)
    g = unrestricted_scf(ctx, method, m.Z, m.positions, 0, 2, 1e-10, 1e-8, 300)
    ctx.close()
    o = oracles[method]
    _, H, erep = o.density_independent(m.Z, m.positions)
    H = np.tril(H) + np.tril(H, -1).T
    Pa, Pb = g["Pa"], g["Pb"]
    for _ in range(200):
        Ga, Gb, E = o.gmatrix_unrestricted(m.Z, m.positions, Pa, Pb)
        Fa = H + np.tril(Ga) + np.tril(Ga, -1).T
        Fb = H + np.tril(Gb) + np.tril(Gb, -1).T
        Ca, Cb = np.linalg.eigh(Fa)[1], np.linalg.eigh(Fb)[1]
        Pa_new = Ca[:, :g["n_alpha"]] @ Ca[:, :g["n_alpha"]].T
        Pb_new = Cb[:, :g["n_beta"]] @ Cb[:, :g["n_beta"]].T
        d = max(np.max(np.abs(Pa_new - Pa)), np.max(np.abs(Pb_new - Pb)))
        Pa, Pb = Pa_new, Pb_new
        if d < 1e-10:
            break
    assert d < 1e-8
    assert abs(g["energy"] - (E + erep)) <= 1e-8
    assert abs(r.energy - g["energy"]) <= 1e-9
