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


METHYL = "4\n\nC 0 0 0\nH 1.08 0 0\nH -0.54 0.935 0\nH -0.54 -0.935 0.05\n"
CH2CL = "4\n\nC 0 0 0\nCl 1.75 0 0.02\nH -0.52 0.93 0\nH -0.52 -0.93 0.1\n"


def _uhf(ctx, method, text, mult=2, dE=1e-11, rmsd=1e-9):
    from sparrow_amd.uhf import unrestricted_scf
    m = AtomCollection.from_xyz_string(text)
    g = unrestricted_scf(ctx, method, m.Z, m.positions, 0, mult, dE, rmsd, 300)
    assert g["converged"]
    return m.Z, m.positions, g


@pytest.mark.parametrize("method,text", [("PM6", METHYL), ("MNDO", METHYL), ("AM1", CH2CL), ("PM6", CH2CL)])
def test_unrestricted_gradients_match_oracle_at_fixed_density(oracles, method, text):
    """Pair-energy derivatives of the spin-polarised energy at fixed P^alpha, P^beta against
    central differences (h = 1e-4 bohr) of the oracle's own unrestricted E_el + E_rep: pins the
    per-spin exchange term (-1/2 sum (P^alpha^2 + P^beta^2)) of every geometry-dependent integral,
    the d-shell classes included (PM6 and AM1 chlorine)."""
    from sparrow_amd import _native
    ctx = _native.Context(0)
    try:
        Z, X, g = _uhf(ctx, method, text)
        grad = ctx.gradients_unrestricted(method, Z, X, g["Pa"], g["Pb"])
    finally:
        ctx.close()
    o = oracles[method]

    def energy(Xd):
        _, _, erep = o.density_independent(Z, Xd)
        _, _, eel = o.gmatrix_unrestricted(Z, Xd, g["Pa"], g["Pb"])
        return eel + erep

    h = 1e-4
    ref = np.zeros_like(grad)
    for a in range(len(Z)):
        for k in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[a, k] += h
            Xm[a, k] -= h
            ref[a, k] = (energy(Xp) - energy(Xm)) / (2 * h)
    print(f"{method} UHF: max |g - g_oracle| {np.max(np.abs(grad - ref)):.2e}, max |g| {np.max(np.abs(grad)):.2e}",
          flush=True)
    assert np.max(np.abs(grad - ref)) <= 1e-7
    assert np.max(np.abs(grad.sum(axis=0))) <= 1e-10


def test_unrestricted_gradients_reduce_to_restricted(ctx):
    """P^alpha = P^beta = P/2 makes the spin density vanish: the unrestricted gradient equals the
    restricted one at the same total density."""
    m = AtomCollection.from_xyz_string("4\n\nC 0 0 0\nO 0 0 1.2122\nH 0.937197 0 -0.584262\nH -0.937197 0 -0.584262\n")
    r = ctx.scf("PM6", m.Z, m.positions)
    g_r = ctx.gradients("PM6", m.Z, m.positions, r["P"])
    g_u = ctx.gradients_unrestricted("PM6", m.Z, m.positions, 0.5 * r["P"], 0.5 * r["P"])
    assert np.max(np.abs(g_u - g_r)) <= 1e-10


@pytest.mark.parametrize("method", ["PM6", "RM1"])
def test_unrestricted_gradients_are_derivatives_of_the_converged_energy(method):
    """Central differences of the converged UHF energy (tight SCF at each displaced geometry)
    equal the fixed-density gradient: the unrestricted energy is stationary in both spin
    densities.  Also the calculator reports the same gradients."""
    from sparrow_amd import _native
    ctx = _native.Context(0)
    try:
        Z, X, g = _uhf(ctx, method, METHYL, dE=1e-12, rmsd=1e-10)
        grad = ctx.gradients_unrestricted(method, Z, X, g["Pa"], g["Pb"])
        h = 1e-4
        ref = np.zeros_like(grad)
        from sparrow_amd.uhf import unrestricted_scf
        for a in range(len(Z)):
            for k in range(3):
                e = []
                for sgn in (1, -1):
                    Xd = X.copy()
                    Xd[a, k] += sgn * h
                    rd = unrestricted_scf(ctx, method, Z, Xd, 0, 2, 1e-12, 1e-10, 300)
                    assert rd["converged"]
                    e.append(rd["energy"])
                ref[a, k] = (e[0] - e[1]) / (2 * h)
    finally:
        ctx.close()
    print(f"{method} UHF: max |g - dE/dR| {np.max(np.abs(grad - ref)):.2e}", flush=True)
    assert np.max(np.abs(grad - ref)) <= 1e-6
    c = _calc(method, METHYL, 2)
    c.set_required_properties(["energy", "gradients"])
    res = c.calculate()
    assert res.successful_calculation
    assert np.max(np.abs(res["gradients"] - grad)) <= 1e-6


def test_density_projector_matches_the_eigenvector_density(ctx):
    """sa_density_projector (FP64 SP2 on the GPU, the Lanczos-narrowed bounds from 400 AOs on)
    against C_occ C_occ^T of numpy's eigensolver for random symmetric matrices; a degenerate
    HOMO/LUMO has no idempotent projector of that trace and is reported as not converged."""
    rng = np.random.default_rng(3)
    for n, nocc in ((37, 11), (450, 200)):
        A = rng.standard_normal((n, n))
        F = 0.5 * (A + A.T) / np.sqrt(n)
        X, ok = ctx.density_projector(F, nocc)
        w, V = np.linalg.eigh(F)
        assert ok, n
        assert np.max(np.abs(X - V[:, :nocc] @ V[:, :nocc].T)) <= 1e-10, n
    X, ok = ctx.density_projector(np.diag([0.0, 1.0, 1.0, 2.0]), 2)
    assert not ok
