"""Unrestricted (spin-polarised) SCF around the GPU's spin-polarised Fock build.

The reference runs the spin-polarised NDDO Fock build (NU/TwoElectronMatrix.cpp:82-108,142-160,
NU/NDDOElectronicEnergyCalculator.cpp:69-89) inside the Utils::ScfMethod loop of scine_utilities,
which is not part of its sources.  This loop is host-orchestrated; every O(n^2)-and-above step of
an iteration runs on the GPU through the C ABI:
  - the density-independent part once (a molecule handle, sa_mol_create), then per iteration
    F^alpha, F^beta and E_el from the resident integrals (sa_mol_fock_unrestricted);
  - each spin density as the spectral projector of its Fock matrix onto N_s orbitals (FP64 SP2
    on the GPU, sa_density_projector, the restricted SCF's own kernels); a spin whose projection
    does not converge (degenerate HOMO/LUMO, e.g. an open-shell atom) takes the eigenvectors of
    the device eigensolver (sa_eigh) and P^s = C_occ^s C_occ^sT instead;
  - the products F^s P^s of the DIIS errors (sa_gemm_nt);
  - the orbital energies and orbitals once, from the matrices of the last density build.
The DIIS subspace algebra (Pulay, error FP - PF per spin, both spins in one subspace: dot
products and the extrapolated sums, O(n^2) per history entry) is on the host.  Occupations
follow the aufbau principle per spin: N_alpha = (N + M - 1) / 2, N_beta = (N - M + 1) / 2 for multiplicity M.  The start is the
standard NDDO guess split in half, P^alpha = P^beta = P_guess / 2 (the reference's density guess,
NDDODensityGuess.cpp:31-66, handed to the SCF as one density of N electrons).
"""
import numpy as np

from .params import core_charge


def electron_counts(Z, charge, multiplicity):
    nel = int(sum(core_charge(int(z)) for z in Z)) - int(charge)
    m = int(multiplicity)
    if m < 1 or (nel + m - 1) % 2 or m - 1 > nel:
        raise ValueError(f"spin multiplicity {m} is impossible with {nel} electrons")
    return (nel + m - 1) // 2, (nel - m + 1) // 2


def unrestricted_scf(ctx, method, Z, xyz_bohr, charge=0, multiplicity=2, energy_tol=1e-7, rmsd_tol=1e-5,
                     max_iterations=100, diis_size=8, density_guess=None):
    """Returns energy, electronic and repulsion energy, alpha and beta densities, orbital
    energies and coefficients, iterations and convergence.  diis_size < 2 runs without a mixer;
    density_guess (n x n total density, or a pair (P_alpha, P_beta)) replaces the standard
    guess (a total density starts both spins at half of it)."""
    na, nb = electron_counts(Z, charge, multiplicity)
    with ctx.molecule(method, Z, xyz_bohr, charge) as mol:
        return _scf(ctx, mol, na, nb, energy_tol, rmsd_tol, max_iterations, diis_size, density_guess)


def _scf(ctx, mol, na, nb, energy_tol, rmsd_tol, max_iterations, diis_size, density_guess):
    _, H, erep = mol.density_independent()
    n = H.shape[0]
    if na > n:
        raise ValueError("more electrons of one spin than orbitals")

    def spin_density(F, nocc):
        if nocc == 0:
            return np.zeros((n, n))
        X, ok = ctx.density_projector(F, nocc)
        if ok:
            return X
        _, C = ctx.eigh(F)
        return ctx.gemm_nt(C[:, :nocc], C[:, :nocc])

    def densities(Fa, Fb):
        return spin_density(Fa, na), spin_density(Fb, nb)

    last = (H, H)  # the matrices of the last density build (orbitals are reported from them)
    if density_guess is None:
        Pa = Pb = 0.5 * mol.density_guess()
    else:
        if isinstance(density_guess, (tuple, list)):
            Pa, Pb = (np.asarray(p, dtype=np.float64) for p in density_guess)
        else:
            Pa = Pb = 0.5 * np.asarray(density_guess, dtype=np.float64)
        if Pa.shape != (n, n) or Pb.shape != (n, n):
            raise ValueError(f"density guess must be {n} x {n}")
    # DIIS history: (F^alpha, F^beta, error) entries and their error overlaps, kept across
    # iterations (each new entry adds one row: the earlier dot products do not change)
    hist, dots, E_old, converged, it = [], [], None, False, 0
    for it in range(1, max_iterations + 1):
        Fa, Fb, E = mol.fock_unrestricted(Pa, Pb)
        FPa, FPb = ctx.gemm_nt(Fa, Pa), ctx.gemm_nt(Fb, Pb)  # F P (P symmetric); P F = (F P)^T
        err = np.concatenate([(FPa - FPa.T).ravel(), (FPb - FPb.T).ravel()])
        hist.append((Fa, Fb, err))
        dots.append([h[2] @ err for h in hist])  # row of the new entry: overlaps with entries 0..m-1
        if len(hist) > diis_size:
            drop = len(hist) - diis_size
            hist, dots = hist[drop:], [row[drop:] for row in dots[drop:]]
        if diis_size >= 2 and len(hist) >= 2:
            m = len(hist)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for i in range(m):
                for j in range(i + 1):
                    B[i, j] = B[j, i] = dots[i][j]
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            try:
                c = np.linalg.solve(B, rhs)[:m]
                Fa = sum(ci * h[0] for ci, h in zip(c, hist))
                Fb = sum(ci * h[1] for ci, h in zip(c, hist))
            except np.linalg.LinAlgError:
                hist, dots = hist[-1:], [dots[-1][-1:]]
        Pa_new, Pb_new = densities(Fa, Fb)
        last = (Fa, Fb)
        rmsd = np.sqrt(0.5 * (np.mean((Pa_new - Pa) ** 2) + np.mean((Pb_new - Pb) ** 2)))
        Pa, Pb = Pa_new, Pb_new
        if E_old is not None and abs(E - E_old) < energy_tol and rmsd < rmsd_tol:
            converged = True
            break
        E_old = E
    Fa, Fb, E = mol.fock_unrestricted(Pa, Pb)  # energy of the final densities
    wa, Ca = ctx.eigh(last[0])
    wb, Cb = ctx.eigh(last[1])
    return dict(energy=E + erep, electronic_energy=E, repulsion_energy=erep, Pa=Pa, Pb=Pb, eps_alpha=wa,
                eps_beta=wb, C_alpha=Ca, C_beta=Cb, n_alpha=na, n_beta=nb, iterations=it, converged=converged)
