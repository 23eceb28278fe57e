"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle and numpy.

Bars: integrals / S / H / F element-wise to 1e-11 (FP64, only the summation order differs);
converged SCF |dE| <= 1e-8 Eh, max|dP| <= 1e-7, orbital energies <= 1e-7 Eh.
"""
import os

import numpy as np
import pytest

from sparrow_amd import generators as gen
from sparrow_amd.structure import BOHR_PER_ANGSTROM, AtomCollection

pytestmark = pytest.mark.gpu

ETHANOL = """9

H      1.9655905060   -0.0263662325    1.0690084915
C      1.3088788172   -0.0403821764    0.1943189946
H      1.5790293586    0.8034866305   -0.4554748131
H      1.5186511399   -0.9518066799   -0.3824432806
C     -0.1561112248    0.0249676675    0.5877379610
H     -0.4682794700   -0.8500294693    1.1854276282
H     -0.4063173598    0.9562730342    1.1264955766
O     -0.8772416674    0.0083263307   -0.6652828084
H     -1.8356000997    0.0539308952   -0.5014877498
"""
CH2O = """4

C      0.0000000000    0.0000000000    0.0000000000
O      0.0529177211   -0.3175063264    0.2645886053
H     -0.5291772107    0.1058354421   -0.1587531632
H     -0.1058354421    0.1058354421   -0.1587531632
"""
CL2 = "2\n\nCl -0.99 0.0 0.0\nCl 0.99 0.0 0.0\n"
CH3SH = """6

S      0.0000000000    0.0000000000    0.0000000000
C      1.8100000000    0.0000000000    0.0000000000
H     -0.4000000000    1.3000000000    0.0000000000
H      2.1500000000    1.0200000000    0.0000000000
H      2.1500000000   -0.5100000000    0.8800000000
H      2.1500000000   -0.5100000000   -0.8800000000
"""

def mol(text):
    m = AtomCollection.from_xyz_string(text)
    return m.Z, m.positions


def test_gemm_matches_numpy(ctx):
    rng = np.random.default_rng(0)
    for M, N, K in [(1, 1, 1), (17, 5, 3), (70, 50, 13), (130, 129, 64), (64, 64, 0)]:
        A = rng.normal(size=(M, K))
        B = rng.normal(size=(N, K))
        C = ctx.gemm_nt(A, B)
        ref = A @ B.T
        assert np.max(np.abs(C - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref))) * max(K, 1)


@pytest.mark.parametrize("n", [1, 2, 5, 37, 130, 401])
def test_eigh_matches_numpy(ctx, n):
    rng = np.random.default_rng(n)
    X = rng.normal(size=(n, n))
    A = 0.5 * (X + X.T)
    w, V = ctx.eigh(A)
    wr = np.linalg.eigvalsh(A)
    scale = max(1.0, np.max(np.abs(wr)))
    assert np.max(np.abs(w - wr)) <= 1e-12 * scale
    assert np.max(np.abs(A @ V - V * w)) <= 1e-11 * scale
    assert np.max(np.abs(V.T @ V - np.eye(n))) <= 1e-11


@pytest.mark.parametrize("n", [60, 260])  # one-workgroup and many-workgroup tridiagonalisation
def test_eigh_degenerate_spectrum(ctx, n):
    rng = np.random.default_rng(7)
    w0 = np.repeat(np.arange(n // 5, dtype=float), 5)  # every eigenvalue five-fold degenerate
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = (Q * w0) @ Q.T
    A = 0.5 * (A + A.T)
    w, V = ctx.eigh(A, nvec=30)
    assert np.max(np.abs(w - w0)) <= 1e-12 * (n // 5)
    assert np.max(np.abs(V.T @ V - np.eye(30))) <= 1e-10
    assert np.max(np.abs(A @ V - V * w[:30])) <= 1e-10


@pytest.mark.parametrize("ZA,ZB,R", [(1, 1, (0.3, -0.2, 1.4)), (6, 1, (1.1, 0.4, -1.3)), (1, 6, (-0.7, 1.2, 0.5)),
                                     (6, 8, (0.0, 0.0, 2.3)), (8, 6, (0.0, 0.0, -2.3)), (7, 8, (2.1, 0.0, 0.0)),
                                     (16, 6, (1.5, 2.0, -0.8)), (6, 16, (-1.9, 0.6, 2.2)), (16, 16, (0.9, -1.7, 3.1)),
                                     (1, 16, (2.2, 1.1, 0.3)), (16, 1, (0.0, 2.5, 0.0))])
def test_two_center_blocks_match_oracle(ctx, oracles, ZA, ZB, R):
    from oracle import oracle as O
    R = np.array(R) * 1.3
    for method in ("PM6", "MNDO"):
        a = oracles[method].atom_params(ZA)
        b = oracles[method].atom_params(ZB)
        ref = O.global_block(int(a["l"]), int(b["l"]), R, a["D"], b["D"], a["rho"], b["rho"])
        got = ctx.two_center_block(method, ZA, ZB, R)
        assert got.shape == ref.shape
        assert np.max(np.abs(got - ref)) <= 1e-12, method


@pytest.mark.parametrize("method,text", [("PM6", CH2O), ("PM6", ETHANOL), ("PM6", CL2), ("PM6", CH3SH),
                                         ("MNDO", ETHANOL), ("AM1", CH2O), ("PM3", CH3SH), ("RM1", ETHANOL)])
def test_density_independent_matches_oracle(ctx, oracles, method, text):
    Z, X = mol(text)
    S, H, E = ctx.density_independent(method, Z, X)
    So, Ho, Eo = oracles[method].density_independent(Z, X)
    low = np.tril_indices(len(S))
    assert np.max(np.abs(S[low] - So[low])) <= 1e-12
    assert np.max(np.abs(H[low] - Ho[low])) <= 1e-11
    assert np.allclose(S, S.T, atol=0, rtol=0) and np.allclose(H, H.T, atol=0, rtol=0)
    assert abs(E - Eo) <= 1e-11 * max(1.0, abs(Eo))


@pytest.mark.parametrize("method,text", [("PM6", ETHANOL), ("PM6", CH3SH), ("MNDO", CH2O), ("PM6", CL2)])
def test_fock_matches_oracle(ctx, oracles, method, text):
    Z, X = mol(text)
    o = oracles[method]
    P = o.density_guess(Z, X)
    rng = np.random.default_rng(3)
    D = rng.normal(scale=0.05, size=P.shape)
    P = P + 0.5 * (D + D.T)
    F, E = ctx.fock(method, Z, X, P)
    G, Eo = o.gmatrix(Z, X, P)
    S, H, _ = o.density_independent(Z, X)
    low = np.tril_indices(len(P))
    assert np.max(np.abs(F[low] - (H + G)[low])) <= 1e-11
    assert abs(E - Eo) <= 1e-10


def _oracle_fixed_density_energy(o, Z, X, P):
    """E_el(P) + E_rep of the oracle at geometry X for a fixed density P."""
    _, _, erep = o.density_independent(Z, X)
    _, eel = o.gmatrix(Z, X, P)
    return eel + erep


@pytest.mark.parametrize("method,text", [("PM6", ETHANOL), ("PM6", CH3SH), ("PM6", CL2), ("MNDO", CH2O), ("AM1", ETHANOL),
                                         ("PM3", CH3SH), ("RM1", CH2O)])
def test_gradients_match_oracle_at_fixed_density(ctx, oracles, method, text):
    """Pair-energy derivatives at fixed P against central differences (h = 1e-4 bohr) of the
    oracle's own E_el(P) + E_rep: pins every geometry-dependent term (beta S, VuvB, the
    two-centre Coulomb and exchange, core repulsion) through an independent CPU code path."""
    Z, X = mol(text)
    o = oracles[method]
    r = ctx.scf(method, Z, X, settings=_tight())
    g = ctx.gradients(method, Z, X, r["P"])
    h = 1e-4
    ref = np.zeros_like(g)
    for a in range(len(Z)):
        for k in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[a, k] += h
            Xm[a, k] -= h
            ref[a, k] = (_oracle_fixed_density_energy(o, Z, Xp, r["P"]) - _oracle_fixed_density_energy(o, Z, Xm, r["P"])) / (2 * h)
    print(f"{method} {len(Z)} atoms: max |g - g_oracle| {np.max(np.abs(g - ref)):.2e}, max |g| {np.max(np.abs(g)):.2e}",
          flush=True)
    assert np.max(np.abs(g - ref)) <= 1e-7


@pytest.mark.parametrize("method,text", [("PM6", ETHANOL), ("MNDO", CH3SH)])
def test_gradients_are_derivatives_of_the_converged_energy(ctx, oracles, method, text):
    """The full derivative of the converged SCF energy (oracle SCF at displaced geometries, tight
    convergence, central differences) equals the fixed-density derivative: the SCF energy is
    stationary in P.  Also rigid-motion invariance: the forces and torques sum to zero."""
    Z, X = mol(text)
    o = oracles[method]
    r = ctx.scf(method, Z, X, settings=_tight())
    g = ctx.gradients(method, Z, X, r["P"])
    h = 1e-4
    ref = np.zeros_like(g)
    for a in range(len(Z)):
        for k in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[a, k] += h
            Xm[a, k] -= h
            ep = o.scf(Z, Xp, 0, dE=1e-13, rmsd=1e-12, maxit=500, P_guess=r["P"])
            em = o.scf(Z, Xm, 0, dE=1e-13, rmsd=1e-12, maxit=500, P_guess=r["P"])
            assert ep["converged"] and em["converged"]
            ref[a, k] = (ep["energy"] - em["energy"]) / (2 * h)
    print(f"{method}: max |g - dE/dR| {np.max(np.abs(g - ref)):.2e}", flush=True)
    assert np.max(np.abs(g - ref)) <= 1e-6
    assert np.max(np.abs(g.sum(axis=0))) <= 1e-10
    assert np.max(np.abs(np.cross(X, g).sum(axis=0))) <= 1e-6


# MOPAC 2016 gradients (kcal/mol/A) of methane at the reference tests' geometry, as asserted by
# src/Sparrow/Tests/MethodsTests/Pm6/PM6Test.cpp:417-458 and Mndo/MNDOTest.cpp:294-336 (row norm
# 1e-4 Hartree/bohr after conversion); zero-gradient geometries: PM6Test.cpp:461-492,
# MNDOTest.cpp:338-370 (row norm 1e-3).
METHANE = "5\n\nC 0.0 0.0 0.0\nH 0.6287 0.6287 0.6287\nH -0.6287 -0.6287 0.6287\nH -0.6287 0.6287 -0.6287\nH 0.6287 -0.6287 -0.6287\n"
METHANE_GRAD = {
    "PM6": [(0.000009, -0.000017, 0.000005), (1.725729, 1.725682, 1.725542), (-1.742019, -1.742019, 1.756755),
            (-1.742116, 1.757545, -1.741122), (1.758416, -1.741192, -1.741180)],
    "MNDO": [(0.000010, 0.000011, 0.000012), (-8.984101, -8.982663, -8.983207), (8.956718, 8.957105, -8.935515),
             (8.957231, -8.933676, 8.958441), (-8.929858, 8.959223, 8.960270)],
}
METHANE_ENERGY_EV = {"PM6": -177.169, "MNDO": -185.07194}
METHANE_OPT = {
    "PM6": "5\n\nC 0.00000000 0.00000001 -0.00000097\nH 0.62612502 0.62612484 0.62613824\nH -0.62612503 -0.62612486 0.62613824\n"
           "H -0.62612481 0.62612463 -0.62613657\nH 0.62612481 -0.62612464 -0.62613657\n",
    "MNDO": "5\n\nC 0.00003086 0.00006509 -0.00021557\nH 0.63772679 0.63840547 0.63682059\nH -0.63773658 -0.63845649 0.63671348\n"
            "H -0.63765374 0.63821789 -0.63638609\nH 0.63755701 -0.63839156 -0.63640383\n",
}


@pytest.mark.parametrize("method", ["PM6", "MNDO"])
def test_gradients_of_methane_match_reference(ctx, method):
    Z, X = mol(METHANE)
    r = ctx.scf(method, Z, X, settings=_tight())
    g = ctx.gradients(method, Z, X, r["P"])
    ref = np.array(METHANE_GRAD[method]) * 0.52917721067 / 627.5094740631  # kcal/mol/A -> Hartree/bohr
    print(f"{method} methane: E {r['energy'] * 27.21138602:.5f} eV, row errors {np.linalg.norm(g - ref, axis=1)}", flush=True)
    assert abs(r["energy"] * 27.21138602 - METHANE_ENERGY_EV[method]) <= 1e-3
    assert np.all(np.linalg.norm(g - ref, axis=1) < 1e-4)
    Zo, Xo = mol(METHANE_OPT[method])
    ro = ctx.scf(method, Zo, Xo, settings=_tight())
    go = ctx.gradients(method, Zo, Xo, ro["P"])
    assert np.all(np.linalg.norm(go, axis=1) < 1e-3)


def _scf_pair(ctx, oracle, method, Z, X, charge=0):
    r = ctx.scf(method, Z, X, charge, settings=_tight(), want_orbitals=True)
    o = oracle.scf(Z, X, charge, dE=1e-11, rmsd=1e-10, maxit=500)
    return r, o


def _tight(builder="sp2"):
    from sparrow_amd import _native
    return _native.settings_struct(1e-11, 1e-10, 500, "diis", 8, density_builder=builder)


@pytest.mark.parametrize("method,text", [("PM6", "2\n\nH -0.529177 0.105835 -0.158753\nH -0.105835 0.105835 -0.158753\n"),
                                         ("PM6", CH2O), ("PM6", ETHANOL), ("PM6", CL2), ("PM6", CH3SH),
                                         ("MNDO", ETHANOL), ("AM1", CH2O), ("PM3", ETHANOL), ("RM1", CH3SH)])
def test_scf_matches_oracle(ctx, oracles, method, text):
    Z, X = mol(text)
    r, o = _scf_pair(ctx, oracles[method], method, Z, X)
    assert r["converged"] and o["converged"]
    assert abs(r["energy"] - o["energy"]) <= 1e-8
    assert abs(r["repulsion_energy"] - o["repulsion_energy"]) <= 1e-10
    assert np.max(np.abs(r["P"] - o["P"])) <= 1e-7
    assert np.max(np.abs(r["eps"] - o["eps"])) <= 1e-7
    C = r["C"]
    assert np.max(np.abs(C.T @ C - np.eye(len(C)))) <= 1e-10


def test_water_cluster_mndo(ctx, oracles):
    """MNDO (H2O)8, the CPU-runnable reference workload (24 atoms, 48 AOs)."""
    Z, X = gen.water_cluster(seed=0)
    r, o = _scf_pair(ctx, oracles["MNDO"], "MNDO", Z, X * BOHR_PER_ANGSTROM)
    assert r["converged"] and r["nao"] == 48
    assert abs(r["energy"] - o["energy"]) <= 1e-8
    assert np.max(np.abs(r["P"] - o["P"])) <= 1e-7


@pytest.mark.parametrize("n,seed", [(12, 1), (40, 2), (90, 3)])
def test_random_pm6_molecules(ctx, oracles, n, seed):
    Z, X = gen.random_molecule(n, seed=seed)
    r, o = _scf_pair(ctx, oracles["PM6"], "PM6", Z, X * BOHR_PER_ANGSTROM)
    assert r["converged"] and o["converged"]
    assert abs(r["energy"] - o["energy"]) <= 1e-8
    assert np.max(np.abs(r["P"] - o["P"])) <= 1e-7
    assert np.max(np.abs(r["eps"] - o["eps"])) <= 1e-7


def test_batch_equals_single(ctx, tiles):
    tiles(DEFAULT_SP2_TILES[True])  # one tile table for both forms: compares batching only
    from sparrow_amd import _native
    mols = [gen.random_molecule(n, seed=100 + n) for n in (8, 23, 31)]
    mols = [(Z, X * BOHR_PER_ANGSTROM, 0) for Z, X in mols]
    st = _native.settings_struct(1e-10, 1e-9, 300, "diis", 8)
    batch, _ = ctx.scf_batch("PM6", mols, settings=st)
    for (Z, X, c), b in zip(mols, batch):
        s = ctx.scf("PM6", Z, X, c, settings=st)
        assert b["iterations"] == s["iterations"]
        assert b["energy"] == s["energy"]
        assert np.array_equal(b["P"], s["P"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("method", ["MNDO", "AM1", "PM3", "RM1"])
def test_batched_methods_match_oracle(ctx, oracles, method, precision):
    """The batched path for the other NDDO methods: a water cluster and three seeded C/H/N/O/S
    molecules (15, 30 and 45 atoms, sulfur included) in one batch, each against the oracle's
    tightly converged solution: FP64 within 1e-8 Eh and densities within 1e-7, mixed within the
    metric's 1e-6 Eh."""
    from sparrow_amd import _native
    mols = [gen.water_cluster(seed=1)] + [gen.random_molecule(n, seed=900 + n) for n in (15, 30, 45)]
    mols = [(Z, X * BOHR_PER_ANGSTROM, 0) for Z, X in mols]
    st = _native.settings_struct(1e-11, 1e-10, 500, "diis", 8, precision=precision)
    res, _ = ctx.scf_batch(method, mols, settings=st)
    for k, ((Z, X, c), r) in enumerate(zip(mols, res)):
        o = oracles[method].scf(Z, X, c, dE=1e-11, rmsd=1e-10, maxit=800)
        assert r["converged"] and o["converged"], (method, k)
        dE = abs(r["energy"] - o["energy"])
        if precision == "fp64":
            assert dE <= 1e-8, (method, k, dE)
            assert np.max(np.abs(r["P"] - o["P"])) <= 1e-7, (method, k)
        else:
            assert dE <= 1e-6, (method, k, dE)


def test_warm_start_converges_faster(ctx):
    from sparrow_amd import _native
    Z, X = gen.random_molecule(30, seed=5)
    X = X * BOHR_PER_ANGSTROM
    st = _native.settings_struct(1e-10, 1e-9, 300, "diis", 8)
    cold = ctx.scf("PM6", Z, X, settings=st)
    rng = np.random.default_rng(0)
    X2 = X + rng.normal(scale=0.05 * BOHR_PER_ANGSTROM, size=X.shape)
    warm = ctx.scf("PM6", Z, X2, settings=st, density_guess=cold["P"])
    fresh = ctx.scf("PM6", Z, X2, settings=st)
    assert warm["converged"] and fresh["converged"]
    assert abs(warm["energy"] - fresh["energy"]) <= 1e-8
    assert warm["iterations"] <= fresh["iterations"]


@pytest.mark.parametrize("method,text", [("PM6", ETHANOL), ("MNDO", CH2O), ("PM6", CH3SH), ("AM1", CL2)])
def test_sp2_and_diagonalization_builders_agree(ctx, oracles, method, text):
    """The SP2 projection and the reference's diagonalisation give the same SCF solution."""
    Z, X = mol(text)
    a = ctx.scf(method, Z, X, settings=_tight("sp2"))
    ca = ctx.solve_counters()
    b = ctx.scf(method, Z, X, settings=_tight("diagonalization"))
    cb = ctx.solve_counters()
    assert ca["sp2_steps"] > 0 and ca["sp2_fallbacks"] == 0
    assert cb["sp2_steps"] == 0 and cb["eigensolves"] >= b["iterations"]
    assert a["converged"] and b["converged"]
    o = oracles[method].scf(Z, X, 0, dE=1e-11, rmsd=1e-10, maxit=500)
    dP = np.max(np.abs(a["P"] - b["P"]))
    print(f"{method}: iterations sp2/diag/oracle {a['iterations']}/{b['iterations']}/{o['iterations']}, "
          f"sp2 steps {ca['sp2_steps']}, dE {a['energy'] - b['energy']:.3e}, max dP {dP:.3e}, "
          f"vs oracle: sp2 {np.max(np.abs(a['P'] - o['P'])):.3e} diag {np.max(np.abs(b['P'] - o['P'])):.3e}",
          flush=True)
    for r in (a, b):
        assert abs(r["energy"] - o["energy"]) <= 1e-8
        assert np.max(np.abs(r["P"] - o["P"])) <= 1e-7
    assert abs(a["energy"] - b["energy"]) <= 1e-9
    assert dP <= 1e-7
    # orbital energies are those of the last DIIS-extrapolated matrix; the two runs take different
    # DIIS paths (the convergence test bounds dE and RMSD(dP), not the extrapolated matrix)
    assert np.max(np.abs(a["eps"] - b["eps"])) <= 1e-5


@pytest.mark.parametrize("Z,X,charge", [([2], [[0.0, 0.0, 0.0]], 0),                    # every orbital occupied
                                        ([1, 1], [[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]], 2),  # no electron
                                        ([8, 1, 1], [[0, 0, 0], [0, 1.4, 1.1], [0, -1.4, 1.1]], 0)])
def test_scf_occupation_edge_cases(ctx, oracles, Z, X, charge):
    Z = np.array(Z)
    X = np.array(X, dtype=float)
    r = ctx.scf("PM6", Z, X, charge, settings=_tight())
    o = oracles["PM6"].scf(Z, X, charge, dE=1e-11, rmsd=1e-10, maxit=500)
    assert r["converged"] and o["converged"]
    assert abs(r["energy"] - o["energy"]) <= 1e-8
    assert np.max(np.abs(r["P"] - o["P"])) <= 1e-7
    assert ctx.solve_counters()["sp2_fallbacks"] == 0


def test_sp2_falls_back_on_degenerate_frontier(ctx):
    """A closed-shell O atom starts with 4 electrons in 3 degenerate p orbitals: no gap, so the
    projection cannot converge and that iteration is diagonalised (which picks two of the p
    orbitals and opens a gap for the following iterations), as with the diagonalisation builder."""
    Z = np.array([8])
    X = np.zeros((1, 3))
    a = ctx.scf("PM6", Z, X, settings=_tight("sp2"))
    ca = ctx.solve_counters()
    b = ctx.scf("PM6", Z, X, settings=_tight("diagonalization"))
    assert ca["sp2_fallbacks"] > 0
    assert a["converged"] and b["converged"]
    assert abs(a["energy"] - b["energy"]) <= 1e-10
    assert np.max(np.abs(a["P"] - b["P"])) <= 1e-7


def test_batch_with_fallback_equals_single(ctx, tiles):
    """SP2 in a batch: molecules that need the fallback do not disturb the others."""
    tiles(DEFAULT_SP2_TILES[True])  # one tile table for both forms: compares batching only
    from sparrow_amd import _native
    mols = [(np.array([8]), np.zeros((1, 3)), 0)] + [
        (Z, X * BOHR_PER_ANGSTROM, 0) for Z, X in (gen.random_molecule(n, seed=200 + n) for n in (9, 17))]
    st = _native.settings_struct(1e-10, 1e-9, 300, "diis", 8)
    batch, _ = ctx.scf_batch("PM6", mols, settings=st)
    for (Z, X, c), b in zip(mols, batch):
        s = ctx.scf("PM6", Z, X, c, settings=st)
        assert b["iterations"] == s["iterations"]
        assert b["energy"] == s["energy"]
        assert np.array_equal(b["P"], s["P"])


def _mixed(energy_tol=1e-11, rmsd_tol=1e-10):
    from sparrow_amd import _native
    return _native.settings_struct(energy_tol, rmsd_tol, 300, "diis", 8, precision="mixed")


@pytest.mark.parametrize("method,text,n,seed", [("PM6", ETHANOL, 0, 0), ("PM6", CH3SH, 0, 0), ("MNDO", None, 24, 0),
                                                ("PM6", None, 40, 2), ("PM6", None, 90, 3)])
def test_mixed_precision_matches_oracle(ctx, oracles, method, text, n, seed):
    """Split-BF16 x3 SP2 products until |dE| < 1e-5 Eh, then FP64 to the same tight convergence as
    the FP64 tests: the FP64 refinement iterations bring energy, density and orbital energies to
    the FP64 bars (the mixed-precision bar itself is |dE| <= 1e-6 Eh)."""
    if text is not None:
        Z, X = mol(text)
    elif method == "MNDO":
        Z, X = gen.water_cluster(seed=seed)
        X = X * BOHR_PER_ANGSTROM
    else:
        Z, X = gen.random_molecule(n, seed=seed)
        X = X * BOHR_PER_ANGSTROM
    r = ctx.scf(method, Z, X, settings=_mixed())
    c = ctx.solve_counters()
    o = oracles[method].scf(Z, X, 0, dE=1e-10, rmsd=1e-9, maxit=500)
    dE = abs(r["energy"] - o["energy"])
    dP = np.max(np.abs(r["P"] - o["P"]))
    de = np.max(np.abs(r["eps"] - o["eps"]))
    print(f"mixed {method} {len(Z)} atoms: dE {dE:.2e} max dP {dP:.2e} max d eps {de:.2e}, "
          f"low-precision share {c['sp2_lowp_flops'] / max(c['sp2_flops'], 1):.2f}", flush=True)
    assert r["converged"]
    assert c["sp2_lowp_flops"] > 0 and c["sp2_fallbacks"] == 0
    assert dE <= 1e-8
    assert dP <= 1e-7
    assert de <= 1e-7


def test_mixed_batch_equals_single(ctx, tiles):
    tiles(DEFAULT_SP2_TILES[True])  # one tile table for both forms: compares batching only
    mols = [gen.random_molecule(n, seed=300 + n) for n in (7, 26, 45)]
    mols = [(Z, X * BOHR_PER_ANGSTROM, 0) for Z, X in mols]
    st = _mixed(1e-9, 1e-8)
    batch, _ = ctx.scf_batch("PM6", mols, settings=st)
    for (Z, X, c), b in zip(mols, batch):
        s = ctx.scf("PM6", Z, X, c, settings=st)
        assert b["iterations"] == s["iterations"]
        assert b["energy"] == s["energy"]
        assert np.array_equal(b["P"], s["P"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_equals_eager_launches(ctx, precision):
    """Iteration segments replayed from hipGraphs (captured once a launch sequence repeats, with the
    host's checks, eager SP2 chunks and synchronisations between the replays) give trajectories
    bit-identical to eager launches, alone and in a batch whose lists change as molecules converge."""
    from sparrow_amd import _native
    mols = [gen.random_molecule(n, seed=500 + n) for n in (9, 33, 70)]
    mols = [(Z, X * BOHR_PER_ANGSTROM, 0) for Z, X in mols]
    g = _native.settings_struct(1e-10, 1e-9, 300, "diis", 8, precision=precision, graphs=True)
    e = _native.settings_struct(1e-10, 1e-9, 300, "diis", 8, precision=precision, graphs=False)
    bg, _ = ctx.scf_batch("PM6", mols, settings=g)
    cg = ctx.solve_counters()
    be, _ = ctx.scf_batch("PM6", mols, settings=e)
    ce = ctx.solve_counters()
    print(f"batch: {cg['graph_launches']} graph launches, {cg['graph_captures']} captures", flush=True)
    assert cg["graph_launches"] > 0 and ce["graph_launches"] == 0 and ce["graph_captures"] == 0
    for a, b in zip(bg, be):
        assert a["converged"] and a["iterations"] == b["iterations"]
        assert a["energy"] == b["energy"]
        assert np.array_equal(a["P"], b["P"])
    Z, X, c = mols[2]
    sg = ctx.scf("PM6", Z, X, c, settings=g)
    cs = ctx.solve_counters()
    se = ctx.scf("PM6", Z, X, c, settings=e)
    print(f"single: {sg['iterations']} iterations, {cs['graph_launches']} graph launches, "
          f"{cs['graph_captures']} captures", flush=True)
    assert cs["graph_launches"] >= sg["iterations"]  # most iterations replay both segments
    assert sg["iterations"] == se["iterations"] and sg["energy"] == se["energy"]
    assert np.array_equal(sg["P"], se["P"])


DEFAULT_SP2_TILES = {False: [(2047, 32), (1 << 30, 128)], True: [(511, 32), (1 << 30, 128)]}


def _context_tiles():
    """The tables a context starts with: the tuned ones of sp2_tiles.json, else the C ABI's."""
    import json
    from sparrow_amd import _native
    if os.path.exists(_native.SP2_TILES):
        with open(_native.SP2_TILES) as f:
            t = json.load(f)
        return {False: [tuple(e) for e in t["single"]], True: [tuple(e) for e in t["batched"]]}
    return DEFAULT_SP2_TILES


@pytest.fixture
def tiles(ctx):
    """Sets one SP2 tile table for both forms (single molecule, batched) for the test and restores
    the context's starting tables afterwards."""
    def both(table):
        ctx.set_sp2_tiles(table, batched=False)
        ctx.set_sp2_tiles(table, batched=True)
    yield both
    for batched, table in _context_tiles().items():
        ctx.set_sp2_tiles(table, batched=batched)


@pytest.mark.parametrize("tile", [64, 128])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n,seed", [(40, 2), (130, 4)])
def test_tiled_sp2_matches_oracle(ctx, oracles, tiles, tile, precision, n, seed):
    """The whole-k tiled SP2 kernels (64- and 128-wide tiles: padding past n, a single diagonal
    tile, off-diagonal tiles and their mirrors) reach the oracle's converged solution at the FP64
    bars, in FP64 and in the mixed form."""
    from sparrow_amd import _native
    tiles([(1 << 30, tile)])
    Z, X = gen.random_molecule(n, seed=seed)
    X = X * BOHR_PER_ANGSTROM
    st = _native.settings_struct(1e-11, 1e-10, 500, "diis", 8, precision=precision)
    r = ctx.scf("PM6", Z, X, settings=st)
    c = ctx.solve_counters()
    o = oracles["PM6"].scf(Z, X, 0, dE=1e-11, rmsd=1e-10, maxit=500)
    dE = abs(r["energy"] - o["energy"])
    dP = np.max(np.abs(r["P"] - o["P"]))
    print(f"tile {tile} {precision} {len(Z)} atoms (n={r['nao']}): dE {dE:.2e} max dP {dP:.2e}, "
          f"{r['iterations']} iterations", flush=True)
    assert r["converged"] and c["sp2_fallbacks"] == 0
    assert (c["sp2_lowp_flops"] > 0) == (precision == "mixed")
    assert dE <= 1e-8
    assert dP <= 1e-7


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_with_mixed_tile_edges_equals_single(ctx, tiles, precision):
    """Molecules of one batch on the three SP2 kernels (32, 64 and 128-wide tiles by size) give the
    results of their own solves bit for bit: the kernel and tiling depend on the molecule only."""
    from sparrow_amd import _native
    tiles([(60, 32), (200, 64), (1 << 30, 128)])
    mols = [gen.random_molecule(n, seed=700 + n) for n in (12, 50, 95)]
    mols = [(Z, X * BOHR_PER_ANGSTROM, 0) for Z, X in mols]
    st = _native.settings_struct(1e-9, 1e-8, 300, "diis", 8, precision=precision)
    batch, _ = ctx.scf_batch("PM6", mols, settings=st)
    ns = [b["nao"] for b in batch]
    assert min(ns) <= 60 and any(60 < v <= 200 for v in ns) and max(ns) > 200, ns
    for (Z, X, c), b in zip(mols, batch):
        s = ctx.scf("PM6", Z, X, c, settings=st)
        assert b["converged"] and b["iterations"] == s["iterations"]
        assert b["energy"] == s["energy"]
        assert np.array_equal(b["P"], s["P"])


def _drive(srv, tickets, arrivals=None):
    """Steps the server until every ticket has finished; arrivals: {step: [(Z, X), ...]} submitted
    at that step (appended to tickets).  Returns {ticket: result} and the largest on-GPU count."""
    out, step, most = {}, 0, 0
    while True:
        for Z, X in (arrivals or {}).get(step, []):
            tickets.append(srv.submit(Z, X))
        on_gpu, queued = srv.step()
        most = max(most, on_gpu)
        for t in tickets:
            if t not in out:
                r = srv.poll(t)
                if r is not None:
                    out[t] = r
        step += 1
        if len(out) == len(tickets) and step > max((arrivals or {0: 0}).keys()):
            return out, most
        assert step < 2000


def test_server_results_equal_alone(ctx, tiles):
    """Continuous batching: molecules admitted while others iterate give the results of their own solve."""
    tiles(DEFAULT_SP2_TILES[True])  # one tile table for both forms: compares batching only
    from sparrow_amd import _native
    st = _native.settings_struct(1e-9, 1e-8, 300, "diis", 8)
    mols = [gen.random_molecule(n, seed=400 + n) for n in (14, 37, 8, 52, 23)]
    mols = [(Z, X * BOHR_PER_ANGSTROM) for Z, X in mols]
    srv = _native.Server(ctx, "PM6", st, hbm_budget_gb=2.0, max_molecules=8)
    tickets = [srv.submit(*mols[0]), srv.submit(*mols[1])]
    arrivals = {3: [mols[2]], 7: [mols[3], mols[4]]}  # arrive while the first ones are on the GPU
    out, most = _drive(srv, tickets, arrivals)
    assert most >= 3
    for t, (Z, X) in zip(tickets, mols):
        a = ctx.scf("PM6", Z, X, settings=st)
        r = out[t]
        assert r["converged"] and r["iterations"] == a["iterations"]
        assert r["energy"] == a["energy"]
        assert np.array_equal(r["P"], a["P"])
    srv.close()


def test_server_admission_by_budget(ctx, tiles):
    """A small budget holds few molecules at a time; the queue drains and results are unchanged."""
    tiles(DEFAULT_SP2_TILES[True])  # one tile table for both forms: compares batching only
    from sparrow_amd import _native
    st = _native.settings_struct(1e-9, 1e-8, 300, "diis", 8, precision="mixed")
    mols = [gen.random_molecule(n, seed=500 + n) for n in (200, 201, 202, 203, 204, 205)]
    mols = [(Z, X * BOHR_PER_ANGSTROM) for Z, X in mols]
    srv = _native.Server(ctx, "PM6", st, hbm_budget_gb=0.25, max_molecules=64)
    tickets = [srv.submit(Z, X) for Z, X in mols]
    out, most = _drive(srv, tickets)
    assert 1 <= most < len(mols)
    for t, (Z, X) in zip(tickets, mols):
        a = ctx.scf("PM6", Z, X, settings=st)
        assert out[t]["energy"] == a["energy"] and np.array_equal(out[t]["P"], a["P"])
    srv.close()


def test_server_warm_start_from_cache(ctx, tiles):
    """A perturbed-geometry repeat starts from the cached device density of its earlier request.
    (Seed 0: warm and cold starts reach the same SCF solution; for some random molecules they
    reach different ones, in the oracle as on the GPU.)"""
    tiles(DEFAULT_SP2_TILES[True])  # one tile table for both forms: compares batching only
    from sparrow_amd import _native
    st = _native.settings_struct(1e-9, 1e-8, 300, "diis", 8)
    Z, X = gen.random_molecule(40, seed=0)
    X = X * BOHR_PER_ANGSTROM
    X2 = X + np.random.default_rng(1).normal(scale=0.05 * BOHR_PER_ANGSTROM, size=X.shape)
    srv = _native.Server(ctx, "PM6", st, hbm_budget_gb=1.0, max_molecules=4)
    t1 = srv.submit(Z, X, warm_key=77)
    out, _ = _drive(srv, [t1])
    t2 = srv.submit(Z, X2, warm_key=77)
    out2, _ = _drive(srv, [t2])
    assert srv.warm_hits() == 1
    cold = ctx.scf("PM6", Z, X2, settings=st)
    warm = ctx.scf("PM6", Z, X2, settings=st, density_guess=out[t1]["P"])
    r = out2[t2]
    print(f"warm-start iterations: server {r['iterations']}, host guess {warm['iterations']}, cold {cold['iterations']}",
          flush=True)
    assert r["converged"] and r["iterations"] < cold["iterations"]
    # same start density as a solve given it from the host: identical trajectory and result
    assert r["iterations"] == warm["iterations"] and r["energy"] == warm["energy"]
    assert np.array_equal(r["P"], warm["P"])
    assert abs(r["energy"] - cold["energy"]) <= 1e-8
    srv.close()


@pytest.mark.parametrize("mixer", ["ediis", "ediis_diis"])
@pytest.mark.parametrize("method,text,n,seed", [("PM6", ETHANOL, 0, 0), ("PM6", None, 40, 2), ("MNDO", None, 24, 0)],
                         ids=["pm6-ethanol", "pm6-random40", "mndo-water8"])
def test_ediis_mixers_match_oracle(ctx, oracles, mixer, method, text, n, seed):
    """The reference's ediis / ediis_diis mixers reach the converged solution of the oracle."""
    from sparrow_amd import _native
    if text is not None:
        Z, X = mol(text)
    elif method == "MNDO":
        Z, X = gen.water_cluster(seed=seed)
        X = X * BOHR_PER_ANGSTROM
    else:
        Z, X = gen.random_molecule(n, seed=seed)
        X = X * BOHR_PER_ANGSTROM
    st = _native.settings_struct(1e-11, 1e-10, 500, mixer, 8)
    r = ctx.scf(method, Z, X, settings=st)
    o = oracles[method].scf(Z, X, 0, dE=1e-11, rmsd=1e-10, maxit=500)
    d = _native.settings_struct(1e-11, 1e-10, 500, "diis", 8)
    rd = ctx.scf(method, Z, X, settings=d)
    print(f"{mixer} {method} {len(Z)}: iterations {r['iterations']} (diis {rd['iterations']}), "
          f"dE {abs(r['energy'] - o['energy']):.2e}", flush=True)
    assert r["converged"]
    assert abs(r["energy"] - o["energy"]) <= 1e-8
    # EDIIS alone converges linearly near the solution, so when consecutive densities differ by
    # less than the criterion the density can still be 1e-6 from the fixed point (the energy,
    # stationary there, is off by its square); EDIIS+DIIS ends on DIIS steps.
    assert np.max(np.abs(r["P"] - o["P"])) <= (1e-6 if mixer == "ediis" else 1e-7)



def test_calculator_reports_gradients(ctx):
    """The Calculator API (src/Sparrow/Python/Tests/CalculatorTest.py:13-24 usage) returns the
    gradients property (natoms x 3, Hartree/bohr) when it is required."""
    import sparrow_amd
    calc = sparrow_amd.get_calculator("PM6")
    calc.set_structure(AtomCollection.from_xyz_string(CH2O))
    calc.settings["self_consistence_criterion"] = 1e-10
    calc.settings["density_rmsd_criterion"] = 1e-9
    calc.set_required_properties(["energy", "gradients"])
    res = calc.calculate()
    Z, X = mol(CH2O)
    r = ctx.scf("PM6", Z, X, settings=_tight())
    assert res.successful_calculation and res.gradients.shape == (4, 3)
    assert np.max(np.abs(res.gradients - ctx.gradients("PM6", Z, X, r["P"]))) <= 1e-7


def test_molden_orbitals_match_reference(ctx):
    """Our sft_1 PM6 orbitals written as Molden match the reference's Molden file: orbital
    energies to 5e-9 Eh and, for orbitals at least 1e-3 Eh from their neighbours, coefficients to
    1e-6 up to the sign of the vector (fixture: tests/golden/molden_sft1_pm6.json)."""
    import json
    import os
    from sparrow_amd import molden
    here = os.path.dirname(os.path.abspath(__file__))
    fx = json.load(open(os.path.join(here, "golden", "molden_sft1_pm6.json")))
    gold = json.load(open(os.path.join(here, "golden", "reference_values.json")))
    m = AtomCollection.from_xyz_string(gold["structures"]["sft_1"])
    r = ctx.scf("PM6", m.Z, m.positions, settings=_tight(), want_orbitals=True)
    out = molden.text("PM6", m.Z, m.positions, r["eps"], r["C"], r["nocc"]).split("\n")
    blocks = []
    for ln in out[out.index("[MO]") + 1:]:
        if ln.startswith("Sym="):
            blocks.append([ln])
        elif ln:
            blocks[-1].append(ln)
    assert len(blocks) == fx["n_mo"]
    eps = r["eps"]
    for mo in fx["mos"]:
        k, ref = mo["index"], mo["lines"]
        ours = blocks[k]
        assert ours[0] == ref[0] and ours[2] == ref[2] and ours[3] == ref[3]
        assert abs(float(ours[1].split()[1]) - float(ref[1].split()[1])) <= 5e-9
        gap = min(abs(eps[k] - eps[j]) for j in (k - 1, k + 1) if 0 <= j < len(eps))
        if gap > 1e-3:
            a = np.array([float(ln.split()[1]) for ln in ours[4:]])
            b = np.array([float(ln.split()[1]) for ln in ref[4:]])
            s = 1.0 if a @ b >= 0 else -1.0
            print(f"MO {k}: gap {gap:.1e}, max |dC| {np.max(np.abs(s * a - b)):.1e}", flush=True)
            assert np.max(np.abs(s * a - b)) <= 1e-6


@pytest.mark.parametrize("method,text", [("PM6", ETHANOL), ("PM6", CH3SH), ("MNDO", CH2O), ("AM1", ETHANOL), ("RM1", CH3SH)])
def test_unrestricted_fock_matches_oracle(ctx, oracles, method, text):
    """Spin-polarised F^alpha, F^beta and E_el for unequal spin densities (the converged density
    split with a random symmetric spin part) against the oracle's restatement of the reference's
    spin branches; with equal spin densities they reduce to the restricted F and energy."""
    Z, X = mol(text)
    o = oracles[method]
    r = ctx.scf(method, Z, X, settings=_tight())
    P = r["P"]
    rng = np.random.default_rng(7)
    D = rng.normal(scale=0.03, size=P.shape)
    D = D + D.T
    Pa, Pb = P / 2 + D, P / 2 - D
    Fa, Fb, E = ctx.fock_unrestricted(method, Z, X, Pa, Pb)
    _, H, _ = ctx.density_independent(method, Z, X)
    Ga, Gb, Eo = o.gmatrix_unrestricted(Z, X, Pa, Pb)
    low = np.tril_indices(len(P))
    print(f"{method} {len(Z)} atoms: max dG {max(np.max(np.abs((Fa - H)[low] - Ga[low])), np.max(np.abs((Fb - H)[low] - Gb[low]))):.1e}, "
          f"dE {abs(E - Eo):.1e}", flush=True)
    assert np.max(np.abs((Fa - H)[low] - Ga[low])) <= 1e-11
    assert np.max(np.abs((Fb - H)[low] - Gb[low])) <= 1e-11
    # (i, j) and (j, i) of a diagonal block are the same sum in different orders
    assert np.max(np.abs(Fa - Fa.T)) <= 1e-13 and np.max(np.abs(Fb - Fb.T)) <= 1e-13
    assert abs(E - Eo) <= 1e-10
    Fr, Er = ctx.fock(method, Z, X, P)
    Fa0, Fb0, E0 = ctx.fock_unrestricted(method, Z, X, P / 2, P / 2)
    assert np.max(np.abs(Fa0 - Fr)) <= 1e-12 and np.max(np.abs(Fb0 - Fr)) <= 1e-12 and abs(E0 - Er) <= 1e-10


def test_unrestricted_boron_matches_reference(ctx):
    """PM6 doublet boron (src/Sparrow/Tests/MethodsTests/Pm6/PM6Test.cpp:211-233, MOPAC values,
    1e-5 eV): a UHF SCF loop of the test (numpy eigh) around the GPU's spin-polarised Fock build
    reproduces the energy and the alpha / beta eigenvalues."""
    Z, X = mol("1\n\nB -0.1058354421 0.1058354421 -0.1587531632\n")
    Pa, Pb = np.diag([1.0, 1.0, 0.0, 0.0]), np.diag([1.0, 0.0, 0.0, 0.0])
    E_old = None
    for _ in range(100):
        Fa, Fb, E = ctx.fock_unrestricted("PM6", Z, X, Pa, Pb)
        wa, Ca = np.linalg.eigh(Fa)
        wb, Cb = np.linalg.eigh(Fb)
        Pa_new, Pb_new = Ca[:, :2] @ Ca[:, :2].T, Cb[:, :1] @ Cb[:, :1].T
        d = max(np.max(np.abs(Pa_new - Pa)), np.max(np.abs(Pb_new - Pb)))
        Pa, Pb = Pa_new, Pb_new
        if E_old is not None and abs(E - E_old) < 1e-12 and d < 1e-10:
            break
        E_old = E
    EV = 27.21138602
    print(f"boron: E {E * EV:.6f} eV, alpha {wa * EV}, beta {wb * EV}", flush=True)
    assert abs(E * EV - (-49.53668)) <= 1e-5
    assert np.max(np.abs(wa * EV - np.array([-11.74716, -5.78067, -0.09376, -0.09376]))) <= 1e-5
    assert np.max(np.abs(wb * EV - np.array([-10.49432, 0.62041, 0.62041, 2.04873]))) <= 1e-5
