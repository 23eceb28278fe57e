"""CIS sigma vectors on the GPU (sa_cis_*) against the oracle's restatement of
CISMatrixAOFockBuilder / CISSigmaVectorEvaluator, and the excitation energies of the reference's
dense-CIS tests (tests/golden/cis_pm3.json, TimeDependent/CISTestCalculation.cpp, PM3) from the
GPU's own SCF orbitals and sigma vectors.
"""
import json
import os

import numpy as np
import pytest

from sparrow_amd._native import settings_struct
from sparrow_amd.structure import AtomCollection

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CIS = json.load(open(os.path.join(HERE, "golden", "cis_pm3.json")))
EV = 27.21138602  # eV per Hartree (CODATA 2014)
TIGHT = dict(energy_tol=1e-11, rmsd_tol=1e-9, max_iterations=500, density_builder="diagonalization")

CASES = [
    ("PM3", "ss", None),
    ("PM6", None, "4\n\nS 0 0 0\nO 0 0 1.45\nCl 1.7 0 -0.6\nCl -0.85 1.47 -0.6\n"),
    ("MNDO", "ssOxadiazole", None),
    ("AM1", None, "4\n\nN 0 0 0\nH 0.94 0 0.38\nH -0.47 0.81 0.38\nH -0.47 -0.81 0.38\n"),
]


def _structure(name=None, xyz=None):
    m = AtomCollection.from_xyz_string(CIS["structures"][name] if name else xyz)
    return m.Z, m.positions


def _oracle_orbitals(o, Z, X):
    r = o.scf(Z, X, dE=1e-12, rmsd=1e-11, maxit=500)
    assert r["converged"]
    nel = int(round(sum(o.atom_params(z)["core"] for z in Z)))
    return r["C"], r["eps"], nel // 2


@pytest.mark.parametrize("method,name,xyz", CASES, ids=[f"{c[0]}-{c[1] or c[2].split()[1]}" for c in CASES])
@pytest.mark.parametrize("spin", ["singlet", "triplet"])
def test_cis_sigma_matches_oracle(ctx, oracles, method, name, xyz, spin):
    """Random trial vectors, the same orbitals on both sides; relative 1e-11 (FP64 everywhere,
    different summation orders).  The PM6 case (SOCl2) has d shells on S and Cl."""
    Z, X = _structure(name, xyz)
    o = oracles[method]
    C, eps, nocc = _oracle_orbitals(o, Z, X)
    nconf = (len(eps) - nocc) * nocc
    Y = np.random.default_rng(11).standard_normal((nconf, 7))
    print(f"[cis] {method} {name or xyz.split()[1]} {spin} n={len(eps)} nconf={nconf}", flush=True)
    with ctx.cis_evaluator(method, Z, X, C, eps, nocc, spin) as ev:
        assert ev.nconf == nconf
        got = ev.evaluate(Y)
        one = ev.evaluate(Y[:, 3])
    ref = o.cis_sigma(Z, X, C, eps, nocc, Y, spin)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-11 * scale
    assert np.max(np.abs(one - ref[:, 3])) <= 1e-11 * scale


def test_cis_sigma_blocks_and_scaling(ctx, oracles):
    """More vectors than one device block (64), in calls of different sizes on one evaluator,
    and the excited-states scaling c1, c2 of a parameter file."""
    Z, X = _structure("ss")
    o = oracles["PM3"]
    C, eps, nocc = _oracle_orbitals(o, Z, X)
    nconf = (len(eps) - nocc) * nocc
    Y = np.random.default_rng(2).standard_normal((nconf, 150))
    ref = o.cis_sigma(Z, X, C, eps, nocc, Y, "singlet", c1=0.8, c2=1.3)
    with ctx.cis_evaluator("PM3", Z, X, C, eps, nocc, "singlet", c1=0.8, c2=1.3) as ev:
        a = ev.evaluate(Y)
        b = ev.evaluate(Y[:, :5])
        c = ev.evaluate(Y[:, 5:80])
        assert ev.evaluate(np.zeros((nconf, 0))).shape == (nconf, 0)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(a - ref)) <= 1e-11 * scale
    assert np.max(np.abs(b - ref[:, :5])) <= 1e-11 * scale
    assert np.max(np.abs(c - ref[:, 5:80])) <= 1e-11 * scale


def test_cis_sigma_device_buffers(ctx, oracles):
    """sa_cis_sigma_device: trial and sigma vectors in HBM give the host path's sigma vectors bit
    for bit, over more than one block."""
    from sparrow_amd._native import DeviceArray
    Z, X = _structure(None, CASES[1][2])
    o = oracles["PM6"]
    C, eps, nocc = _oracle_orbitals(o, Z, X)
    nconf = (len(eps) - nocc) * nocc
    Y = np.random.default_rng(4).standard_normal((nconf, 70))
    with ctx.cis_evaluator("PM6", Z, X, C, eps, nocc, "triplet") as ev:
        host = ev.evaluate(Y)
        dev, ms = ev.evaluate_device(DeviceArray.from_numpy(ctx, Y.T))
    assert ms > 0.0
    assert np.array_equal(dev.numpy().reshape(70, nconf).T, host)


def test_cis_unrestricted_sigma_matches_oracle(ctx, oracles):
    """Unrestricted reference with different alpha and beta orbitals and occupations (PM6 SOCl2,
    d shells): the beta orbitals are the alpha ones under a random rotation, nocc_alpha = nocc + 1,
    nocc_beta = nocc - 1."""
    Z, X = _structure(None, CASES[1][2])
    o = oracles["PM6"]
    C, eps, nocc = _oracle_orbitals(o, Z, X)
    n = len(eps)
    Q, _ = np.linalg.qr(np.random.default_rng(8).standard_normal((n, n)))
    Cb, eb = C @ Q, np.sort(eps + 0.01 * np.arange(n))
    na, nb = nocc + 1, nocc - 1
    nconf = (n - na) * na + (n - nb) * nb
    Y = np.random.default_rng(9).standard_normal((nconf, 9))
    with ctx.cis_evaluator_unrestricted("PM6", Z, X, C, Cb, eps, eb, na, nb, c1=0.9, c2=1.1) as ev:
        assert ev.nconf == nconf
        got = ev.evaluate(Y)
    ref = o.cis_sigma_unrestricted(Z, X, C, Cb, eps, eb, na, nb, Y, c1=0.9, c2=1.1)
    assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_cis_rejects_bad_input(ctx):
    Z, X = _structure("ssHH")
    C, eps = np.eye(2), np.array([-0.5, 0.2])
    with pytest.raises(RuntimeError):
        ctx.cis_evaluator("PM3", Z, X, C, eps, 2)  # no virtual orbitals
    with pytest.raises(ValueError):
        ctx.cis_evaluator("PM3", Z, X, np.eye(3), eps, 1)


def _gpu_orbitals(ctx, Z, X):
    r = ctx.scf(CIS["method"], Z, X, settings=settings_struct(**TIGHT), want_orbitals=True)
    assert r["converged"]
    return r["C"], r["eps"], r["nocc"]


def _dense_roots(ctx, Z, X, C, eps, nocc, spin):
    if spin == "unrestricted":
        ev = ctx.cis_evaluator_unrestricted(CIS["method"], Z, X, C, C, eps, eps, nocc, nocc)
    else:
        ev = ctx.cis_evaluator(CIS["method"], Z, X, C, eps, nocc, spin)
    with ev:
        A = ev.evaluate(np.eye(ev.nconf))
    assert np.max(np.abs(A - A.T)) <= 1e-12 * np.max(np.abs(A))  # the CIS matrix is symmetric
    return np.linalg.eigvalsh(0.5 * (A + A.T))


@pytest.mark.parametrize("t", CIS["tests"], ids=[t["name"] for t in CIS["tests"]])
def test_cis_excitation_energies_match_reference(ctx, t):
    """Lowest roots of the dense CIS matrix from the GPU's SCF orbitals and sigma vectors against
    the reference's values and tolerances.  An unrestricted reference of these closed-shell
    molecules converges to the restricted orbitals, so the unrestricted evaluator gets
    C_alpha = C_beta."""
    Z, X = _structure(t["structure"])
    print(f"[cis] {t['name']}", flush=True)
    C, eps, nocc = _gpu_orbitals(ctx, Z, X)
    w = _dense_roots(ctx, Z, X, C, eps, nocc, t["spin"])
    got = w[:t["roots"]] * EV
    assert np.max(np.abs(got - np.array(t["energies_ev"]))) <= t["tolerance_ev"], (got, t["energies_ev"])


@pytest.mark.parametrize("t", [t for t in CIS["tests"] if t["spin"] != "unrestricted" and t["roots"] >= 3],
                         ids=[t["name"] for t in CIS["tests"] if t["spin"] != "unrestricted" and t["roots"] >= 3])
def test_cis_davidson_matches_reference(ctx, t):
    """sparrow_amd.cis.excitation_energies (SCF + block Davidson around the GPU's sigma vectors)
    reproduces the reference's lowest roots without forming the CIS matrix."""
    from sparrow_amd.cis import excitation_energies
    Z, X = _structure(t["structure"])
    print(f"[cis] davidson {t['name']}", flush=True)
    r = excitation_energies(ctx, CIS["method"], Z, X, t["roots"], t["spin"], settings=settings_struct(**TIGHT), tol=1e-7)
    got = r["energies"] * EV
    assert np.max(np.abs(got - np.array(t["energies_ev"]))) <= t["tolerance_ev"], (got, t["energies_ev"])


def test_cis_evaluator_closed_with_its_context():
    """Closing a context first closes its open CIS evaluators; the evaluator then refuses work
    instead of touching a destroyed engine."""
    from sparrow_amd import _native
    Z, X = _structure("ssHH")
    c = _native.Context(0)
    r = c.scf("PM3", Z, X, settings=settings_struct(**TIGHT), want_orbitals=True)
    ev = c.cis_evaluator("PM3", Z, X, r["C"], r["eps"], r["nocc"], "singlet")
    assert ev.evaluate(np.ones(1)).shape == (1,)
    c.close()
    assert not ev.h
    with pytest.raises(_native.NativeError):
        ev.evaluate(np.ones(1))
    ev.close()
