"""Host-side logic that needs no GPU: synthetic generators, structures, calculator front end."""
import numpy as np
import pytest

from sparrow_amd import generators as gen
from sparrow_amd.calculator import Calculator, EmptyMolecularStructureError, DEFAULT_SETTINGS
from sparrow_amd.params import core_charge
from sparrow_amd.structure import AtomCollection, BOHR_PER_ANGSTROM


def test_water_cluster_geometry():
    Z, X = gen.water_cluster(seed=3)
    assert len(Z) == 24 and list(Z[:3]) == [8, 1, 1]
    for k in range(8):
        o, h1, h2 = X[3 * k:3 * k + 3]
        assert np.linalg.norm(h1 - o) == pytest.approx(0.96, abs=1e-12)
        cosang = np.dot(h1 - o, h2 - o) / (0.96 * 0.96)
        assert np.degrees(np.arccos(cosang)) == pytest.approx(104.5, abs=1e-9)
    Z2, X2 = gen.water_cluster(seed=3)
    assert np.array_equal(X, X2)


@pytest.mark.parametrize("n", [3, 4, 7, 10, 33, 150, 301])
def test_random_molecule_is_exact_closed_shell(n):
    Z, X = gen.random_molecule(n, seed=n)
    assert len(Z) == n
    assert set(np.unique(Z)) <= {1, 6, 7, 8, 16}
    assert sum(core_charge(int(z)) for z in Z) % 2 == 0
    heavy = X[Z != 1]
    d = np.linalg.norm(heavy[:, None] - heavy[None], axis=-1)
    iu = np.triu_indices(len(heavy), 1)
    # heavy-heavy distances are either a 1.5 A bond or >= 2.6 A
    dd = d[iu]
    assert np.all((np.abs(dd - 1.5) < 1e-9) | (dd >= 2.6 - 1e-9))
    Z2, X2 = gen.random_molecule(n, seed=n)
    assert np.array_equal(Z, Z2) and np.array_equal(X, X2)


def test_packed_aggregate_spacing():
    Z, X = gen.packed_aggregate(3, 20, seed=2)
    assert len(Z) == 60
    for a in range(3):
        for b in range(a + 1, 3):
            A, B = X[20 * a:20 * a + 20], X[20 * b:20 * b + 20]
            assert np.min(np.linalg.norm(A[:, None] - B[None], axis=-1)) >= 2.2


def test_size_mix_range():
    mols = gen.size_mix(12, 10, 30, seed=4)
    assert len(mols) == 12 and all(10 <= len(z) <= 30 for z, _ in mols)


def test_xyz_units():
    m = AtomCollection.from_xyz_string("2\n\nH 0 0 0\nH 0 0 1.0\n")
    assert m.positions[1, 2] == pytest.approx(BOHR_PER_ANGSTROM)
    assert BOHR_PER_ANGSTROM == pytest.approx(1.8897261254578281)


def test_calculator_settings_and_errors():
    c = Calculator("PM6")
    for k, v in DEFAULT_SETTINGS.items():
        assert c.settings[k] == v
    assert c.settings["self_consistence_criterion"] == 1e-7 and c.settings["density_rmsd_criterion"] == 1e-5
    with pytest.raises(EmptyMolecularStructureError):
        c.calculate()
    c.set_structure(AtomCollection([1, 1], [[0, 0, 0], [0, 0, 1.4]]))
    with pytest.raises(RuntimeError):
        c.modify_positions(np.zeros((3, 3)))
    with pytest.raises(ValueError):
        c.set_required_properties(["hessian"])
    c.settings["spin_multiplicity"] = 3
    with pytest.raises(NotImplementedError):
        c.calculate()
    with pytest.raises(ValueError):
        Calculator("DFTB3")
    d = c.clone()
    assert d.settings == c.settings and d is not c
    assert c.supports_method_family("pm6") and c.name() == "PM6"


def test_serving_requests_stream():
    reqs = gen.serving_requests(200, rate=40.0, lo=10, hi=500, repeat_frac=0.3, seed=5)
    t = np.array([r["arrival"] for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert 200 / t[-1] == pytest.approx(40.0, rel=0.25)
    sizes = np.array([len(r["Z"]) for r in reqs])
    assert sizes.min() >= 3 and sizes.max() <= 500
    fams = [r["family"] for r in reqs]
    repeats = sum(1 for k, f in enumerate(fams) if f in fams[:k])
    assert 0.2 < repeats / len(reqs) < 0.4
    # a repeat has its family's atoms, displaced by about sigma
    for k, r in enumerate(reqs):
        first = next(q for q in reqs if q["family"] == r["family"])
        assert np.array_equal(first["Z"], r["Z"])
        if r is not first:
            d = r["xyz"] - first["xyz"]
            assert 0.02 < np.std(d) < 0.08


def test_autotune_table_from_winners():
    """The tuner's per-size winners become {nmax: tile} ranges split at the midpoints."""
    from sparrow_amd.autotune import _table
    assert _table([(150, 32), (300, 32), (500, 128), (750, 128)]) == [(400, 32), (1 << 30, 128)]
    assert _table([(150, 32), (300, 64), (500, 128)]) == [(225, 32), (400, 64), (1 << 30, 128)]
    assert _table([(150, 128)]) == [(1 << 30, 128)]


def test_molden_writer_reproduces_reference_text():
    """Fed the reference's own structure and orbitals, the writer reproduces the reference Molden
    file line for line: header, [Atoms] (bohr, 10 decimals), the PM6 STO-6G [GTO] block, and the
    [MO] blocks of the orbitals kept in the fixture (tests/golden/molden_sft1_pm6.json, extracted
    from src/Sparrow/Tests/AppTests/molden_input_reference.molden)."""
    import json
    import os
    from sparrow_amd import molden
    fx = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "molden_sft1_pm6.json")))
    Z, xyz = [], []
    for ln in fx["atoms"][1:]:
        t = ln.split()
        Z.append(int(t[2]))
        xyz.append([float(v) for v in t[3:6]])
    n = fx["n_mo"]
    eps, C = np.zeros(n), np.zeros((n, n))
    naos = [1 if z == 1 else 4 for z in Z]  # sft_1: H, C, O
    first = np.concatenate([[0], np.cumsum(naos)[:-1]])
    for mo in fx["mos"]:
        k, lines = mo["index"], mo["lines"]
        eps[k] = float(lines[1].split()[1])
        vals = [float(ln.split()[1]) for ln in lines[4:]]
        pos = 0
        for a, na in enumerate(naos):
            for m in range(na):
                C[first[a] + molden.MOLDEN_TO_AO[m], k] = vals[pos]
                pos += 1
    occupied = [mo["index"] for mo in fx["mos"] if mo["lines"][3].split()[1] == "2.0"]
    nocc = min(mo["index"] for mo in fx["mos"] if mo["lines"][3].split()[1] == "0.0")
    assert max(occupied) == nocc - 1  # HOMO and LUMO are both in the fixture
    out = molden.text("PM6", Z, xyz, eps, C, nocc).split("\n")
    i_mo = out.index("[MO]")
    assert out[:i_mo + 1] == fx["header"] + fx["atoms"] + fx["gto"] + fx["mo_header"]
    blocks = []
    for ln in out[i_mo + 1:]:
        if ln.startswith("Sym="):
            blocks.append([ln])
        elif ln:
            blocks[-1].append(ln)
    assert len(blocks) == n
    for mo in fx["mos"]:
        assert blocks[mo["index"]] == mo["lines"], mo["index"]


def test_davidson_finds_the_lowest_roots():
    """Block Davidson with the orbital-energy-difference preconditioner against a dense
    eigensolver on a diagonally dominant symmetric operator (with a degenerate pair and a
    subspace small enough to force collapses)."""
    from sparrow_amd.cis import davidson
    rng = np.random.default_rng(0)
    n = 400
    d = np.sort(rng.uniform(0.2, 3.0, n))
    d[1] = d[0]
    B = rng.standard_normal((n, n)) * 0.01
    A = np.diag(d) + 0.5 * (B + B.T)
    w, X, it = davidson(lambda V: A @ V, d, 6, tol=1e-8, max_subspace=20)
    ref = np.linalg.eigvalsh(A)[:6]
    assert np.max(np.abs(w - ref)) <= 1e-10
    assert np.max(np.abs(A @ X - X * w)) <= 1e-7
    assert np.max(np.abs(X.T @ X - np.eye(6))) <= 1e-10


def test_unrestricted_electron_counts():
    from sparrow_amd.uhf import electron_counts
    assert electron_counts([5], 0, 2) == (2, 1)          # boron doublet
    assert electron_counts([8, 8], 0, 3) == (7, 5)       # triplet O2
    assert electron_counts([6, 1, 1, 1], 0, 2) == (4, 3)  # methyl radical
    with pytest.raises(ValueError):
        electron_counts([5], 0, 1)                       # odd electron count as a singlet
    with pytest.raises(ValueError):
        electron_counts([1], 0, 3)


def test_davidson_expansion_stays_orthogonal_to_the_subspace():
    """New directions are orthonormal and orthogonal to the subspace: a correction inside the
    subspace is replaced by its residual, a correction that duplicates an earlier one is dropped
    (not completed by an arbitrary vector), and a generic one is kept after orthogonalisation."""
    from sparrow_amd.cis import _expansion
    rng = np.random.default_rng(2)
    n = 12
    V, _ = np.linalg.qr(rng.standard_normal((n, 4)))
    r = rng.standard_normal(n)
    r -= V @ (V.T @ r)
    t_in = V @ rng.standard_normal(4)                 # correction inside span(V)
    t_out = rng.standard_normal(n)                    # a generic correction
    t_dup = 2.0 * t_out + V @ rng.standard_normal(4)  # the same new direction again
    T = np.stack([t_in, t_out, t_dup], axis=1)
    R = np.stack([r, r, r], axis=1)
    Q = _expansion(V, T, R)
    assert Q.shape[1] == 2
    assert abs(abs(Q[:, 0] @ r) - np.linalg.norm(r)) <= 1e-12   # the residual's direction
    assert np.max(np.abs(V.T @ Q)) <= 1e-14
    assert np.max(np.abs(Q.T @ Q - np.eye(2))) <= 1e-14


def test_davidson_raises_when_residuals_cannot_fall_below_tol():
    """A tolerance below the attainable precision: the corrections fall into the subspace (QR keeps
    no new vector) while the residual norms are still above tol; that is an error, not a result."""
    from sparrow_amd.cis import davidson
    rng = np.random.default_rng(1)
    n = 30
    d = np.linspace(0.5, 3.0, n)
    B = rng.standard_normal((n, n)) * 0.01
    A = np.diag(d) + 0.5 * (B + B.T)
    with pytest.raises(RuntimeError, match="cannot reduce the residuals"):
        davidson(lambda V: A @ V, d, 2, tol=1e-30, max_subspace=6)


def test_server_calls_after_close_raise():
    """A server uses its context's engine: once the server or the context is closed, every call
    raises instead of reaching freed device state (no GPU needed: the check precedes the call)."""
    from sparrow_amd import _native

    class ClosedContext:
        h = None

    srv = object.__new__(_native.Server)
    srv.ctx, srv.h, srv._nao = ClosedContext(), 1, {}
    for call in (lambda: srv.step(), lambda: srv.poll(0), lambda: srv.warm_hits(),
                 lambda: srv.submit([1, 1], np.zeros((2, 3)))):
        with pytest.raises(_native.NativeError, match="closed"):
            call()
    srv.h = None  # (nothing to destroy)


def test_cis_device_output_buffer_size_is_checked():
    from sparrow_amd import _native

    class Buf:
        def __init__(self, size):
            self.size, self.ptr = size, None

    ev = object.__new__(_native.CisEvaluator)
    ev.nconf, ev.h = 10, None
    with pytest.raises(ValueError, match="output buffer"):
        ev.evaluate_device(Buf(20), Buf(10))
    with pytest.raises(ValueError, match="positive multiple"):
        ev.evaluate_device(Buf(15))
    with pytest.raises(ValueError, match="positive multiple"):
        ev.evaluate_device(Buf(0))


def test_molden_unrestricted_blocks():
    """Unrestricted orbitals: an Alpha block then a Beta block, occupied orbitals at 1.0
    (MoldenFileGenerator.cpp:133-148)."""
    from sparrow_amd import molden
    Z, X = [6, 1, 1, 1], np.array([[0.0, 0, 0], [2.0, 0, 0], [-1.0, 1.7, 0], [-1.0, -1.7, 0]])
    n = 7
    rng = np.random.default_rng(0)
    ea, eb = np.sort(rng.normal(size=n)), np.sort(rng.normal(size=n))
    Ca, Cb = rng.normal(size=(n, n)), rng.normal(size=(n, n))
    out = molden.text("PM6", Z, X, {"alpha": ea, "beta": eb}, {"alpha": Ca, "beta": Cb}, (4, 3)).split("\n")
    spins = [ln.split()[1] for ln in out if ln.startswith("Spin=")]
    occ = [ln.split()[1] for ln in out if ln.startswith("Occup=")]
    assert spins == ["Alpha"] * n + ["Beta"] * n
    assert occ == ["1.0"] * 4 + ["0.0"] * 3 + ["1.0"] * 3 + ["0.0"] * 4
    ene = [float(ln.split()[1]) for ln in out if ln.startswith("Ene=")]
    assert np.allclose(ene, np.concatenate([ea, eb]), atol=1e-10)


def test_unrestricted_scf_driver_with_host_stand_ins(oracles):
    """The unrestricted SCF driver (sparrow_amd.uhf: DIIS bookkeeping with cached overlaps, spin
    densities from the projector with the eigenvector fallback, final orbitals) run on the CPU with
    numpy/oracle stand-ins for the device calls reproduces the reference's doublet boron
    (PM6Test.cpp:211-233, MOPAC values, 1e-5 eV) and the methyl radical's fixed point."""
    from sparrow_amd.uhf import _scf, electron_counts
    from sparrow_amd.structure import AtomCollection
    ev = 27.21138602

    def sym(L):
        L = np.tril(L)
        return L + L.T - np.diag(np.diag(L))

    class Ctx:
        def gemm_nt(self, A, B):
            return A @ B.T

        def eigh(self, F):
            return np.linalg.eigh(F)

        def density_projector(self, F, nocc):
            w, C = np.linalg.eigh(F)
            return C[:, :nocc] @ C[:, :nocc].T, bool(w[nocc] - w[nocc - 1] > 1e-6)

    class Mol:
        def __init__(self, o, Z, X):
            self.o, self.Z, self.X = o, Z, X

        def density_independent(self):
            S, H, erep = self.o.density_independent(self.Z, self.X)
            return S, sym(H), erep

        def density_guess(self):
            return self.o.density_guess(self.Z, self.X)

        def fock_unrestricted(self, Pa, Pb):
            Ga, Gb, E = self.o.gmatrix_unrestricted(self.Z, self.X, Pa, Pb)
            _, H, _ = self.density_independent()
            return H + sym(Ga), H + sym(Gb), E

    for text, mult, check in (("1\n\nB -0.1058354421 0.1058354421 -0.1587531632\n", 2, -49.53668),
                              ("4\n\nC 0 0 0\nH 1.08 0 0\nH -0.54 0.935 0\nH -0.54 -0.935 0.05\n", 2, None)):
        m = AtomCollection.from_xyz_string(text)
        na, nb = electron_counts(m.Z, 0, mult)
        r = _scf(Ctx(), Mol(oracles["PM6"], m.Z, m.positions), na, nb, 1e-10, 1e-8, 300, 8, None)
        assert r["converged"]
        if check is not None:
            assert abs(r["energy"] * ev - check) <= 1e-5
        # fixed point: one more Fock build and aufbau per spin reproduce the densities
        Fa, Fb, _ = Mol(oracles["PM6"], m.Z, m.positions).fock_unrestricted(r["Pa"], r["Pb"])
        Ca, Cb = np.linalg.eigh(Fa)[1], np.linalg.eigh(Fb)[1]
        assert np.max(np.abs(Ca[:, :na] @ Ca[:, :na].T - r["Pa"])) <= 1e-6
        assert np.max(np.abs(Cb[:, :nb] @ Cb[:, :nb].T - r["Pb"])) <= 1e-6
