"""Seeded synthetic molecules for the benchmark configurations (all coordinates in Angstrom).

  water_cluster(seed)            8 H2O: O on a 2x2x2 cubic lattice (3.0 A) with U(+-0.2 A) jitter,
                                 O-H 0.96 A, HOH 104.5 deg, uniformly random orientation.
  random_molecule(n, seed)       n-atom C/H/N/O/S molecule: heavy atoms (weights C .6 N .15 O .15
                                 S .1) by a self-avoiding random walk with 1.5 A bonds and a
                                 >= 2.6 A non-bonded heavy-atom distance; every heavy atom
                                 saturated to valence (C4 N3 O2 S2) with H at 1.09 A in random
                                 free directions.  Neutral and closed shell by construction.
  packed_aggregate(k, m, seed)   k molecules of m atoms packed in a cube at 0.8 g/cm^3 with a
                                 >= 2.2 A intermolecular distance.
Everything uses numpy's PCG64 generator, so the same seed gives the same molecule everywhere.
"""
import numpy as np

VALENCE = {6: 4, 7: 3, 8: 2, 16: 2}
HEAVY = np.array([6, 7, 8, 16])
WEIGHTS = np.array([0.6, 0.15, 0.15, 0.1])
MASS = {1: 1.008, 6: 12.011, 7: 14.007, 8: 15.999, 16: 32.06}


def _unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(seed=0):
    rng = np.random.default_rng(seed)
    Z, xyz = [], []
    half = np.deg2rad(104.5) / 2
    local = np.array([[0, 0, 0], [0.96 * np.sin(half), 0, 0.96 * np.cos(half)],
                      [-0.96 * np.sin(half), 0, 0.96 * np.cos(half)]])
    for ix in range(2):
        for iy in range(2):
            for iz in range(2):
                o = np.array([ix, iy, iz], dtype=float) * 3.0 + rng.uniform(-0.2, 0.2, size=3)
                R = _random_rotation(rng)
                for k, zz in enumerate((8, 1, 1)):
                    Z.append(zz)
                    xyz.append(o + R @ local[k])
    return np.array(Z, dtype=np.int32), np.array(xyz)


def _heavy_chain(n_target, rng):
    """Heavy atoms and heavy-heavy bonds such that heavy atoms + saturating H = n_target."""
    pos, el, bonds = [], [], []
    free = []

    def total():
        return len(el) + sum(free)

    def pick(remaining):
        if remaining >= 4:
            return int(rng.choice(HEAVY, p=WEIGHTS))
        # a new atom attached to a saturated site adds (valence - 1) atoms in total
        return {3: 6, 2: 7, 1: int(rng.choice([8, 16], p=[0.6, 0.4]))}[remaining]

    z0 = pick(n_target - 0 if n_target >= 5 else n_target)
    if n_target < 5:  # tiny molecules: one heavy atom plus hydrogens
        z0 = {3: 8, 4: 7, 5: 6}.get(n_target, 6)
    el.append(z0)
    pos.append(np.zeros(3))
    free.append(VALENCE[z0])
    last = 0
    guard = 0
    while total() < n_target:
        guard += 1
        if guard > 100000:
            raise RuntimeError("random walk did not terminate")
        remaining = n_target - total()
        z = pick(remaining)
        if VALENCE[z] - 1 > remaining:
            continue
        parents = [last] if free[last] > 0 else []
        if not parents:
            parents = [i for i in range(len(el)) if free[i] > 0]
            rng.shuffle(parents)
        placed = False
        for par in parents[:20]:
            for _ in range(200):
                p = pos[par] + 1.5 * _unit(rng)
                d = np.linalg.norm(np.array(pos) - p, axis=1)
                d[par] = 10.0
                if d.min() >= 2.6:
                    placed = True
                    break
            if placed:
                pos.append(p)
                el.append(z)
                free.append(VALENCE[z] - 1)
                free[par] -= 1
                bonds.append((par, len(el) - 1))
                last = len(el) - 1
                break
        if not placed:
            last = int(rng.integers(len(el)))
    return np.array(el), np.array(pos), free


def random_molecule(n_atoms, seed=0):
    """Neutral closed-shell C/H/N/O/S molecule with exactly n_atoms atoms (n_atoms >= 3)."""
    if n_atoms < 3:
        raise ValueError("need at least 3 atoms")
    rng = np.random.default_rng(seed)
    el, pos, free = _heavy_chain(n_atoms, rng)
    allpos = list(pos)
    Z = list(el)
    for i in range(len(el)):
        for _ in range(free[i]):
            best, bestd = None, -1.0
            for trial in range(400):
                p = pos[i] + 1.09 * _unit(rng)
                d = np.linalg.norm(np.array(allpos) - p, axis=1)
                d[i] = 10.0
                if d.min() >= 1.7:
                    best = p
                    break
                if d.min() > bestd:
                    best, bestd = p, d.min()
            allpos.append(best)
            Z.append(1)
    Z = np.array(Z, dtype=np.int32)
    assert len(Z) == n_atoms, (len(Z), n_atoms)
    return Z, np.array(allpos)


def packed_aggregate(n_mol=20, atoms_per_mol=100, seed=0, density=0.8):
    """n_mol random molecules packed at `density` g/cm^3 with >= 2.2 A intermolecular minimum."""
    rng = np.random.default_rng(seed)
    mols = [random_molecule(atoms_per_mol, seed=int(rng.integers(1 << 31))) for _ in range(n_mol)]
    mass = sum(MASS[int(z)] for Z, _ in mols for z in Z)  # g/mol
    volume = mass / 0.602214076 / density  # A^3  (1 g/cm^3 = 0.6022 amu/A^3)
    L = volume ** (1.0 / 3.0)
    Zs, Xs = [], []
    for Z, X in mols:
        X = X - X.mean(axis=0)
        for _ in range(2000):
            R = _random_rotation(rng)
            Y = X @ R.T + rng.uniform(0, L, size=3)
            if not Xs:
                break
            prev = np.concatenate(Xs)
            d2 = ((prev[:, None, :] - Y[None, :, :]) ** 2).sum(-1)
            if d2.min() >= 2.2 ** 2:
                break
        else:
            raise RuntimeError("could not pack aggregate")
        Zs.append(Z)
        Xs.append(Y)
    return np.concatenate(Zs), np.concatenate(Xs)


def size_mix(n_mol=256, lo=10, hi=300, seed=0):
    """n_mol molecules with atom counts uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, size=n_mol)
    return [random_molecule(int(s), seed=int(rng.integers(1 << 31))) for s in sizes]


def serving_requests(n_req=300, rate=50.0, lo=10, hi=500, repeat_frac=0.3, sigma=0.05, seed=0):
    """Request stream of the serving workload: Poisson arrivals at `rate` per second, atom counts
    log-uniform in [lo, hi]; a fraction `repeat_frac` repeats an earlier molecule (its family)
    with N(0, sigma) A displacements of every coordinate.  Returns a list of dicts with arrival
    time (s), Z, xyz (A) and the family id (the warm-start key)."""
    rng = np.random.default_rng(seed)
    t, out, families = 0.0, [], []
    for k in range(n_req):
        t += rng.exponential(1.0 / rate)
        if families and rng.random() < repeat_frac:
            fam = int(rng.integers(len(families)))
            Z, X = families[fam]
            X = X + rng.normal(scale=sigma, size=X.shape)
        else:
            n = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
            Z, X = random_molecule(max(n, 3), seed=int(rng.integers(1 << 31)))
            fam = len(families)
            families.append((Z, X))
        out.append({"arrival": t, "Z": Z, "xyz": X, "family": fam})
    return out
