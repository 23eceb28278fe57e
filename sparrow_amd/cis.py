"""Lowest CIS excitation energies by a block Davidson iteration around the GPU's sigma vectors.

The reference drives its CIS sigma-vector evaluator with the Davidson diagonalizer of
scine_utilities (CISLinearResponseTimeDependentCalculator::solve,
src/Sparrow/Sparrow/Implementations/Nddo/TimeDependent/LinearResponse/CISLinearResponseTimeDependentCalculator.cpp:126-160),
preconditioned by the orbital-energy differences
(src/Sparrow/Sparrow/Implementations/TimeDependent/DiagonalPreconditionerEvaluator.cpp:50-65:
r_i / (lambda - de_i), left unchanged where |lambda - de_i| < 1e-3).  That diagonalizer is not
part of the reference's sources; this is the textbook block Davidson method (E. R. Davidson,
J. Comput. Phys. 17, 87 (1975)) with the same preconditioner, guess vectors on the lowest
orbital-energy differences and collapse to the current Ritz vectors when the subspace is full.
The subspace algebra runs on the host (it is small); every sigma vector is computed on the GPU.
"""
import numpy as np

from . import _native


def _expansion(V, T, R):
    """Orthonormal new directions for the subspace V (orthonormal columns) from the preconditioned
    corrections T, column by column: each is orthogonalised against V and the directions already
    taken (repeated while a pass removes more than half of it, at most four passes) and kept if
    what remains is above rounding level (1e-10 of the largest correction, or absolute 1e-10).
    A correction that lies (almost) inside V -- the diagonal preconditioner can map a residual into
    the subspace -- is replaced by its residual R, which is orthogonal to V and so still widens it.
    A column is only ever dropped, never completed by an arbitrary vector, so the result stays
    orthogonal to V."""
    scale = max(1.0, float(np.max(np.linalg.norm(T, axis=0))) if T.size else 1.0)
    B = V
    out = []
    for c in range(T.shape[1]):
        for cand in (T[:, c], R[:, c]):
            t = cand.copy()
            t0 = np.linalg.norm(t)
            if t0 == 0.0:
                continue
            prev = t0
            for _ in range(4):
                t -= B @ (B.T @ t)
                cur = np.linalg.norm(t)
                if cur > 0.5 * prev:
                    break
                prev = cur
            nt = np.linalg.norm(t)
            if nt > 1e-3 * t0 and nt > 1e-10 * scale:
                q = t / nt
                out.append(q)
                B = np.hstack([B, q[:, None]])
                break
    return np.stack(out, axis=1) if out else np.zeros((V.shape[0], 0))


def davidson(sigma, diag, nroots, tol=1e-5, max_iterations=200, guess_dim=None, max_subspace=None):
    """Lowest nroots eigenpairs of the symmetric operator sigma (maps an n x k block to n x k)
    with diagonal estimate diag.  Converged when every residual norm is below tol.
    Returns (eigenvalues, eigenvectors n x nroots, iterations)."""
    diag = np.asarray(diag, dtype=np.float64)
    n = diag.size
    nroots = int(nroots)
    if not 0 < nroots <= n:
        raise ValueError(f"need 0 < nroots <= {n}")
    k0 = min(n, max(nroots, guess_dim or nroots))
    max_subspace = min(n, max(max_subspace or 8 * nroots + k0, k0 + nroots))
    V = np.zeros((n, k0))
    V[np.argsort(diag, kind="stable")[:k0], np.arange(k0)] = 1.0
    AV = sigma(V)
    for it in range(1, max_iterations + 1):
        H = V.T @ AV
        w, y = np.linalg.eigh(0.5 * (H + H.T))
        w, y = w[:nroots], y[:, :nroots]
        X = V @ y
        R = AV @ y - X * w
        norms = np.linalg.norm(R, axis=0)
        if np.all(norms < tol) or V.shape[1] == n:
            return w, X, it
        T = []
        open_ = np.nonzero(norms >= tol)[0]
        for k in open_:
            d = w[k] - diag
            t = R[:, k].copy()
            far = np.abs(d) >= 1e-3
            t[far] /= d[far]
            T.append(t)
        T = np.stack(T, axis=1)
        if V.shape[1] + T.shape[1] > max_subspace:  # collapse onto the Ritz vectors
            V, AV = X, AV @ y
        Q = _expansion(V, T, R[:, open_])
        if Q.shape[1] == 0:  # no new direction: the residuals are at rounding level above tol
            raise RuntimeError(f"Davidson cannot reduce the residuals {norms} below {tol}: the corrections lie "
                               f"in the subspace (tolerance below the attainable precision?)")
        V = np.hstack([V, Q])
        AV = np.hstack([AV, sigma(Q)])
    raise RuntimeError(f"Davidson did not converge in {max_iterations} iterations (residuals {norms})")


def excitation_energies(ctx, method, Z, xyz_bohr, nroots, spin="singlet", settings=None, tol=1e-5, guess_dim=None,
                        c1=1.0, c2=1.0):
    """Restricted SCF on the GPU (final diagonalization for the orbitals), then the lowest nroots
    CIS excitation energies (Hartree) of the singlet or triplet block.  Returns a dict with
    energies, the eigenvectors (configurations x nroots, virtual index fastest), the Davidson
    iterations and the SCF result."""
    st = settings or _native.settings_struct(energy_tol=1e-8, rmsd_tol=1e-6, max_iterations=300,
                                             density_builder="diagonalization")
    r = ctx.scf(method, Z, xyz_bohr, settings=st, want_orbitals=True)
    if not r["converged"]:
        raise RuntimeError("the reference SCF did not converge")
    C, eps, nocc = r["C"], r["eps"], r["nocc"]
    de = (eps[nocc:][:, None] - eps[:nocc][None, :]).reshape(-1, order="F")
    with ctx.cis_evaluator(method, Z, xyz_bohr, C, eps, nocc, spin, c1, c2) as ev:
        w, X, it = davidson(ev.evaluate, de, nroots, tol=tol, guess_dim=guess_dim)
    return {"energies": w, "vectors": X, "iterations": it, "scf": r}
