"""sa_fock_noalloc (device buffers from the static scratch array) against sa_fock (allocating
arena): bit-identical F and E_el on two calls in a row, the second with a smaller molecule (the
array keeps the first call's contents), and a clear error for a molecule that does not fit."""
import numpy as np
import pytest

from sparrow_amd import _native
from sparrow_amd import generators as gen
from sparrow_amd.structure import BOHR_PER_ANGSTROM

pytestmark = pytest.mark.gpu


def _density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * 0.05
    return np.asfortranarray(A + A.T + np.eye(n))


def test_fock_noalloc_equals_fock_on_consecutive_calls(ctx):
    cases = [gen.random_molecule(150, seed=0), gen.random_molecule(40, seed=2)]  # second call smaller
    for k, (Z, X) in enumerate(cases):
        X = X * BOHR_PER_ANGSTROM
        P = _density(ctx.nao("PM6", Z), k)
        F0, e0 = ctx.fock("PM6", Z, X, P)
        F1, e1 = ctx.fock_noalloc("PM6", Z, X, P)
        assert np.array_equal(F0, F1) and e0 == e1, k
    # MNDO (sp basis) after PM6 (spd): other elements' one-centre blocks in the array
    Z, X = cases[1]
    X = X * BOHR_PER_ANGSTROM
    P = _density(ctx.nao("MNDO", Z), 3)
    F0, e0 = ctx.fock("MNDO", Z, X, P)
    F1, e1 = ctx.fock_noalloc("MNDO", Z, X, P)
    assert np.array_equal(F0, F1) and e0 == e1


def test_fock_noalloc_rejects_a_molecule_beyond_the_scratch_size(ctx):
    Z, X = gen.random_molecule(600, seed=1)
    X = X * BOHR_PER_ANGSTROM
    n = ctx.nao("PM6", Z)
    with pytest.raises(_native.NativeError, match="static scratch array holds"):
        ctx.fock_noalloc("PM6", Z, X, np.eye(n, order="F"))
    # and the array is still usable afterwards
    Z, X = gen.random_molecule(20, seed=4)
    X = X * BOHR_PER_ANGSTROM
    P = _density(ctx.nao("PM6", Z), 5)
    F0, e0 = ctx.fock("PM6", Z, X, P)
    F1, e1 = ctx.fock_noalloc("PM6", Z, X, P)
    assert np.array_equal(F0, F1) and e0 == e1
