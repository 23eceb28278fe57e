"""sparrow_amd: MI355X-native NDDO (MNDO/AM1/PM3/RM1/PM6) restricted SCF.

The compute path is libsparrow_amd.so (HIP kernels for gfx950 behind the C ABI in
include/sparrow_amd.h); this package holds the host-side mirror of the reference calculator API.
"""
from .structure import AtomCollection, BOHR_PER_ANGSTROM, ANGSTROM_PER_BOHR  # noqa: F401
from .calculator import Calculator, get_calculator, State, Results  # noqa: F401

__all__ = ["AtomCollection", "Calculator", "get_calculator", "State", "Results"]
