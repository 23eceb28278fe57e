"""Calculator front end mirroring the reference's Core::Calculator for the NDDO methods.

Reference surface (src/Sparrow/Sparrow/Implementations/GenericMethodWrapper.h:51-135,
Nddo/NDDOMethodWrapper.cpp:68-103, SparrowModule.cpp:29-49):
  get_calculator("PM6")                ~ ModuleManager::get<Calculator>("PM6")
  calc.settings[...]                   ~ calc.settings() with the same keys and defaults:
      molecular_charge 0, spin_multiplicity 1, spin_mode "restricted",
      self_consistence_criterion 1e-7, density_rmsd_criterion 1e-5,
      max_scf_iterations 100, scf_mixer "diis", method_parameters ""
  calc.set_structure(atoms)            ~ setStructure(AtomCollection)
  calc.modify_positions(pos)           ~ modifyPositions (raises on size mismatch)
  calc.set_required_properties([...])  ~ setRequiredProperties
  calc.calculate()                     ~ calculate() -> Results (energy, successful_calculation,
                                          density_matrix, orbital_energies, coefficient_matrix,
                                          overlap_matrix, one/two_electron_matrix on request)
  calc.get_state() / load_state(s)     ~ getState/loadState (density only, SparrowState.h:32-42)
  calc.clone()                         ~ clone()
Non-convergence is reported through successful_calculation = False, not an exception.
The restricted closed-shell SCF runs entirely on the GPU; spin_mode "unrestricted" (or "any" with
a multiplicity above 1) runs the spin-polarised SCF of sparrow_amd.uhf, whose Fock builds,
eigensolves and products run on the GPU (it diagonalises every iteration whatever the
density_builder, in FP64, with DIIS or no mixer; a loaded state's density starts both spins at
half of it).  Other settings raise.
"""
import copy

import numpy as np

from . import _native
from .structure import AtomCollection

METHODS = ("MNDO", "AM1", "PM3", "RM1", "PM6")
DEFAULT_SETTINGS = {
    "molecular_charge": 0,
    "spin_multiplicity": 1,
    "spin_mode": "restricted",
    "self_consistence_criterion": 1e-7,
    "density_rmsd_criterion": 1e-5,
    "max_scf_iterations": 100,
    "scf_mixer": "diis",
    "method_parameters": "",
    "diis_size": 8,
    # GPU-only keys (no counterpart in the reference): how the density of each iteration is built
    # ("sp2" projection or "diagonalization") and the arithmetic ("fp64" or "mixed")
    "density_builder": "sp2",
    "precision": "fp64",
}
PROPERTIES = ("energy", "gradients", "density_matrix", "orbital_energies", "coefficient_matrix", "overlap_matrix",
              "one_electron_matrix", "two_electron_matrix", "successful_calculation")

_shared_context = {}


def _context(device=0):
    if device not in _shared_context:
        _shared_context[device] = _native.Context(device)
    return _shared_context[device]


class EmptyMolecularStructureError(RuntimeError):
    pass


class IncompatibleStateError(RuntimeError):
    pass


class State:
    """SparrowState: the density matrix and its electron count."""

    def __init__(self, density, n_electrons):
        self.density = np.array(density, copy=True)
        self.n_electrons = n_electrons


class Results(dict):
    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e


class Calculator:
    def __init__(self, method="PM6", device=0):
        m = method.upper()
        if m not in METHODS:
            raise ValueError(f"no NDDO calculator for method {method!r}")
        self._method = m
        self._device = device
        self.settings = dict(DEFAULT_SETTINGS)
        self._structure = None
        self._required = {"energy"}
        self._state_density = None
        self.results = Results()

    def name(self):
        return self._method

    def supports_method_family(self, family):
        return family.upper() == self._method

    def set_structure(self, atoms):
        if not isinstance(atoms, AtomCollection):
            Z, xyz = atoms
            atoms = AtomCollection(Z, xyz)
        self._structure = atoms.copy()
        self._state_density = None

    @property
    def structure(self):
        return self._structure

    def get_positions(self):
        return self._structure.positions.copy()

    def modify_positions(self, positions):
        positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        if self._structure is None or len(positions) != len(self._structure):
            raise RuntimeError("modify_positions: number of positions does not match the structure")
        self._structure.positions = positions.copy()

    def set_required_properties(self, props):
        for p in props:
            if p not in PROPERTIES:
                raise ValueError(f"property {p!r} is not available")
        self._required = set(props) | {"energy"}

    def get_required_properties(self):
        return sorted(self._required)

    @staticmethod
    def possible_properties():
        return list(PROPERTIES)

    def _unrestricted(self):
        s = self.settings
        if s["spin_mode"] == "unrestricted":
            return True
        if s["spin_mode"] not in ("restricted", "any"):
            raise NotImplementedError(f"spin mode {s['spin_mode']!r} is not supported (restricted, unrestricted, any)")
        if s["spin_multiplicity"] != 1:
            if s["spin_mode"] == "restricted":
                raise NotImplementedError("restricted open-shell is not supported; use spin_mode 'unrestricted'")
            return True
        return False

    def _scf_settings(self):
        s = self.settings
        if s["method_parameters"]:
            raise NotImplementedError("custom parameter files are not supported; the built-in sets are used")
        return _native.settings_struct(s["self_consistence_criterion"], s["density_rmsd_criterion"],
                                       s["max_scf_iterations"], s["scf_mixer"], s["diis_size"],
                                       density_builder=s["density_builder"], precision=s["precision"])

    def calculate(self, description=""):
        if self._structure is None or len(self._structure) == 0:
            raise EmptyMolecularStructureError("empty molecular structure")
        unrestricted = self._unrestricted()  # validates the spin settings before the GPU is touched
        st = self._scf_settings()
        ctx = _context(self._device)
        if unrestricted:
            return self._calculate_unrestricted(ctx, description)
        want_C = "coefficient_matrix" in self._required
        # orbital energies cost a final eigensolve: only when asked for
        r = ctx.scf(self._method, self._structure.Z, self._structure.positions, self.settings["molecular_charge"],
                    st, self._state_density, want_orbitals=want_C, want_eps="orbital_energies" in self._required)
        res = Results(description=description, energy=r["energy"], successful_calculation=r["converged"],
                      electronic_energy=r["electronic_energy"], repulsion_energy=r["repulsion_energy"],
                      scf_iterations=r["iterations"])
        if "gradients" in self._required:  # Hartree/bohr, natoms x 3 (Utils::Property::Gradients)
            res["gradients"] = ctx.gradients(self._method, self._structure.Z, self._structure.positions, r["P"])
        if "density_matrix" in self._required:
            res["density_matrix"] = r["P"]
        if "orbital_energies" in self._required:
            res["orbital_energies"] = r["eps"]
        if want_C:
            res["coefficient_matrix"] = r["C"]
        if {"overlap_matrix", "one_electron_matrix"} & self._required:
            S, H, _ = ctx.density_independent(self._method, self._structure.Z, self._structure.positions)
            if "overlap_matrix" in self._required:
                res["overlap_matrix"] = S
            if "one_electron_matrix" in self._required:
                res["one_electron_matrix"] = H
        if "two_electron_matrix" in self._required:
            F, _ = ctx.fock(self._method, self._structure.Z, self._structure.positions, r["P"])
            S, H, _ = ctx.density_independent(self._method, self._structure.Z, self._structure.positions)
            res["two_electron_matrix"] = F - H
        self._last_density = r["P"]
        self._n_electrons = 2 * r["nocc"]
        self.results = res
        return res

    def _calculate_unrestricted(self, ctx, description):
        """Spin-polarised SCF (sparrow_amd.uhf): density_matrix is P^alpha + P^beta, orbital
        energies and coefficients come as {"alpha": ..., "beta": ...}."""
        from .uhf import unrestricted_scf
        unsupported = {"two_electron_matrix"} & self._required
        if unsupported:
            raise NotImplementedError(f"{sorted(unsupported)} are not available for the unrestricted SCF")
        s = self.settings
        if s["scf_mixer"] not in ("diis", "no_mixer", "none"):
            raise NotImplementedError(f"scf_mixer {s['scf_mixer']!r} is not available for the unrestricted SCF "
                                      f"(diis, no_mixer)")
        if s["precision"] != "fp64":
            raise NotImplementedError("the unrestricted SCF runs in FP64 only")
        if s["density_builder"] not in ("sp2", "diagonalization"):
            raise ValueError(f"unsupported density builder {s['density_builder']!r}")
        Z, X = self._structure.Z, self._structure.positions
        diis = 0 if s["scf_mixer"] in ("no_mixer", "none") else s["diis_size"]
        r = unrestricted_scf(ctx, self._method, Z, X, s["molecular_charge"], s["spin_multiplicity"],
                             s["self_consistence_criterion"], s["density_rmsd_criterion"], s["max_scf_iterations"],
                             diis, density_guess=self._state_density)
        res = Results(description=description, energy=r["energy"], successful_calculation=r["converged"],
                      electronic_energy=r["electronic_energy"], repulsion_energy=r["repulsion_energy"],
                      scf_iterations=r["iterations"])
        if "density_matrix" in self._required:
            res["density_matrix"] = r["Pa"] + r["Pb"]
        if "orbital_energies" in self._required:
            res["orbital_energies"] = {"alpha": r["eps_alpha"], "beta": r["eps_beta"]}
        if "coefficient_matrix" in self._required:
            res["coefficient_matrix"] = {"alpha": r["C_alpha"], "beta": r["C_beta"]}
        if "gradients" in self._required:  # Hartree/bohr, natoms x 3
            res["gradients"] = ctx.gradients_unrestricted(self._method, Z, X, r["Pa"], r["Pb"],
                                                          charge=s["molecular_charge"])
        if {"overlap_matrix", "one_electron_matrix"} & self._required:
            S, H, _ = ctx.density_independent(self._method, Z, X)
            if "overlap_matrix" in self._required:
                res["overlap_matrix"] = S
            if "one_electron_matrix" in self._required:
                res["one_electron_matrix"] = H
        self._last_density = r["Pa"] + r["Pb"]
        self._n_electrons = r["n_alpha"] + r["n_beta"]
        self._n_alpha, self._n_beta = r["n_alpha"], r["n_beta"]
        self.results = res
        return res

    def write_molden(self, path):
        """Molden file of the last calculation (MoldenFileGenerator.cpp:44-165); it needs the
        coefficient_matrix and orbital_energies properties among the required ones."""
        from . import molden
        r = self.results
        if "coefficient_matrix" not in r or "orbital_energies" not in r:
            raise RuntimeError("write_molden: require coefficient_matrix and orbital_energies, then calculate()")
        eps, C = r["orbital_energies"], r["coefficient_matrix"]
        nocc = (self._n_alpha, self._n_beta) if isinstance(eps, dict) else self._n_electrons // 2
        molden.write(path, self._method, self._structure.Z, self._structure.positions, eps, C, nocc)

    def get_state(self):
        if getattr(self, "_last_density", None) is None:
            return State(np.zeros((0, 0)), 0)
        return State(self._last_density, self._n_electrons)

    def load_state(self, state):
        if not isinstance(state, State):
            raise TypeError("state is not a SparrowState")
        if state.n_electrons == 0:
            return
        ctx = _context(self._device)
        n = ctx.nao(self._method, self._structure.Z)
        nel = self._electron_count()
        if state.n_electrons != nel or state.density.shape != (n, n):
            raise IncompatibleStateError("the state's electron count does not match the structure")
        self._state_density = np.asfortranarray(state.density)

    def _electron_count(self):
        from .params import core_charge
        return int(sum(core_charge(int(z)) for z in self._structure.Z)) - int(self.settings["molecular_charge"])

    def clone(self):
        c = Calculator(self._method, self._device)
        c.settings = copy.deepcopy(self.settings)
        c._structure = self._structure.copy() if self._structure is not None else None
        c._required = set(self._required)
        c._state_density = None if self._state_density is None else self._state_density.copy()
        return c


def get_calculator(method="PM6", device=0):
    return Calculator(method, device)
