"""ctypes binding of libsparrow_amd.so (the C ABI declared in include/sparrow_amd.h).

There is no fallback: if the library is missing or no GPU is present, calls raise.
"""
import ctypes
import json
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsparrow_amd.so")
PARAM_DIR = os.path.join(HERE, "params")
SP2_TILES = os.path.join(HERE, "sp2_tiles.json")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class ScfSettings(ctypes.Structure):
    _fields_ = [("self_consistence_criterion", ctypes.c_double), ("density_rmsd_criterion", ctypes.c_double),
                ("max_scf_iterations", ctypes.c_int), ("scf_mixer", ctypes.c_int), ("diis_size", ctypes.c_int),
                ("density_builder", ctypes.c_int), ("precision", ctypes.c_int), ("graph_replay", ctypes.c_int)]


class ScfResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("electronic_energy", ctypes.c_double),
                ("repulsion_energy", ctypes.c_double), ("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("nao", ctypes.c_int), ("nocc", ctypes.c_int)]


# name -> (restype, argtypes); mirrors include/sparrow_amd.h
SIGNATURES = {
    "sa_last_error": (ctypes.c_char_p, []),
    "sa_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_char_p]),
    "sa_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sa_nao": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _ip]),
    "sa_scf": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp, ctypes.c_int,
                              ctypes.POINTER(ScfSettings), _dp, ctypes.POINTER(ScfResult), _dp, _dp, _dp]),
    "sa_scf_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _ip, _dp, _ip,
                                    ctypes.POINTER(ScfSettings), ctypes.POINTER(_dp), ctypes.POINTER(ScfResult),
                                    ctypes.POINTER(_dp), ctypes.POINTER(_dp), ctypes.POINTER(_dp), _dp]),
    "sa_density_independent": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp, _dp, _dp,
                                              _dp]),
    "sa_fock": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp, _dp, _dp, _dp]),
    "sa_mol_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_void_p)]),
    "sa_mol_info": (ctypes.c_int, [ctypes.c_void_p, _ip, _dp, _dp]),
    "sa_mol_density_independent": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "sa_mol_density_guess": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "sa_mol_fock": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "sa_mol_fock_unrestricted": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "sa_mol_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sa_fock_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp, _dp, _dp, _dp]),
    "sa_fock_unrestricted": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _dp,
                                            _dp]),
    "sa_gradients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp, _dp, ctypes.c_double, _dp]),
    "sa_gradients_unrestricted": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp, ctypes.c_int,
                                                 _dp, _dp, ctypes.c_double, _dp]),
    "sa_cis_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)]),
    "sa_cis_create_unrestricted": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp, _dp, _dp, _dp,
                                                  _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                  ctypes.POINTER(ctypes.c_void_p)]),
    "sa_cis_configurations": (ctypes.c_int, [ctypes.c_void_p, _ip]),
    "sa_cis_sigma": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp]),
    "sa_cis_sigma_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, _dp]),
    "sa_cis_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sa_two_center_block": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "sa_eigh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
    "sa_density_projector": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, _ip]),
    "sa_gemm_nt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
    "sa_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "sa_reset_profile": (ctypes.c_int, [ctypes.c_void_p]),
    "sa_kernel_time": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "sa_solve_counters": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]),
    "sa_work_counts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _ip, _dp]),
    "sa_set_sp2_tiles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip, _ip]),
    "sa_server_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ScfSettings), ctypes.c_double,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "sa_server_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sa_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_longlong,
                                        ctypes.POINTER(ctypes.c_longlong)]),
    "sa_server_step": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip]),
    "sa_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _ip, ctypes.POINTER(ScfResult), _dp]),
    "sa_server_warm_hits": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]),
}

_lib = None


def load_library():
    """Load libsparrow_amd.so and declare the ABI.  Loading needs no GPU."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or python -m sparrow_amd.build")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class NativeError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise NativeError(load_library().sa_last_error().decode())


DENSITY_BUILDERS = {"sp2": 0, "diagonalization": 1}


PRECISIONS = {"fp64": 0, "mixed": 1}


def settings_struct(energy_tol=1e-7, rmsd_tol=1e-5, max_iterations=100, mixer="diis", diis_size=8,
                    density_builder="sp2", precision="fp64", graphs=False):
    mixers = {"no_mixer": 0, "none": 0, "diis": 1, "ediis": 2, "ediis_diis": 3}
    if mixer not in mixers:
        raise ValueError(f"unsupported scf_mixer {mixer!r} (supported: no_mixer, diis, ediis, ediis_diis)")
    if density_builder not in DENSITY_BUILDERS:
        raise ValueError(f"unsupported density builder {density_builder!r} (supported: sp2, diagonalization)")
    if precision not in PRECISIONS:
        raise ValueError(f"unsupported precision {precision!r} (supported: fp64, mixed)")
    return ScfSettings(float(energy_tol), float(rmsd_tol), int(max_iterations), mixers[mixer], int(diis_size),
                       DENSITY_BUILDERS[density_builder], PRECISIONS[precision], 1 if graphs else 0)


class Context:
    """One GPU context (stream + device memory); calls on it are serialised by the caller."""

    def __init__(self, device=0, param_dir=PARAM_DIR, apply_tuning=True):
        lib = load_library()
        h = ctypes.c_void_p()
        check(lib.sa_create(ctypes.byref(h), int(device), param_dir.encode()))
        self.h = h
        self.lib = lib
        self.device = int(device)
        self._cis = weakref.WeakSet()  # open CIS evaluators: closed before the context
        self._servers = weakref.WeakSet()  # open servers: closed before the context
        self._molecules = weakref.WeakSet()  # open molecule handles: closed before the context
        if apply_tuning and os.path.exists(SP2_TILES):  # tables written by sparrow_amd.autotune
            with open(SP2_TILES) as f:
                tuned = json.load(f)
            for key, batched in (("single", False), ("batched", True)):
                if key in tuned:
                    self.set_sp2_tiles([tuple(e) for e in tuned[key]], batched=batched)

    def close(self):
        if self.h:
            for ev in list(getattr(self, "_cis", ())):
                ev.close()
            for srv in list(getattr(self, "_servers", ())):
                srv.close()
            for mol in list(getattr(self, "_molecules", ())):
                mol.close()
            self.lib.sa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nao(self, method, Z):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        n = ctypes.c_int(0)
        check(self.lib.sa_nao(self.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), ctypes.byref(n)))
        return n.value

    def scf(self, method, Z, xyz_bohr, charge=0, settings=None, density_guess=None, want_orbitals=False,
            want_eps=True, want_density=True):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(method, Z)
        st = settings or settings_struct()
        P = np.zeros((n, n), order="F") if want_density else None
        eps = np.zeros(n) if want_eps or want_orbitals else None
        C = np.zeros((n, n), order="F") if want_orbitals else None
        g = np.asfortranarray(density_guess, dtype=np.float64) if density_guess is not None else None
        r = ScfResult()
        check(self.lib.sa_scf(self.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(xyz), int(charge),
                              ctypes.byref(st), _dptr(g), ctypes.byref(r), _dptr(P), _dptr(eps), _dptr(C)))
        return _result_dict(r, P, eps, C)

    def scf_batch(self, method, molecules, settings=None, density_guesses=None, want_orbitals=False,
                  want_density=True, want_eps=False):
        """molecules: list of (Z, xyz_bohr, charge).  An empty list gives no results (and no device
        work); an empty structure inside the list is an error, as for a single solve."""
        nm = len(molecules)
        if nm == 0:
            return [], dict.fromkeys(["setup", "fock", "diis", "eigensolve", "density", "total"], 0.0)
        natoms = np.array([len(m[0]) for m in molecules], dtype=np.int32)
        Z = np.ascontiguousarray(np.concatenate([np.asarray(m[0]) for m in molecules]), dtype=np.int32)
        xyz = np.ascontiguousarray(np.concatenate([np.asarray(m[1], dtype=np.float64).reshape(-1, 3)
                                                   for m in molecules]))
        charge = np.array([m[2] if len(m) > 2 else 0 for m in molecules], dtype=np.int32)
        ns = [self.nao(method, m[0]) for m in molecules]
        st = settings or settings_struct()
        Ps = [np.zeros((n, n), order="F") for n in ns] if want_density else None
        es = [np.zeros(n) for n in ns] if want_eps or want_orbitals else None
        Cs = [np.zeros((n, n), order="F") for n in ns] if want_orbitals else None
        arr = lambda lst: (_dp * nm)(*[_dptr(a) for a in lst]) if lst is not None else None
        guesses = None
        if density_guesses is not None:
            gl = [np.asfortranarray(g, dtype=np.float64) if g is not None else None for g in density_guesses]
            guesses = (_dp * nm)(*[_dptr(g) if g is not None else None for g in gl])
        res = (ScfResult * nm)()
        phases = np.zeros(6)
        check(self.lib.sa_scf_batch(self.h, method.encode(), nm, natoms.ctypes.data_as(_ip), Z.ctypes.data_as(_ip),
                                    _dptr(xyz), charge.ctypes.data_as(_ip), ctypes.byref(st), guesses, res, arr(Ps),
                                    arr(es), arr(Cs), _dptr(phases)))
        out = [_result_dict(res[i], Ps[i] if Ps else None, es[i] if es else None, Cs[i] if Cs else None)
               for i in range(nm)]
        return out, dict(zip(["setup", "fock", "diis", "eigensolve", "density", "total"], phases))

    def density_independent(self, method, Z, xyz_bohr):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(method, Z)
        S = np.zeros((n, n), order="F")
        H = np.zeros((n, n), order="F")
        e = np.zeros(1)
        check(self.lib.sa_density_independent(self.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(xyz),
                                              _dptr(S), _dptr(H), _dptr(e)))
        return S, H, float(e[0])

    def fock(self, method, Z, xyz_bohr, P):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(method, Z)
        P = np.asfortranarray(P, dtype=np.float64)
        F = np.zeros((n, n), order="F")
        e = np.zeros(1)
        check(self.lib.sa_fock(self.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(xyz), _dptr(P),
                               _dptr(F), _dptr(e)))
        return F, float(e[0])

    def molecule(self, method, Z, xyz_bohr, charge=0):
        """Molecule handle: the density-independent part computed once on the device (sa_mol_*)."""
        return Molecule(self, method, Z, xyz_bohr, charge)

    def fock_noalloc(self, method, Z, xyz_bohr, P):
        """fock() without device allocations (static device scratch array; see sa_fock_noalloc)."""
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(method, Z)
        P = np.asfortranarray(P, dtype=np.float64)
        if P.shape != (n, n):
            raise ValueError(f"density must be {n} x {n}")
        F = np.zeros((n, n), order="F")
        e = np.zeros(1)
        check(self.lib.sa_fock_noalloc(self.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(xyz), _dptr(P),
                                       _dptr(F), _dptr(e)))
        return F, float(e[0])

    def fock_unrestricted(self, method, Z, xyz_bohr, Pa, Pb):
        """Spin-polarised F^alpha, F^beta (full symmetric) and E_el for spin densities Pa, Pb."""
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(method, Z)
        Pa = np.asfortranarray(Pa, dtype=np.float64)
        Pb = np.asfortranarray(Pb, dtype=np.float64)
        if Pa.shape != (n, n) or Pb.shape != (n, n):
            raise ValueError(f"densities must be {n} x {n}")
        Fa = np.zeros((n, n), order="F")
        Fb = np.zeros((n, n), order="F")
        e = np.zeros(1)
        check(self.lib.sa_fock_unrestricted(self.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(xyz),
                                            _dptr(Pa), _dptr(Pb), _dptr(Fa), _dptr(Fb), _dptr(e)))
        return Fa, Fb, float(e[0])

    def gradients(self, method, Z, xyz_bohr, P, h=1e-4):
        """dE/dR (Hartree/bohr, natoms x 3) at the converged density P: derivatives of the pair
        energies at fixed P, central differences of step h bohr."""
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        P = np.asfortranarray(P, dtype=np.float64)
        n = self.nao(method, Z)
        if P.shape != (n, n):
            raise ValueError(f"density must be {n} x {n}")
        g = np.zeros((len(Z), 3))
        check(self.lib.sa_gradients(self.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(xyz), _dptr(P),
                                    float(h), _dptr(g)))
        return g

    def gradients_unrestricted(self, method, Z, xyz_bohr, Pa, Pb, charge=0, h=1e-4):
        """dE/dR (Hartree/bohr, natoms x 3) of a converged unrestricted solution with spin densities
        Pa, Pb (each spin's exchange at fixed densities)."""
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        Pa = np.asfortranarray(Pa, dtype=np.float64)
        Pb = np.asfortranarray(Pb, dtype=np.float64)
        n = self.nao(method, Z)
        if Pa.shape != (n, n) or Pb.shape != (n, n):
            raise ValueError(f"spin densities must be {n} x {n}")
        g = np.zeros((len(Z), 3))
        check(self.lib.sa_gradients_unrestricted(self.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(xyz),
                                                 int(charge), _dptr(Pa), _dptr(Pb), float(h), _dptr(g)))
        return g

    def cis_evaluator(self, method, Z, xyz_bohr, C, eps, nocc, spin="singlet", c1=1.0, c2=1.0):
        """Device-resident CIS sigma-vector evaluator of one molecule (the reference's
        CISSigmaVectorEvaluator, restricted reference): C, eps the converged orbitals (columns
        ascending in energy, the first nocc occupied), spin "singlet" or "triplet"."""
        return CisEvaluator(self, method, Z, xyz_bohr, C, eps, nocc, spin, c1, c2)

    def cis_evaluator_unrestricted(self, method, Z, xyz_bohr, Ca, Cb, eps_a, eps_b, nocc_a, nocc_b, c1=1.0, c2=1.0):
        """Unrestricted-reference CIS sigma vectors (trial vector: alpha part, then beta part)."""
        return CisEvaluator(self, method, Z, xyz_bohr, (Ca, Cb), (eps_a, eps_b), (nocc_a, nocc_b), "unrestricted", c1,
                            c2)

    def two_center_block(self, method, ZA, ZB, Rab):
        dims = {1: 1, 4: 10, 9: 40}
        dA, dB = dims[self.nao(method, [ZA])], dims[self.nao(method, [ZB])]
        out = np.zeros((dA, dB))
        R = np.ascontiguousarray(Rab, dtype=np.float64)
        check(self.lib.sa_two_center_block(self.h, method.encode(), int(ZA), int(ZB), _dptr(R), _dptr(out)))
        return out

    def eigh(self, A, nvec=None):
        A = np.asfortranarray(A, dtype=np.float64)
        n = A.shape[0]
        nvec = n if nvec is None else nvec
        w = np.zeros(n)
        V = np.zeros((n, nvec), order="F")
        check(self.lib.sa_eigh(self.h, n, nvec, _dptr(A), _dptr(w), _dptr(V)))
        return w, V

    def density_projector(self, F, nocc):
        """(X, converged): the projector onto the nocc lowest eigenvectors of the symmetric F by
        FP64 SP2 on the GPU; converged is False for a degenerate HOMO/LUMO (use eigh then)."""
        F = np.asfortranarray(F, dtype=np.float64)
        n = F.shape[0]
        if F.shape != (n, n):
            raise ValueError("F must be square")
        X = np.zeros((n, n), order="F")
        ok = ctypes.c_int(0)
        check(self.lib.sa_density_projector(self.h, n, int(nocc), _dptr(F), _dptr(X), ctypes.byref(ok)))
        return X, bool(ok.value)

    def set_profiling(self, on=True):
        check(self.lib.sa_set_profiling(self.h, 1 if on else 0))

    def reset_profile(self):
        check(self.lib.sa_reset_profile(self.h))

    def set_sp2_tiles(self, table, batched=True):
        """table: [(nmax, tile), ...] with nmax ascending and tile in {32, 64, 128}; batched selects the
        table of scf_batch and the server (True) or of scf (False)."""
        nmax = np.ascontiguousarray([int(a) for a, _ in table], dtype=np.int32)
        tile = np.ascontiguousarray([int(b) for _, b in table], dtype=np.int32)
        check(self.lib.sa_set_sp2_tiles(self.h, 1 if batched else 0, len(table), nmax.ctypes.data_as(_ip),
                                        tile.ctypes.data_as(_ip)))

    def solve_counters(self):
        """{iterations, sp2_steps, sp2_fallbacks, eigensolves, sp2_flops, sp2_lowp_flops, graph_launches,
        graph_captures} of the last SCF call."""
        out = (ctypes.c_longlong * 8)()
        check(self.lib.sa_solve_counters(self.h, out, 8))
        return dict(zip(["iterations", "sp2_steps", "sp2_fallbacks", "eigensolves", "sp2_flops", "sp2_lowp_flops",
                         "graph_launches", "graph_captures"], list(out)))

    def work_counts(self, method, Z):
        """Algorithmic work of one molecule: {twocenter_flops, gbuild_bytes, pairs, nao}."""
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        out = np.zeros(4)
        check(self.lib.sa_work_counts(self.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(out)))
        return dict(zip(["twocenter_flops", "gbuild_bytes", "pairs", "nao"], out.tolist()))

    def kernel_time(self, name):
        """(total milliseconds, launches) of a profiled kernel group."""
        t = ctypes.c_double(0)
        c = ctypes.c_longlong(0)
        check(self.lib.sa_kernel_time(self.h, name.encode(), ctypes.byref(t), ctypes.byref(c)))
        return t.value, c.value

    def gemm_nt(self, A, B):
        A = np.asfortranarray(A, dtype=np.float64)
        B = np.asfortranarray(B, dtype=np.float64)
        M, K = A.shape
        N = B.shape[0]
        C = np.zeros((M, N), order="F")
        check(self.lib.sa_gemm_nt(self.h, M, N, K, _dptr(A), _dptr(B), _dptr(C)))
        return C


class Molecule:
    """One geometry's density-independent part resident on the GPU (integrals, S, H, E_rep, guess);
    Fock builds reuse it (FockMatrix::calculateDensityIndependentPart once, then
    calculateDensityDependentPart per density)."""

    def __init__(self, ctx, method, Z, xyz_bohr, charge=0):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        self.ctx, self.lib, self.h = ctx, ctx.lib, ctypes.c_void_p()
        check(self.lib.sa_mol_create(ctx.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(xyz), int(charge),
                                     ctypes.byref(self.h)))
        n, ne, er = ctypes.c_int(0), ctypes.c_double(0), ctypes.c_double(0)
        check(self.lib.sa_mol_info(self.h, ctypes.byref(n), ctypes.byref(ne), ctypes.byref(er)))
        self.n, self.n_electrons, self.repulsion_energy = n.value, ne.value, er.value
        ctx._molecules.add(self)

    def _live(self):
        if not self.h or not getattr(self.ctx, "h", None):
            raise NativeError("the molecule handle or its context is closed")

    def _square(self, P):
        P = np.asfortranarray(P, dtype=np.float64)
        if P.shape != (self.n, self.n):
            raise ValueError(f"density must be {self.n} x {self.n}")
        return P

    def density_independent(self):
        self._live()
        S, H, e = np.zeros((self.n, self.n), order="F"), np.zeros((self.n, self.n), order="F"), ctypes.c_double(0)
        check(self.lib.sa_mol_density_independent(self.h, _dptr(S), _dptr(H), ctypes.byref(e)))
        return S, H, e.value

    def density_guess(self):
        self._live()
        P = np.zeros((self.n, self.n), order="F")
        check(self.lib.sa_mol_density_guess(self.h, _dptr(P)))
        return P

    def fock(self, P):
        self._live()
        P = self._square(P)
        F, e = np.zeros((self.n, self.n), order="F"), ctypes.c_double(0)
        check(self.lib.sa_mol_fock(self.h, _dptr(P), _dptr(F), ctypes.byref(e)))
        return F, e.value

    def fock_unrestricted(self, Pa, Pb):
        self._live()
        Pa, Pb = self._square(Pa), self._square(Pb)
        Fa, Fb = np.zeros((self.n, self.n), order="F"), np.zeros((self.n, self.n), order="F")
        e = ctypes.c_double(0)
        check(self.lib.sa_mol_fock_unrestricted(self.h, _dptr(Pa), _dptr(Pb), _dptr(Fa), _dptr(Fb), ctypes.byref(e)))
        return Fa, Fb, e.value

    def close(self):
        if self.h and getattr(self.ctx, "h", None):
            check(self.lib.sa_mol_destroy(self.h))
        self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching SCF server on one context (see sa_server_* in include/sparrow_amd.h)."""

    def __init__(self, ctx, method, settings=None, hbm_budget_gb=16.0, max_molecules=256):
        self.ctx = ctx
        self.lib = ctx.lib
        self.method = method
        h = ctypes.c_void_p()
        st = settings or settings_struct()
        check(self.lib.sa_server_create(ctx.h, method.encode(), ctypes.byref(st), float(hbm_budget_gb) * 2.0 ** 30,
                                        int(max_molecules), ctypes.byref(h)))
        self.h = h
        self._nao = {}
        ctx._servers.add(self)

    def close(self):
        if self.h:
            self.lib.sa_server_destroy(self.h)
            self.h = None

    def _live(self):
        # a server uses its context's engine: after either is closed every call is an error
        if not self.h or not getattr(self.ctx, "h", None):
            raise NativeError("the server or its context is closed")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, Z, xyz_bohr, charge=0, warm_key=-1):
        self._live()
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        t = ctypes.c_longlong(0)
        check(self.lib.sa_server_submit(self.h, len(Z), Z.ctypes.data_as(_ip), _dptr(xyz), int(charge), int(warm_key),
                                        ctypes.byref(t)))
        self._nao[t.value] = self.ctx.nao(self.method, Z)
        return t.value

    def step(self):
        """One admission + iteration; returns (molecules on the GPU, queued requests)."""
        self._live()
        g, q = ctypes.c_int(0), ctypes.c_int(0)
        check(self.lib.sa_server_step(self.h, ctypes.byref(g), ctypes.byref(q)))
        return g.value, q.value

    def poll(self, ticket, want_density=True):
        """Result dict of a finished request, or None."""
        self._live()
        n = self._nao[ticket]
        P = np.zeros((n, n), order="F") if want_density else None
        done = ctypes.c_int(0)
        r = ScfResult()
        check(self.lib.sa_server_poll(self.h, int(ticket), ctypes.byref(done), ctypes.byref(r), _dptr(P)))
        if not done.value:
            return None
        del self._nao[ticket]
        return _result_dict(r, P, None, None)

    def warm_hits(self):
        self._live()
        h = ctypes.c_longlong(0)
        check(self.lib.sa_server_warm_hits(self.h, ctypes.byref(h)))
        return h.value


def _result_dict(r, P, eps, C):
    return dict(energy=r.energy, electronic_energy=r.electronic_energy, repulsion_energy=r.repulsion_energy,
                iterations=r.iterations, converged=bool(r.converged), nao=r.nao, nocc=r.nocc, P=P, eps=eps, C=C)


class CisEvaluator:
    """sigma = (e_a - e_i) x + vec(C_vir^T F C_occ) for trial vectors x (columns of X, nconf =
    nvir * nocc rows, virtual index fastest), F = 2 c1 J(D) - c2 K(D^T) (singlet) or -c2 K(D^T)
    (triplet), D = C_occ x^T C_vir^T (CISSigmaVectorEvaluator.cpp:96-125)."""

    def __init__(self, ctx, method, Z, xyz_bohr, C, eps, nocc, spin, c1, c2):
        if spin not in ("singlet", "triplet", "unrestricted"):
            raise ValueError("spin must be 'singlet', 'triplet' or 'unrestricted'")
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = ctx.nao(method, Z)
        Cs = C if spin == "unrestricted" else (C,)
        es = eps if spin == "unrestricted" else (eps,)
        Cs = [np.asfortranarray(c, dtype=np.float64) for c in Cs]
        es = [np.ascontiguousarray(e, dtype=np.float64) for e in es]
        if any(c.shape != (n, n) for c in Cs) or any(e.shape != (n,) for e in es):
            raise ValueError(f"orbitals must be {n} x {n} with {n} energies")
        self.ctx, self.h = ctx, ctypes.c_void_p()
        if spin == "unrestricted":
            check(ctx.lib.sa_cis_create_unrestricted(ctx.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(xyz),
                                                     _dptr(Cs[0]), _dptr(Cs[1]), _dptr(es[0]), _dptr(es[1]),
                                                     int(nocc[0]), int(nocc[1]), float(c1), float(c2),
                                                     ctypes.byref(self.h)))
        else:
            check(ctx.lib.sa_cis_create(ctx.h, method.encode(), len(Z), Z.ctypes.data_as(_ip), _dptr(xyz), _dptr(Cs[0]),
                                        _dptr(es[0]), int(nocc), 1 if spin == "triplet" else 0, float(c1), float(c2),
                                        ctypes.byref(self.h)))
        nc = ctypes.c_int(0)
        check(ctx.lib.sa_cis_configurations(self.h, ctypes.byref(nc)))
        self.nconf = nc.value
        ctx._cis.add(self)

    def evaluate(self, X):
        X = np.asarray(X, dtype=np.float64)
        one = X.ndim == 1
        X = np.asfortranarray(X.reshape(self.nconf, -1))
        S = np.zeros_like(X, order="F")
        check(self.ctx.lib.sa_cis_sigma(self.h, X.shape[1], _dptr(X), _dptr(S)))
        return S[:, 0] if one else S

    def evaluate_device(self, X, S=None):
        """Trial vectors resident on the GPU: X a DeviceArray of nvec * nconf doubles (vector k at
        k * nconf).  Returns (sigma as a DeviceArray of the same size, GPU ms); S, when given, is
        the output buffer."""
        if X.size == 0 or X.size % self.nconf:
            raise ValueError(f"device buffer size must be a positive multiple of {self.nconf}")
        if S is not None and S.size != X.size:
            raise ValueError(f"output buffer holds {S.size} doubles, the trial vectors {X.size}")
        S = S or DeviceArray(self.ctx, X.size)
        ms = ctypes.c_double(0.0)
        check(self.ctx.lib.sa_cis_sigma_device(self.h, X.size // self.nconf, X.ptr, S.ptr, ctypes.byref(ms)))
        return S, ms.value

    def close(self):
        if self.h and self.ctx.h:
            check(self.ctx.lib.sa_cis_destroy(self.h))
        self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """float64 buffer in device memory, allocated through the HIP runtime the library itself uses
    (plumbing for the *_device entry points)."""

    _hip = None

    @classmethod
    def _runtime(cls):
        if cls._hip is None:
            load_library()
            hip = ctypes.CDLL("libamdhip64.so.7")  # the instance libsparrow_amd.so already loaded
            hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            hip.hipFree.argtypes = [ctypes.c_void_p]
            hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            hip.hipSetDevice.argtypes = [ctypes.c_int]
            cls._hip = hip
        return cls._hip

    def __init__(self, ctx, size, host=None):
        hip = self._runtime()
        self.size, self.ptr = int(size), ctypes.c_void_p()
        if hip.hipSetDevice(ctx.device) or hip.hipMalloc(ctypes.byref(self.ptr), max(self.size, 1) * 8):
            raise RuntimeError("hipMalloc failed")
        if host is not None:
            self.upload(host)

    @classmethod
    def from_numpy(cls, ctx, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        return cls(ctx, a.size, a)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != self.size or self._runtime().hipMemcpy(self.ptr, a.ctypes.data, self.size * 8, 1):
            raise RuntimeError("hipMemcpy to the device failed")

    def numpy(self):
        out = np.empty(self.size)
        if self._runtime().hipMemcpy(out.ctypes.data, self.ptr, self.size * 8, 2):
            raise RuntimeError("hipMemcpy from the device failed")
        return out

    def free(self):
        if self.ptr:
            self._runtime().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
