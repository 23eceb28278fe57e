"""ctypes front end of the FP64 CPU oracle (oracle/nddo_oracle.cpp).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and only as the checker / CPU baseline.  The product path (sparrow_amd) never
imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
PARAM_DIR = os.path.join(os.path.dirname(HERE), "sparrow_amd", "params")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    lib.orc_last_error.restype = ctypes.c_char_p
    lib.orc_load.restype = ctypes.c_void_p
    lib.orc_load.argtypes = [ctypes.c_char_p]
    lib.orc_set_ev_per_hartree.argtypes = [ctypes.c_double]
    lib.orc_get_ev_per_hartree.restype = ctypes.c_double
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_atom_params.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp]
    lib.orc_one_center.restype = ctypes.c_double
    lib.orc_one_center.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5
    lib.orc_charge_separations.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [_dp]
    lib.orc_klopman.argtypes = [ctypes.c_int, _dp, _dp]
    lib.orc_one_center_explicit.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                            _dp]
    lib.orc_local_integral.restype = ctypes.c_double
    lib.orc_local_integral.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp]
    lib.orc_local_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp, _dp]
    lib.orc_rotation.argtypes = [_dp, _dp]
    lib.orc_global_block.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.orc_nao.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
    lib.orc_density_independent.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
    lib.orc_gmatrix.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
    lib.orc_gmatrix_unrestricted.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.orc_cis_fock.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                 ctypes.c_double, _dp]
    lib.orc_density_guess.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp]
    lib.orc_eigh.argtypes = [ctypes.c_int, _dp, _dp, _dp]
    lib.orc_scf.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                            ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def set_ev_per_hartree(v):
    lib().orc_set_ev_per_hartree(float(v))


def set_threads(t):
    lib().orc_set_threads(int(t))


def max_threads():
    return lib().orc_max_threads()


MULTIPOLE_PAIRS = ["sp1", "pd1", "pp2", "sd2", "dd2", "ss0", "pp0", "dd0"]  # order of D[5] (first 5) and rho[8]


def charge_separations(ns, np_, nd, zs, zp, zd=0.0):
    """ChargeSeparationParameter::computeFromExponents: D (bohr) by multipole pair name."""
    D = np.zeros(5)
    _check(lib().orc_charge_separations(int(ns), int(np_), int(nd), float(zs), float(zp), float(zd), _d(D)))
    return dict(zip(MULTIPOLE_PAIRS[:5], D))


def klopman(order, gss, hsp=0.0, D1sp=0.0, hpp=0.0, D2pp=0.0, F0dd=0.0, G1pd=0.0, D1pd=0.0, G2sd=0.0, D2sd=0.0,
            F2dd=0.0, D2dd=0.0):
    """KlopmanParameter::generateUpToS/P/D (order 0/1/2), energies in Hartree: rho (bohr) by name."""
    a = np.array([gss, hsp, D1sp, hpp, D2pp, F0dd, G1pd, D1pd, G2sd, D2sd, F2dd, D2dd], dtype=np.float64)
    rho = np.zeros(8)
    _check(lib().orc_klopman(int(order), _d(a), _d(rho)))
    return dict(zip(MULTIPOLE_PAIRS, rho))


def one_center_explicit(Z, n1c, zint=(0.0, 0.0, 0.0), f0sd=0.0, g2sd=0.0, gss=0.0, gsp=0.0, gpp=0.0, gp2=0.0, hsp=0.0):
    """The n1c (1, 6 or 58) unique one-centre integrals (Hartree) of element Z from explicit inputs."""
    z = np.ascontiguousarray(zint, dtype=np.float64)
    g = np.array([gss, gsp, gpp, gp2, hsp], dtype=np.float64)
    sc = np.zeros(17)
    out = np.zeros(58)
    _check(lib().orc_one_center_explicit(int(Z), int(n1c), _d(z), float(f0sd), float(g2sd), _d(g), _d(sc), _d(out)))
    return out[:n1c]


def local_integral(t1, t2, R, D1, D2, rho1, rho2):
    """Local-frame (t1|t2) (Hartree) of charge distributions t1, t2 (0..39) at distance R (bohr);
    D and rho as 5- and 8-vectors in MULTIPOLE_PAIRS order."""
    d1, d2 = np.ascontiguousarray(D1, dtype=np.float64), np.ascontiguousarray(D2, dtype=np.float64)
    r1, r2 = np.ascontiguousarray(rho1, dtype=np.float64), np.ascontiguousarray(rho2, dtype=np.float64)
    return lib().orc_local_integral(int(t1), int(t2), float(R), _d(d1), _d(d2), _d(r1), _d(r2))


class Oracle:
    """One NDDO parameter set (MNDO, AM1, PM3, RM1 or PM6)."""

    def __init__(self, method="PM6"):
        self.method = method.upper()
        path = os.path.join(PARAM_DIR, self.method.lower() + ".txt")
        self.h = lib().orc_load(path.encode())
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())

    def atom_params(self, Z):
        D = np.zeros(5)
        rho = np.zeros(8)
        oc = np.zeros(58)
        misc = np.zeros(13)
        _check(lib().orc_atom_params(self.h, int(Z), _d(D), _d(rho), _d(oc), _d(misc)))
        keys = ["nao", "l", "core", "pcore", "pcore_specified", "Uss", "Upp", "Udd", "betaS", "betaP", "betaD",
                "alpha", "n1c"]
        return dict(D=D, rho=rho, oneC=oc, **dict(zip(keys, misc)))

    def one_center(self, Z, i, j, k, l):
        return lib().orc_one_center(self.h, int(Z), i, j, k, l)

    def nao(self, Z):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        n = lib().orc_nao(self.h, len(Z), Z.ctypes.data_as(_ip))
        if n < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return n

    def density_independent(self, Z, xyz_bohr):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(Z)
        S = np.zeros((n, n), order="F")
        H = np.zeros((n, n), order="F")
        E = np.zeros(1)
        _check(lib().orc_density_independent(self.h, len(Z), Z.ctypes.data_as(_ip), _d(xyz), _d(S), _d(H), _d(E)))
        return S, H, float(E[0])

    def gmatrix(self, Z, xyz_bohr, P):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(Z)
        P = np.asfortranarray(P, dtype=np.float64)
        G = np.zeros((n, n), order="F")
        E = np.zeros(1)
        _check(lib().orc_gmatrix(self.h, len(Z), Z.ctypes.data_as(_ip), _d(xyz), _d(P), _d(G), _d(E)))
        return G, float(E[0])

    def gmatrix_unrestricted(self, Z, xyz_bohr, Pa, Pb):
        """G^alpha, G^beta (lower triangles) and the unrestricted electronic energy."""
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(Z)
        Pa = np.asfortranarray(Pa, dtype=np.float64)
        Pb = np.asfortranarray(Pb, dtype=np.float64)
        Ga = np.zeros((n, n), order="F")
        Gb = np.zeros((n, n), order="F")
        E = np.zeros(1)
        _check(lib().orc_gmatrix_unrestricted(self.h, len(Z), Z.ctypes.data_as(_ip), _d(xyz), _d(Pa), _d(Pb), _d(Ga),
                                              _d(Gb), _d(E)))
        return Ga, Gb, float(E[0])

    def cis_fock(self, Z, xyz_bohr, DJ, DK, cJ, cK):
        """CIS AO Fock matrices F_k = cJ J(DJ_k) - cK K(DK_k^T) (CISMatrixAOFockBuilder.cpp:171-330) for
        stacks DJ, DK of shape (nvec, n, n) (each matrix indexed [row, col])."""
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(Z)
        DJ = np.asarray(DJ, dtype=np.float64).reshape(-1, n, n)
        DK = np.asarray(DK, dtype=np.float64).reshape(-1, n, n)
        nv = DJ.shape[0]
        dj = np.ascontiguousarray(DJ.transpose(0, 2, 1))  # column-major per matrix
        dk = np.ascontiguousarray(DK.transpose(0, 2, 1))
        F = np.zeros((nv, n, n))
        _check(lib().orc_cis_fock(self.h, len(Z), Z.ctypes.data_as(_ip), _d(xyz), nv, _d(dj), _d(dk), float(cJ),
                                  float(cK), _d(F)))
        return F.transpose(0, 2, 1).copy()

    def cis_sigma(self, Z, xyz_bohr, C, eps, nocc, X, spin="singlet", c1=1.0, c2=1.0):
        """Restricted CIS sigma vectors (CISSigmaVectorEvaluator.cpp:96-125, pseudo-density of
        CISPseudoDensityBuilder.cpp:99-105): for each column x of X (nvir*nocc, virtual index fastest),
        D = C_occ x^T C_vir^T with x viewed as nvir x nocc, F = 2 c1 J(D) - c2 K(D^T) (singlet) or
        -c2 K(D^T) (triplet), sigma = vec(C_vir^T F C_occ) + (eps_a - eps_i) x."""
        C = np.asarray(C, dtype=np.float64)
        eps = np.asarray(eps, dtype=np.float64)
        n = C.shape[0]
        nvir = n - nocc
        X = np.asarray(X, dtype=np.float64).reshape(nvir * nocc, -1)
        Co, Cv = C[:, :nocc], C[:, nocc:]
        D = np.stack([Co @ X[:, v].reshape(nocc, nvir) @ Cv.T for v in range(X.shape[1])])
        cJ = 2.0 * c1 if spin == "singlet" else 0.0
        F = self.cis_fock(Z, xyz_bohr, D, D, cJ, c2)
        dE = (eps[nocc:][:, None] - eps[:nocc][None, :]).reshape(-1, order="F")
        S = np.stack([(Cv.T @ F[v] @ Co).reshape(-1, order="F") for v in range(X.shape[1])], axis=1)
        return S + dE[:, None] * X

    def cis_sigma_unrestricted(self, Z, xyz_bohr, Ca, Cb, eps_a, eps_b, nocc_a, nocc_b, X, c1=1.0, c2=1.0):
        """Unrestricted CIS sigma vectors (CISSigmaVectorEvaluator.cpp:128-165, the unrestricted
        buildFock of CISMatrixAOFockBuilder.cpp:250-320): each column of X is x^alpha
        (nvir_a*nocc_a) then x^beta; F^s = c1 J(D^alpha + D^beta) - c2 K(D^sT),
        sigma^s = vec(C_vir^sT F^s C_occ^s) + (e_a - e_i) x^s."""
        parts = []
        for C, eps, nocc in ((Ca, eps_a, nocc_a), (Cb, eps_b, nocc_b)):
            C = np.asarray(C, dtype=np.float64)
            parts.append((C, np.asarray(eps, dtype=np.float64), nocc, C.shape[0] - nocc))
        na = parts[0][3] * parts[0][2]
        X = np.asarray(X, dtype=np.float64).reshape(na + parts[1][3] * parts[1][2], -1)
        nv = X.shape[1]
        D = []
        for s, (C, eps, nocc, nvir) in enumerate(parts):
            Xs = X[:na] if s == 0 else X[na:]
            D.append(np.stack([C[:, :nocc] @ Xs[:, v].reshape(nocc, nvir) @ C[:, nocc:].T for v in range(nv)]))
        DJ = D[0] + D[1]
        out = []
        for s, (C, eps, nocc, nvir) in enumerate(parts):
            Xs = X[:na] if s == 0 else X[na:]
            F = self.cis_fock(Z, xyz_bohr, DJ, D[s], c1, c2)
            dE = (eps[nocc:][:, None] - eps[:nocc][None, :]).reshape(-1, order="F")
            Ss = np.stack([(C[:, nocc:].T @ F[v] @ C[:, :nocc]).reshape(-1, order="F") for v in range(nv)], axis=1)
            out.append(Ss + dE[:, None] * Xs)
        return np.vstack(out)

    def density_guess(self, Z, xyz_bohr, charge=0):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(Z)
        P = np.zeros((n, n), order="F")
        _check(lib().orc_density_guess(self.h, len(Z), Z.ctypes.data_as(_ip), _d(xyz), int(charge), _d(P)))
        return P

    def scf(self, Z, xyz_bohr, charge=0, dE=1e-10, rmsd=1e-9, maxit=500, mixer="diis", diis_max=8, P_guess=None):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = self.nao(Z)
        P = np.zeros((n, n), order="F")
        C = np.zeros((n, n), order="F")
        eps = np.zeros(n)
        E = np.zeros(3)
        it = ctypes.c_int(0)
        conv = ctypes.c_int(0)
        pg = None
        if P_guess is not None:
            pg = np.asfortranarray(P_guess, dtype=np.float64)
        _check(lib().orc_scf(self.h, len(Z), Z.ctypes.data_as(_ip), _d(xyz), int(charge), float(dE), float(rmsd),
                             int(maxit), 1 if mixer == "diis" else 0, int(diis_max),
                             _d(pg) if pg is not None else None, _d(E), _d(P), _d(eps), _d(C),
                             ctypes.byref(it), ctypes.byref(conv)))
        return dict(energy=E[0], electronic_energy=E[1], repulsion_energy=E[2], P=P, eps=eps, C=C,
                    iterations=it.value, converged=bool(conv.value))


def local_integral(t1, t2, R, D1, D2, r1, r2):
    D1, D2, r1, r2 = (np.ascontiguousarray(v, dtype=np.float64) for v in (D1, D2, r1, r2))
    return lib().orc_local_integral(int(t1), int(t2), float(R), _d(D1), _d(D2), _d(r1), _d(r2))


def local_matrix(l1, l2, R, D1, D2, r1, r2):
    dims = [1, 10, 40]
    out = np.zeros((dims[l1], dims[l2]))
    D1, D2, r1, r2 = (np.ascontiguousarray(v, dtype=np.float64) for v in (D1, D2, r1, r2))
    _check(lib().orc_local_matrix(int(l1), int(l2), float(R), _d(D1), _d(D2), _d(r1), _d(r2), _d(out)))
    return out


def rotation(Rab):
    Rab = np.ascontiguousarray(Rab, dtype=np.float64)
    out = np.zeros((9, 9))
    lib().orc_rotation(_d(Rab), _d(out))
    return out


def global_block(l1, l2, Rab, D1, D2, r1, r2):
    dims = [1, 10, 40]
    out = np.zeros((dims[l1], dims[l2]))
    Rab = np.ascontiguousarray(Rab, dtype=np.float64)
    D1, D2, r1, r2 = (np.ascontiguousarray(v, dtype=np.float64) for v in (D1, D2, r1, r2))
    _check(lib().orc_global_block(int(l1), int(l2), _d(Rab), _d(D1), _d(D2), _d(r1), _d(r2), _d(out)))
    return out


def eigh(A):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n), order="F")
    _check(lib().orc_eigh(n, _d(A), _d(w), _d(V)))
    return w, V
