// FP64 CPU ORACLE for the NDDO (MNDO / AM1 / PM3 / RM1 / PM6) restricted SCF.
//
// This file is TEST INFRASTRUCTURE.  It restates, in plain serial/OpenMP C++, the reference
// algorithm of qcscine/sparrow for the NDDO SCF hot path, so that the HIP product path can be
// checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it; the product library (sparrow_amd/) never links or calls it.
//
// Citations are file:line into the reference tree (src/Sparrow/Sparrow/Implementations/Nddo/...;
// abbreviations: IE = Utils/IntegralsEvaluationUtils, PU = Utils/ParameterUtils, NU = Utils).
//
// What pins each function (the tests named here exist; CPU suite unless marked):
//   chargeSeparations, klopman, rootVar1/2 ... tests/test_oracle_units.py: the literal inputs and
//                                          values of Tests/Nddo/MultipoleParametersTest.cpp:39-146
//   slaterCondon, oneCenterIntegrals ....... test_oracle_units.py: oneCenterTwoElectronIntegralsTest.cpp
//                                          :27-135 (H, Ga, and the 58 spd integrals of V)
//   localIntegral, localMatrix ............. test_oracle_units.py: LocalIntegralTest.cpp:53-385 (C-O
//                                          10 x 10 block) and :387-531 (H-V, 35 d entries)
//   computeHcore (VuvB) .................... test_oracle_units.py: VuvB = -Z_B (mu nu|s_B s_B) as in
//                                          VuvBMatrixTest.cpp:43-55; tests/test_oracle_golden.py: the
//                                          H cores of PM6Test.cpp:94-168 and MNDOTest.cpp
//   rotationCoefficients, globalBlock ...... tests/test_oracle_invariants.py: rigid-rotation
//                                          invariance of E with S atoms (d rotation), a rotated-frame
//                                          block equals the local block for R along z
//   shellOverlap, atomPairOverlap .......... test_oracle_invariants.py: S(R -> 0) = I; test_oracle_units.py:
//                                          the STO-6G records equal Resources/*/*-STO-6G.basis
//                                          (Sto6gBasisTest.cpp:127-160) and H2 overlap closed form
//   densityGuess ........................... test_oracle_units.py: one atom (diag core/nAO) and H2
//                                          (closed-form contracted-Gaussian overlap)
//   pm6PairRepulsion, mndoPairRepulsion .... test_oracle_golden.py: E_rep of PM6Test.cpp:260-415,
//                                          MNDOTest.cpp, AM1Test.cpp
//   computeG, electronicEnergy, runScf ..... test_oracle_golden.py: converged E_el / E / eigenvalues
//                                          of the method tests, the 70 sft_1 Molden orbital energies,
//                                          the CLI and MOPAC sft energies; test_oracle_invariants.py:
//                                          G linear in P, P^2 = 2P, [F, P] = 0, atom-order invariance
//   computeGUnrestricted ................... tests/test_oracle_cis.py / test_oracle_golden.py: the
//                                          reference's doublet boron (PM6Test.cpp:211-233), the
//                                          restricted limit
//   computeCisFock ......................... tests/test_oracle_cis.py: the PM3 dense-CIS roots of
//                                          TimeDependent/CISTestCalculation.cpp, H2 closed forms
//   symmetricEigen ......................... test_oracle_invariants.py: numpy.linalg.eigh
// The per-iteration SCF trajectory (DIIS subspace, start rule) lives in the absent scine_utilities
// library: "parity unpinned" for the trajectory; only converged results are pinned.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------------------------
// Units.  Utils::Constants (scine_utilities, absent) uses CODATA 2014: a0 = 0.52917721067 A.
// The eV/Hartree factor is settable so the tests can decide between CODATA 2010 / 2014 against
// the reference's molden orbital energies.
// ---------------------------------------------------------------------------------------------
static const double kAngPerBohr = 0.52917721067;
static const double kBohrPerAng = 1.0 / 0.52917721067;
static double gEvPerHartree = 27.21138602;
inline double hartreePerEv() { return 1.0 / gEvPerHartree; }

// ---------------------------------------------------------------------------------------------
// Raw parameter tables (sparrow_amd/params/<method>.txt, converted from Resources/*).
// ---------------------------------------------------------------------------------------------
struct RawAtom {
  int Z = 0;
  double U[3]{}, beta[3]{}, zeta[3]{}, zint[3]{};
  double gss = 0, gpp = 0, gsp = 0, gp2 = 0, hsp = 0, pcore = 0, f0sd = 0, g2sd = 0, alpha = 0;
  std::vector<std::array<double, 3>> gauss;
};
struct Shell {
  int l = 0;
  std::vector<double> e, c;  // exponents (bohr^-2), primitive-normalised coefficients
};
struct ParamSet {
  std::string name;
  bool spd = false, diatomic = false;
  std::map<int, RawAtom> atoms;
  std::map<std::pair<int, int>, std::pair<double, double>> pairs;  // (exponent, factor)
  std::map<int, std::vector<Shell>> gtos;
};

static ParamSet loadParams(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open parameter file " + path);
  ParamSet ps;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    std::string kind;
    is >> kind;
    if (kind == "method") {
      std::string b, basis, d;
      int diat;
      is >> ps.name >> b >> basis >> d >> diat;
      ps.spd = (basis == "spd");
      ps.diatomic = diat != 0;
    } else if (kind == "atom") {
      RawAtom a;
      is >> a.Z;
      for (double& v : a.U) is >> v;
      for (double& v : a.beta) is >> v;
      for (double& v : a.zeta) is >> v;
      for (double& v : a.zint) is >> v;
      is >> a.gss >> a.gpp >> a.gsp >> a.gp2 >> a.hsp >> a.pcore >> a.f0sd >> a.g2sd >> a.alpha;
      int ng;
      is >> ng;
      for (int i = 0; i < ng; ++i) {
        std::array<double, 3> g;
        is >> g[0] >> g[1] >> g[2];
        a.gauss.push_back(g);
      }
      ps.atoms[a.Z] = a;
    } else if (kind == "pair") {
      int z1, z2;
      double e, x;
      is >> z1 >> z2 >> e >> x;
      ps.pairs[{std::min(z1, z2), std::max(z1, z2)}] = {e, x};
    } else if (kind == "gto") {
      int z, l, n;
      is >> z >> l >> n;
      Shell s;
      s.l = l;
      for (int i = 0; i < n; ++i) {
        double e, c;
        is >> e >> c;
        s.e.push_back(e);
        s.c.push_back(c);
      }
      ps.gtos[z].push_back(s);
    }
  }
  return ps;
}

// ---------------------------------------------------------------------------------------------
// Element data: PM6Elements (PU/PrincipalQuantumNumbers.cpp:25-167).
// ---------------------------------------------------------------------------------------------
static int nS(int Z) {  // :25-41
  if (Z <= 2) return 1;
  if (Z <= 9) return 2;
  if (Z <= 17) return 3;
  if (Z <= 35) return 4;
  if (Z <= 53) return 5;
  if (Z <= 83) return 6;
  return 7;
}
static int nP(int Z) {  // :43-58
  if (Z <= 10) return 2;
  if (Z <= 18) return 3;
  if (Z <= 36) return 4;
  if (Z <= 54) return 5;
  if (Z <= 83) return 6;
  return 7;
}
static int nD(int Z) {  // :60-71
  if (Z <= 32) return 3;
  if (Z <= 50) return 4;
  if (Z <= 83) return 5;
  return 6;
}
static int numAOs(int Z, bool spd) {  // :74-124
  if (Z == 1) return 1;
  if (!spd) return 4;
  if (Z <= 12) return 4;
  if (Z <= 17) return 9;
  if (Z <= 20) return 4;
  if (Z <= 29) return 9;
  if (Z <= 32) return 4;
  if (Z <= 33) return 9;
  if (Z <= 34) return 4;
  if (Z <= 35) return 9;
  if (Z <= 38) return 4;
  if (Z <= 47) return 9;
  if (Z <= 50) return 4;
  if (Z <= 51) return 9;
  if (Z <= 52) return 4;
  if (Z <= 53) return 9;
  if (Z <= 56) return 4;
  if (Z <= 79) return 9;
  if (Z <= 83) return 4;
  return 0;
}
static double coreCharge(int Z) {  // :138-163
  if (Z == 10 || Z == 18 || Z == 36 || Z == 54) return 6;
  int core = 0;
  if (Z > 86) core = 86;
  else if (Z > 79) core = 78;
  else if (Z > 70) core = 68;
  else if (Z > 54) core = 54;
  else if (Z > 47) core = 46;
  else if (Z > 36) core = 36;
  else if (Z > 29) core = 28;
  else if (Z > 18) core = 18;
  else if (Z > 10) core = 10;
  else if (Z > 2) core = 2;
  return Z - core;
}

static double fact(int n) {
  double r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}

// MultipolePair order (IE/multipoleTypes.h:25): sp1 pd1 pp2 sd2 dd2 ss0 pp0 dd0
enum MP { sp1 = 0, pd1, pp2, sd2, dd2, ss0, pp0, dd0 };

// ChargeSeparationParameter::A / computeFromExponents (PU/ChargeSeparationParameter.cpp:24-51)
static double chargeSepA(int n1, int n2, double z1, double z2, int L) {
  double f1 = std::pow(2 * z1, n1 + 0.5);
  double f2 = std::pow(2 * z2, n2 + 0.5);
  double f3 = std::pow(z1 + z2, -1.0 * (n1 + n2 + L + 1));
  double f4 = 1.0 / std::sqrt(fact(2 * n1) * fact(2 * n2));
  double f5 = fact(n1 + n2 + L);
  return f1 * f2 * f3 * f4 * f5;
}
static void chargeSeparations(int ns, int np, int nd, double zs, double zp, double zd, bool withD, double D[5]) {
  for (int i = 0; i < 5; ++i) D[i] = 0;
  int n = ns <= np ? ns : np;  // the reference uses min(ns,np) for both (PM6 convention)
  D[sp1] = chargeSepA(n, n, zs, zp, 1) / std::sqrt(3.0);
  D[pp2] = std::sqrt(chargeSepA(n, n, zp, zp, 2)) / std::sqrt(5.0);
  if (withD) {
    D[pd1] = chargeSepA(np, nd, zp, zd, 1) / std::sqrt(5.0);
    D[sd2] = std::sqrt(chargeSepA(ns, nd, zs, zd, 2)) / std::pow(15.0, 0.25);
    D[dd2] = std::sqrt(chargeSepA(nd, nd, zd, zd, 2)) / std::sqrt(7.0);
  }
}

// KlopmanParameter::findRootVar1/2 (PU/KlopmanParameter.cpp:49-92): Newton from 0.001, |dx|<=1e-8
static double rootVar1(double D, double shift) {
  double g = 0.001, prev = 0;
  while (std::abs(g - prev) > 1e-8) {
    double r = g, r2 = r * r, d2 = D * D, sq = std::sqrt(r2 + d2);
    double f = 1.0 / r - 1.0 / sq - shift;
    double der = -1.0 / r2 + r / (sq * sq * sq);
    prev = g;
    g = r - f / der;
  }
  return g;
}
static double rootVar2(double D, double shift) {
  double g = 0.001, prev = 0;
  while (std::abs(g - prev) > 1e-8) {
    double r = g, r2 = r * r, d2 = D * D;
    double sq1 = std::sqrt(r2 + d2), sq2 = std::sqrt(r2 + 2 * d2);
    double f = 1.0 / r - 2.0 / sq1 + 1.0 / sq2 - shift;
    double der = -1.0 / r2 + 2 * r / (sq1 * sq1 * sq1) - r / (sq2 * sq2 * sq2);
    prev = g;
    g = r - f / der;
  }
  return g;
}

// Slater-Condon parameters (PU/SlaterCondonParameters.cpp:43-82) from one-centre Slater
// integrals (IE/oneCenterSlaterIntegral.cpp:38-109).  Order of sc_t: SlaterCondonParameters.h:20-38
enum SC { F0ss, F0pp, F0dd, F0sp, F0sd, F0pd, F2pp, F2dd, F2pd, F4dd, G1sp, G1pd, G2sd, G3pd, R1sppd, R2sdpp, R2sddd };
static double slaterIntegral(int l, const int n[4], const double e[4]) {
  int na = n[0], nb = n[1], nc = n[2], nd = n[3];
  double ea = e[0], eb = e[1], ec = e[2], ed = e[3];
  int n1 = na + nb, n2 = nc + nd;
  double e1 = ea + eb, e2 = ec + ed;
  double x = std::pow(2 * ec, nc + 0.5) * std::pow(2 * ed, nd + 0.5) / std::sqrt(fact(2 * nc) * fact(2 * nd));
  double y = std::pow(2 * ea, na + 0.5) * std::pow(2 * eb, nb + 0.5) / std::sqrt(fact(2 * na) * fact(2 * nb));
  double c = x * y * fact(n2 + l) / std::pow(e2, n2 + l + 1);
  double s1 = fact(n1 - l - 1) / std::pow(e1, n1 - l);
  double s2 = 0;
  for (int ll = 1; ll <= n2 + l + 1; ++ll)
    s2 += std::pow(e2, n2 + l - ll + 1) / fact(n2 + l - ll + 1) * fact(n1 + n2 - ll) / std::pow(e1 + e2, n1 + n2 - ll + 1);
  double s3 = 0;
  for (int ll = 1; ll <= n2 - l; ++ll) {
    double num = std::pow(e2, n2 + l - ll + 1) * fact(n2 - l - 1) * fact(n1 + n2 - ll);
    double den = fact(n2 + l) * fact(n2 - l - ll) * std::pow(e1 + e2, n1 + n2 - ll + 1);
    s3 += num / den;
  }
  return c * (s1 - s2 + s3);
}
static void slaterCondon(int Z, const double zint[3], double f0sd_Eh, double g2sd_Eh, double sc[17]) {
  int nq[3] = {nS(Z), nP(Z), nD(Z)};
  auto U = [&](int l, int a, int b, int c, int d) {
    int ang[4] = {a, b, c, d};
    int n[4];
    double e[4];
    for (int i = 0; i < 4; ++i) {
      n[i] = nq[ang[i]];
      e[i] = zint[ang[i]];
    }
    return slaterIntegral(l, n, e);
  };
  sc[F0ss] = U(0, 0, 0, 0, 0);
  sc[F0pp] = U(0, 1, 1, 1, 1);
  sc[F0dd] = U(0, 2, 2, 2, 2);
  sc[F0sp] = U(0, 0, 0, 1, 1);
  sc[F0sd] = U(0, 0, 0, 2, 2);
  sc[F0pd] = U(0, 1, 1, 2, 2);
  sc[F2pp] = U(2, 1, 1, 1, 1);
  sc[F2dd] = U(2, 2, 2, 2, 2);
  sc[F2pd] = U(2, 1, 1, 2, 2);
  sc[F4dd] = U(4, 2, 2, 2, 2);
  sc[G1sp] = U(1, 0, 1, 1, 0);
  sc[G1pd] = U(1, 1, 2, 2, 1);
  sc[G2sd] = U(2, 0, 2, 2, 0);
  sc[G3pd] = U(3, 1, 2, 2, 1);
  sc[R2sddd] = U(2, 0, 2, 2, 2);
  sc[R2sdpp] = U(2, 0, 2, 1, 1);
  sc[R1sppd] = U(1, 0, 1, 1, 2);
  if (f0sd_Eh != 0) sc[F0sd] = f0sd_Eh;
  if (g2sd_Eh != 0) sc[G2sd] = g2sd_Eh;
}

// ---------------------------------------------------------------------------------------------
// One-centre two-electron integrals: 58 unique (ij|kl) over a 9^4 index table
// (IE/OneCenterTwoElectronCalculator.cpp:31-348).  Orbital order s x y z x2y2 xz z2 yz xy.
// Each unique integral is c1*SC[e1] + c2*SC[e2] + c3*SC[e3].
// ---------------------------------------------------------------------------------------------
enum Orb { s_ = 0, x_, y_, z_, x2y2_, xz_, z2_, yz_, xy_ };
struct OneCExpr {
  double c[3];
  int e[3];
};
static int g1cIndex[9][9][9][9];
static OneCExpr g1cExpr[58];
static bool g1cReady = false;

static void build1cTable() {
  if (g1cReady) return;
  for (auto& a : g1cIndex)
    for (auto& b : a)
      for (auto& c : b)
        for (auto& d : c) d = 100;
  const double r3 = std::sqrt(3.0);
  auto E = [](int k, double c1, int e1, double c2 = 0, int e2 = F0ss, double c3 = 0, int e3 = F0ss) {
    g1cExpr[k] = OneCExpr{{c1, c2, c3}, {e1, e2, e3}};
  };
  auto I = [](int k, std::initializer_list<std::array<int, 4>> lst) {
    for (auto& q : lst) g1cIndex[q[0]][q[1]][q[2]][q[3]] = k;
  };
  // Unique integrals and their Slater-Condon expressions (standard spd one-centre table).
  I(0, {{s_, s_, s_, s_}}); E(0, 1, F0ss);
  I(1, {{s_, s_, z_, z_}, {s_, s_, x_, x_}, {s_, s_, y_, y_}}); E(1, 1, F0sp);
  I(2, {{s_, z_, s_, z_}, {s_, x_, s_, x_}, {s_, y_, s_, y_}}); E(2, 1.0 / 3.0, G1sp);
  I(3, {{z_, z_, z_, z_}, {x_, x_, x_, x_}, {y_, y_, y_, y_}}); E(3, 1, F0pp, 4.0 / 25.0, F2pp);
  I(4, {{x_, x_, z_, z_}, {y_, y_, z_, z_}, {x_, x_, y_, y_}}); E(4, 1, F0pp, -2.0 / 25.0, F2pp);
  I(5, {{x_, z_, x_, z_}, {y_, z_, y_, z_}, {x_, y_, x_, y_}}); E(5, 3.0 / 25.0, F2pp);
  I(6, {{x_, z2_, x_, z2_}, {y_, z2_, y_, z2_}}); E(6, 1.0 / 15.0, G1pd, 18.0 / 245.0, G3pd);
  I(7, {{x_, z2_, x_, x2y2_}, {x_, z2_, y_, xy_}, {x_, xy_, y_, z2_}}); E(7, -r3 / 15.0, G1pd, -std::sqrt(27.0) / 245.0, G3pd);
  I(8, {{x_, x2y2_, y_, xy_}}); E(8, 1.0 / 5.0, G1pd, -21.0 / 245.0, G3pd);
  I(9, {{x_, xy_, y_, x2y2_}}); E(9, -1.0 / 5.0, G1pd, 21.0 / 245.0, G3pd);
  I(10, {{x_, x_, z2_, z2_}, {y_, y_, z2_, z2_}}); E(10, 1, F0pd, -2.0 / 35.0, F2pd);
  I(11, {{x_, x_, z2_, x2y2_}, {x_, y_, z2_, xy_}}); E(11, -std::sqrt(12.0) / 35.0, F2pd);
  I(12, {{y_, z2_, y_, x2y2_}}); E(12, r3 / 15.0, G1pd, std::sqrt(27.0) / 245.0, G3pd);
  I(13, {{z_, z2_, z_, z2_}}); E(13, 4.0 / 15.0, G1pd, 27.0 / 245.0, G3pd);
  I(14, {{z_, x2y2_, z_, x2y2_}, {z_, xy_, z_, xy_}, {x_, yz_, x_, yz_}, {y_, xz_, y_, xz_},
         {x_, xz_, z_, x2y2_}, {x_, yz_, z_, xy_}, {y_, xz_, z_, xy_}, {x_, yz_, y_, xz_}});
  E(14, 3.0 / 49.0, G3pd);
  I(15, {{z_, xz_, z_, xz_}, {z_, yz_, z_, yz_}, {x_, xy_, x_, xy_}, {x_, x2y2_, x_, x2y2_},
         {x_, xz_, x_, xz_}, {y_, xy_, y_, xy_}, {y_, yz_, y_, yz_}, {y_, x2y2_, y_, x2y2_}});
  E(15, 1.0 / 5.0, G1pd, 24.0 / 245.0, G3pd);
  I(16, {{x_, xz_, z_, z2_}, {y_, yz_, z_, z2_}}); E(16, std::sqrt(12.0) / 15.0, G1pd, -std::sqrt(243.0) / 245.0, G3pd);
  I(17, {{x_, x2y2_, z_, xz_}, {x_, xy_, z_, yz_}, {y_, xy_, z_, xz_}, {x_, xz_, y_, yz_}});
  E(17, 1.0 / 5.0, G1pd, -6.0 / 245.0, G3pd);
  I(18, {{y_, yz_, z_, x2y2_}}); E(18, -3.0 / 49.0, G3pd);
  I(19, {{y_, z2_, z_, yz_}, {x_, z2_, z_, xz_}}); E(19, -r3 / 15.0, G1pd, std::sqrt(432.0) / 245.0, G3pd);
  I(20, {{y_, x2y2_, z_, yz_}}); E(20, -1.0 / 5.0, G1pd, 6.0 / 245.0, G3pd);
  I(21, {{z_, z_, z2_, z2_}}); E(21, 1, F0pd, 4.0 / 35.0, F2pd);
  I(22, {{x_, z_, z2_, xz_}, {y_, z_, z2_, yz_}}); E(22, r3 / 35.0, F2pd);
  I(23, {{z_, z_, x2y2_, x2y2_}, {z_, z_, xy_, xy_}, {x_, x_, yz_, yz_}, {y_, y_, xz_, xz_}});
  E(23, 1, F0pd, -4.0 / 35.0, F2pd);
  I(24, {{x_, z_, x2y2_, xz_}, {x_, z_, xy_, yz_}, {y_, z_, xy_, xz_}, {x_, y_, xz_, yz_}}); E(24, 3.0 / 35.0, F2pd);
  I(25, {{y_, z_, x2y2_, yz_}}); E(25, -3.0 / 35.0, F2pd);
  I(26, {{z_, z_, xz_, xz_}, {z_, z_, yz_, yz_}, {x_, x_, xy_, xy_}, {x_, x_, x2y2_, x2y2_},
         {x_, x_, xz_, xz_}, {y_, y_, xy_, xy_}, {y_, y_, yz_, yz_}, {y_, y_, x2y2_, x2y2_}});
  E(26, 1, F0pd, 2.0 / 35.0, F2pd);
  I(27, {{y_, y_, z2_, x2y2_}}); E(27, std::sqrt(12.0) / 35.0, F2pd);
  I(28, {{s_, z2_, s_, z2_}, {s_, x2y2_, s_, x2y2_}, {s_, xy_, s_, xy_}, {s_, xz_, s_, xz_}, {s_, yz_, s_, yz_}});
  E(28, 1.0 / 5.0, G2sd);
  I(29, {{s_, s_, z2_, z2_}, {s_, s_, x2y2_, x2y2_}, {s_, s_, xy_, xy_}, {s_, s_, xz_, xz_}, {s_, s_, yz_, yz_}});
  E(29, 1, F0sd);
  I(30, {{z2_, z2_, z2_, z2_}, {x2y2_, x2y2_, x2y2_, x2y2_}, {xy_, xy_, xy_, xy_}, {xz_, xz_, xz_, xz_}, {yz_, yz_, yz_, yz_}});
  E(30, 1, F0dd, 4.0 / 49.0, F2dd, 36.0 / 441.0, F4dd);
  I(31, {{z2_, x2y2_, z2_, x2y2_}, {z2_, xy_, z2_, xy_}}); E(31, 4.0 / 49.0, F2dd, 15.0 / 441.0, F4dd);
  I(32, {{z2_, xz_, z2_, xz_}, {z2_, yz_, z2_, yz_}}); E(32, 1.0 / 49.0, F2dd, 30.0 / 441.0, F4dd);
  I(33, {{z2_, z2_, x2y2_, x2y2_}, {z2_, z2_, xy_, xy_}}); E(33, 1, F0dd, -4.0 / 49.0, F2dd, 6.0 / 441.0, F4dd);
  I(34, {{z2_, xz_, x2y2_, xz_}, {z2_, xz_, xy_, yz_}, {z2_, yz_, xy_, xz_}});
  E(34, r3 / 49.0, F2dd, -std::sqrt(75.0) / 441.0, F4dd);
  I(35, {{z2_, yz_, x2y2_, yz_}}); E(35, -r3 / 49.0, F2dd, std::sqrt(75.0) / 441.0, F4dd);
  I(36, {{z2_, z2_, xz_, xz_}, {z2_, z2_, yz_, yz_}}); E(36, 1, F0dd, 2.0 / 49.0, F2dd, -24.0 / 441.0, F4dd);
  I(37, {{z2_, x2y2_, xz_, xz_}, {z2_, xy_, xz_, yz_}}); E(37, -std::sqrt(12.0) / 49.0, F2dd, std::sqrt(300.0) / 441.0, F4dd);
  I(38, {{z2_, x2y2_, yz_, yz_}}); E(38, std::sqrt(12.0) / 49.0, F2dd, -std::sqrt(300.0) / 441.0, F4dd);
  I(39, {{x2y2_, xy_, x2y2_, xy_}}); E(39, 35.0 / 441.0, F4dd);
  I(40, {{xy_, xz_, xy_, xz_}, {xy_, yz_, xy_, yz_}, {xz_, yz_, xz_, yz_}, {x2y2_, xz_, x2y2_, xz_}, {x2y2_, yz_, x2y2_, yz_}});
  E(40, 3.0 / 49.0, F2dd, 20.0 / 441.0, F4dd);
  I(41, {{x2y2_, x2y2_, xy_, xy_}}); E(41, 1, F0dd, 4.0 / 49.0, F2dd, -34.0 / 441.0, F4dd);
  I(42, {{x2y2_, xz_, xy_, yz_}}); E(42, 3.0 / 49.0, F2dd, -15.0 / 441.0, F4dd);
  I(43, {{x2y2_, yz_, xy_, xz_}}); E(43, -3.0 / 49.0, F2dd, 15.0 / 441.0, F4dd);
  I(44, {{xy_, xy_, xz_, xz_}, {xy_, xy_, yz_, yz_}, {xz_, xz_, yz_, yz_}, {x2y2_, x2y2_, xz_, xz_}, {x2y2_, x2y2_, yz_, yz_}});
  E(44, 1, F0dd, -2.0 / 49.0, F2dd, -4.0 / 441.0, F4dd);
  const double s245 = std::sqrt(245.0), s125 = std::sqrt(125.0), s45 = std::sqrt(45.0);
  I(45, {{s_, z2_, z2_, z2_}}); E(45, 2.0 / s245, R2sddd);
  I(46, {{s_, x2y2_, z2_, x2y2_}, {s_, xy_, z2_, xy_}, {s_, z2_, x2y2_, x2y2_}, {s_, z2_, xy_, xy_}}); E(46, -2.0 / s245, R2sddd);
  I(47, {{s_, xz_, z2_, xz_}, {s_, yz_, z2_, yz_}, {s_, z2_, xz_, xz_}, {s_, z2_, yz_, yz_}}); E(47, 1.0 / s245, R2sddd);
  I(48, {{s_, xz_, x2y2_, xz_}, {s_, xz_, xy_, yz_}, {s_, yz_, xy_, xz_}, {s_, xy_, xz_, yz_}, {s_, x2y2_, xz_, xz_}});
  E(48, r3 / s245, R2sddd);
  I(49, {{s_, yz_, x2y2_, yz_}, {s_, x2y2_, yz_, yz_}}); E(49, -r3 / s245, R2sddd);
  I(50, {{s_, z2_, x_, x_}, {s_, z2_, y_, y_}}); E(50, -1.0 / s125, R2sdpp);
  I(51, {{s_, y_, y_, x2y2_}}); E(51, -r3 / s45, R1sppd);
  I(52, {{s_, x2y2_, y_, y_}}); E(52, -r3 / s125, R2sdpp);
  I(53, {{s_, z_, z_, z2_}}); E(53, 2.0 / s45, R1sppd);
  I(54, {{s_, z_, x_, xz_}, {s_, x_, z_, xz_}, {s_, z_, y_, yz_}, {s_, y_, z_, yz_}, {s_, x_, x_, x2y2_}, {s_, x_, y_, xy_}, {s_, y_, x_, xy_}});
  E(54, r3 / s45, R1sppd);
  I(55, {{s_, z2_, z_, z_}}); E(55, 2.0 / s125, R2sdpp);
  I(56, {{s_, xz_, x_, z_}, {s_, yz_, y_, z_}, {s_, x2y2_, x_, x_}, {s_, xy_, x_, y_}}); E(56, r3 / s125, R2sdpp);
  I(57, {{s_, x_, x_, z2_}, {s_, y_, y_, z2_}}); E(57, -1.0 / s45, R1sppd);
  // Permutational expansion in the reference's order (OneCenterTwoElectronCalculator.cpp:46-63)
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j)
      for (int k = 0; k < 9; ++k)
        for (int l = 0; l < 9; ++l) {
          int v = g1cIndex[i][j][k][l];
          if (v == 100) continue;
          g1cIndex[j][i][k][l] = v;
          g1cIndex[i][j][l][k] = v;
          g1cIndex[j][i][l][k] = v;
          g1cIndex[k][l][i][j] = v;
          g1cIndex[k][l][j][i] = v;
          g1cIndex[l][k][i][j] = v;
          g1cIndex[l][k][j][i] = v;
        }
  g1cReady = true;
}

// ---------------------------------------------------------------------------------------------
// Runtime atomic parameters (RawParameterProcessor::processAtomicParameters,
// PU/RawParameterProcessor.cpp:67-99; setKlopman :126-159).
// KlopmanParameter::generateUpToS / P / D (PU/KlopmanParameter.cpp:26-47); order 0, 1, 2.
// in = {gss, hsp, D1sp, hpp, D2pp, F0dd, G1pd, D1pd, G2sd, D2sd, F2dd, D2dd} (Hartree, bohr)
static void klopman(int order, const double in[12], double rho[8]) {
  for (int i = 0; i < 8; ++i) rho[i] = 0;
  rho[ss0] = 1.0 / (2.0 * in[0]);
  if (order >= 1) {
    rho[pp0] = rho[ss0];
    rho[sp1] = rootVar1(in[2], 4 * in[1]);
    rho[pp2] = rootVar2(in[4], 8 * in[3]);
  }
  if (order >= 2) {
    rho[dd0] = 1.0 / (2.0 * in[5]);
    rho[pd1] = rootVar1(in[7], 16.0 / 15.0 * in[6]);
    rho[sd2] = rootVar2(in[9], 8.0 / 5.0 * in[8]);
    rho[dd2] = rootVar2(in[11], 24.0 / 49.0 * in[10]);
  }
}

// OneCenterTwoElectronIntegrals::calculateIntegrals (IE/oneCenterTwoElectronIntegrals.cpp:35-87)
// with the values RawParameterProcessor::get1c2eIntegrals sets (PU/RawParameterProcessor.cpp:42-65):
// the non-zero ones of g5 = {(ss|ss), (ss|xx), (xx|xx), (xx|yy), (sx|sx)} (Hartree) are taken as
// given, (xy|xy) = 1/2[(xx|xx) - (xx|yy)] when both are, every other of the n1c (1, 6, 58) unique
// integrals from the Slater-Condon parameters sc.
static void oneCenterIntegrals(int n1c, const double g5[5], const double sc[17], double* out) {
  build1cTable();
  bool given[58] = {};
  for (int i = 0; i < 58; ++i) out[i] = 0;
  auto set = [&](int i, int j, int k, int l, double v) {
    int idx = g1cIndex[i][j][k][l];
    out[idx] = v;
    given[idx] = true;
  };
  if (g5[0] != 0) set(s_, s_, s_, s_, g5[0]);
  if (g5[1] != 0) set(s_, s_, x_, x_, g5[1]);
  if (g5[2] != 0) set(x_, x_, x_, x_, g5[2]);
  if (g5[3] != 0) set(x_, x_, y_, y_, g5[3]);
  if (g5[4] != 0) set(s_, x_, s_, x_, g5[4]);
  if (n1c >= 6 && given[3] && given[4]) set(x_, y_, x_, y_, 0.5 * (out[3] - out[4]));
  for (int i = 0; i < n1c; ++i)
    if (!given[i]) {
      const OneCExpr& e = g1cExpr[i];
      out[i] = e.c[0] * sc[e.e[0]] + e.c[1] * sc[e.e[1]] + e.c[2] * sc[e.e[2]];
    }
}

// ---------------------------------------------------------------------------------------------
struct AtomPar {
  int Z = 0, nao = 0, l = 0;
  double core = 0;
  double U[3]{}, beta[3]{};
  double alpha = 0;          // bohr^-1
  double D[5]{}, rho[8]{};   // bohr
  double pcore = 0;
  bool pcoreSpecified = false;
  std::vector<std::array<double, 3>> gauss;  // (a [eV*bohr], b [bohr^-2], c [bohr])
  int n1c = 0;
  double oneC[58]{};         // Hartree
  std::vector<Shell> shells;
};

struct Method {
  ParamSet raw;
  std::map<int, AtomPar> atoms;
  double scStale[17]{};  // RawParameterProcessor keeps its last Slater-Condon set as a member
};

static const AtomPar& atomPar(Method& m, int Z) {
  auto it = m.atoms.find(Z);
  if (it != m.atoms.end()) return it->second;
  auto rit = m.raw.atoms.find(Z);
  if (rit == m.raw.atoms.end()) throw std::runtime_error("parameters do not exist for element Z=" + std::to_string(Z));
  const RawAtom& r = rit->second;
  const double hpe = hartreePerEv();
  AtomPar a;
  a.Z = Z;
  a.nao = numAOs(Z, m.raw.spd);
  a.l = a.nao == 1 ? 0 : a.nao == 4 ? 1 : 2;
  a.core = coreCharge(Z);
  double sc[17];
  std::memcpy(sc, m.scStale, sizeof(sc));
  if (a.nao == 9) {
    slaterCondon(Z, r.zint, r.f0sd * hpe, r.g2sd * hpe, sc);
    std::memcpy(m.scStale, sc, sizeof(sc));
  }
  for (int i = 0; i < 3; ++i) {
    a.beta[i] = r.beta[i] * hpe;
    a.U[i] = r.U[i] * hpe;
  }
  a.alpha = r.alpha * kAngPerBohr;
  // charge separations (setChargeSeparations :161-174)
  if (a.nao == 4) chargeSeparations(nS(Z), nP(Z), 0, r.zeta[0], r.zeta[1], 0, false, a.D);
  else if (a.nao == 9) chargeSeparations(nS(Z), nP(Z), nD(Z), r.zeta[0], r.zeta[1], r.zeta[2], true, a.D);
  // Klopman-Ohno additive terms (setKlopman :126-159)
  {
    double hpp = 0.5 * (r.gpp - r.gp2) * hpe;
    if (a.nao == 4) hpp = std::max(0.1 * hpe, hpp);  // REPRODUCE_PM6_ERRORS clamp (:134-137)
    const double in[12] = {r.gss * hpe, r.hsp * hpe, a.D[sp1], hpp,        a.D[pp2], sc[F0dd],
                           sc[G1pd],    a.D[pd1],    sc[G2sd], a.D[sd2],   sc[F2dd], a.D[dd2]};
    klopman(a.nao == 1 ? 0 : a.nao == 4 ? 1 : 2, in, a.rho);
  }
  a.pcore = r.pcore == 0 ? a.rho[ss0] : r.pcore;
  a.pcoreSpecified = r.pcore != 0;
  for (auto& g : r.gauss) a.gauss.push_back({g[0] * kBohrPerAng, g[1] * kAngPerBohr * kAngPerBohr, g[2] * kBohrPerAng});
  // one-centre integrals (get1c2eIntegrals :42-65)
  a.n1c = a.nao == 1 ? 1 : a.nao == 4 ? 6 : 58;
  {
    const double g5[5] = {r.gss * hpe, r.gsp * hpe, r.gpp * hpe, r.gp2 * hpe, r.hsp * hpe};
    oneCenterIntegrals(a.n1c, g5, sc, a.oneC);
  }
  auto git = m.raw.gtos.find(Z);
  if (git != m.raw.gtos.end()) a.shells = git->second;
  return m.atoms[Z] = a;
}

static double oneC(const AtomPar& a, int i, int j, int k, int l) {
  int idx = g1cIndex[i][j][k][l];
  if (idx >= a.n1c) return 0;
  return a.oneC[idx];
}

// ---------------------------------------------------------------------------------------------
// Multipole expansion of the local two-centre integrals.
//   charge configurations: IE/ChargesInMultipoles.cpp:27-125
//   term lists per charge-distribution pair: IE/Local2c2eIntegralCalculator.cpp:69-140,142-292
//   point-charge interaction: IE/MultipoleMultipoleTerm.h:57-64 (merged in MMTermCreator.cpp:19-58)
// ---------------------------------------------------------------------------------------------
enum Mp { M00 = 0, Qxx, Qyy, Qzz, M1m1, M10, M11, M2m2, M2m1, M20, M21, M22, Qzx };
struct Charge {
  double x, y, z, q;
};
static std::vector<Charge> gCharges[13];
static int mpM(int m) {
  static const int M[13] = {0, 0, 0, 0, -1, 0, 1, -2, -1, 0, 1, 2, 2};
  return M[m];
}
static int mpL(int m) {
  static const int L[13] = {0, 2, 2, 2, 1, 1, 1, 2, 2, 2, 2, 2, 2};
  return L[m];
}
static void buildCharges() {
  if (!gCharges[0].empty()) return;
  const double r2 = std::sqrt(2.0);
  gCharges[M00] = {{0, 0, 0, 1}};
  gCharges[Qxx] = {{-2, 0, 0, 0.25}, {2, 0, 0, 0.25}, {0, 0, 0, -0.5}};
  gCharges[Qyy] = {{0, -2, 0, 0.25}, {0, 2, 0, 0.25}, {0, 0, 0, -0.5}};
  gCharges[Qzz] = {{0, 0, -2, 0.25}, {0, 0, 2, 0.25}, {0, 0, 0, -0.5}};
  gCharges[M2m2] = {{1, 1, 0, 0.25}, {-1, -1, 0, 0.25}, {1, -1, 0, -0.25}, {-1, 1, 0, -0.25}};
  gCharges[M2m1] = {{0, 1, 1, 0.25}, {0, -1, -1, 0.25}, {0, 1, -1, -0.25}, {0, -1, 1, -0.25}};
  gCharges[M21] = {{1, 0, 1, 0.25}, {-1, 0, -1, 0.25}, {1, 0, -1, -0.25}, {-1, 0, 1, -0.25}};
  gCharges[M20] = {{0, 0, r2, 0.25}, {0, 0, -r2, 0.25}, {r2, 0, 0, -0.375}, {-r2, 0, 0, -0.375}, {0, r2, 0, 0.125}, {0, -r2, 0, 0.125}};
  gCharges[M22] = {{r2, 0, 0, 0.25}, {-r2, 0, 0, 0.25}, {0, r2, 0, -0.25}, {0, -r2, 0, -0.25}};
  gCharges[Qzx] = {{0, 0, r2, 0.25}, {0, 0, -r2, 0.25}, {r2, 0, 0, -0.25}, {-r2, 0, 0, -0.25}};
  gCharges[M10] = {{0, 0, -1, -0.5}, {0, 0, 1, 0.5}};
  gCharges[M11] = {{-1, 0, 0, -0.5}, {1, 0, 0, 0.5}};
  gCharges[M1m1] = {{0, -1, 0, -0.5}, {0, 1, 0, 0.5}};
}
static double mmInteraction(int m1, int m2, double R, double D1, double D2, double rho2) {
  double sum = 0;
  for (const Charge& c1 : gCharges[m1])
    for (const Charge& c2 : gCharges[m2]) {
      double dx = D2 * c2.x - D1 * c1.x;
      double dy = D2 * c2.y - D1 * c1.y;
      double dz = R + D2 * c2.z - D1 * c1.z;
      sum += c1.q * c2.q / std::sqrt(dx * dx + dy * dy + dz * dz + rho2);
    }
  return sum;
}

// charge-distribution index (twoElIntegral_t, IE/GeneralTypes.h:44-55) -> orbital pair
static const int kPairOrb[40][2] = {
    {s_, s_}, {s_, x_}, {x_, x_}, {s_, y_}, {x_, y_}, {y_, y_}, {s_, z_}, {x_, z_}, {y_, z_}, {z_, z_},
    {s_, z2_}, {s_, xz_}, {s_, yz_}, {s_, x2y2_}, {s_, xy_}, {x_, z2_}, {x_, xz_}, {x_, x2y2_}, {x_, xy_},
    {y_, z2_}, {y_, yz_}, {y_, x2y2_}, {y_, xy_}, {z_, z2_}, {z_, xz_}, {z_, yz_}, {z2_, z2_}, {z2_, xz_},
    {z2_, yz_}, {z2_, x2y2_}, {z2_, xy_}, {xz_, xz_}, {xz_, yz_}, {xz_, x2y2_}, {xz_, xy_}, {yz_, yz_},
    {yz_, x2y2_}, {yz_, xy_}, {x2y2_, x2y2_}, {xy_, xy_}};
static int orbL(int o) { return o == 0 ? 0 : o < 4 ? 1 : 2; }
static int gPairIndex[9][9];
static void buildPairIndex() {
  for (auto& r : gPairIndex)
    for (auto& c : r) c = 100;
  for (int p = 0; p < 40; ++p) {
    gPairIndex[kPairOrb[p][0]][kPairOrb[p][1]] = p;
    gPairIndex[kPairOrb[p][1]][kPairOrb[p][0]] = p;
  }
}

struct PoleTerm {
  double f;
  int m;
};
// getMultipoles (IE/Local2c2eIntegralCalculator.cpp:142-292)
static std::vector<PoleTerm> multipolesOf(int t, bool hasD) {
  const double is3 = 1 / std::sqrt(3.0);
  switch (t) {
    case 0: return {{1, M00}};
    case 6: return {{1, M10}};
    case 1: return {{1, M11}};
    case 3: return {{1, M1m1}};
    case 7: return {{1, M21}};
    case 8: return {{1, M2m1}};
    case 4: return {{1, M2m2}};
    case 9: return hasD ? std::vector<PoleTerm>{{1, M00}, {4.0 / 3.0, M20}} : std::vector<PoleTerm>{{1, M00}, {1, Qzz}};
    case 2: return hasD ? std::vector<PoleTerm>{{1, M00}, {-2.0 / 3.0, M20}, {1, M22}} : std::vector<PoleTerm>{{1, M00}, {1, Qxx}};
    case 5: return hasD ? std::vector<PoleTerm>{{1, M00}, {-2.0 / 3.0, M20}, {-1, M22}} : std::vector<PoleTerm>{{1, M00}, {1, Qyy}};
    case 10: return {{2 * is3, M20}};
    case 11: return {{1, M21}};
    case 12: return {{1, M2m1}};
    case 13: return {{1, M22}};
    case 14: return {{1, M2m2}};
    case 23: return {{2 * is3, M10}};
    case 24: return {{1, M11}};
    case 25: return {{1, M1m1}};
    case 15: return {{-is3, M11}};
    case 16: return {{1, M10}};
    case 17: return {{1, M11}};
    case 18: return {{1, M1m1}};
    case 19: return {{-is3, M1m1}};
    case 20: return {{1, M10}};
    case 21: return {{-1, M1m1}};
    case 22: return {{1, M11}};
    case 26: return {{1, M00}, {4.0 / 3.0, M20}};
    case 27: return {{is3, M21}};
    case 28: return {{is3, M2m1}};
    case 29: return {{-2 * is3, M22}};
    case 30: return {{-2 * is3, M2m2}};
    case 31: return {{1, M00}, {2.0 / 3.0, M20}, {1, M22}};
    case 32: return {{1, M2m2}};
    case 33: return {{1, M21}};
    case 34: return {{1, M2m1}};
    case 35: return {{1, M00}, {2.0 / 3.0, M20}, {-1, M22}};
    case 36: return {{-1, M2m1}};
    case 37: return {{1, M21}};
    case 38: return {{1, M00}, {-4.0 / 3.0, M20}};
    case 39: return {{1, M00}, {-4.0 / 3.0, M20}};
  }
  return {};
}
// pairType (IE/multipoleTypes.cpp:186-205)
static int pairType(int l1, int l2, int l) {
  if (l1 == 0 && l2 == 1 && l == 1) return sp1;
  if (l1 == 1 && l2 == 2 && l == 1) return pd1;
  if (l1 == 1 && l2 == 1 && l == 2) return pp2;
  if (l1 == 0 && l2 == 2 && l == 2) return sd2;
  if (l1 == 2 && l2 == 2 && l == 2) return dd2;
  if (l1 == 0 && l2 == 0 && l == 0) return ss0;
  if (l1 == 1 && l2 == 1 && l == 0) return pp0;
  if (l1 == 2 && l2 == 2 && l == 0) return dd0;
  throw std::runtime_error("invalid quantum numbers");
}
struct LocalTerm {
  double f;
  int p1, p2, m1, m2;
};
static std::vector<LocalTerm> gLocalTerms[40][40];
static void buildLocalTerms() {
  static bool done = false;
  if (done) return;
  buildCharges();
  buildPairIndex();
  for (int i = 0; i < 40; ++i)
    for (int j = 0; j < 40; ++j) {
      bool hasD = !(i < 10 && j < 10);
      auto l1 = multipolesOf(i, hasD), l2 = multipolesOf(j, hasD);
      int l1a = orbL(kPairOrb[i][0]), l1b = orbL(kPairOrb[i][1]);
      int l2a = orbL(kPairOrb[j][0]), l2b = orbL(kPairOrb[j][1]);
      for (auto& a : l1)
        for (auto& b : l2) {
          int mp1 = pairType(l1a, l1b, mpL(a.m)), mp2 = pairType(l2a, l2b, mpL(b.m));
          if (hasD && a.m == M20 && b.m == M20) {
            // REPRODUCE_PM6_ERRORS: [M20,M20] = [Qzx,Qzx] - 1/4 [M22,M22]  (:87-104)
            gLocalTerms[i][j].push_back({a.f * b.f, mp1, mp2, Qzx, Qzx});
            gLocalTerms[i][j].push_back({a.f * b.f * -0.25, mp1, mp2, M22, M22});
          } else if (mpM(a.m) == mpM(b.m)) {
            gLocalTerms[i][j].push_back({a.f * b.f, mp1, mp2, a.m, b.m});
          }
        }
    }
  done = true;
}
// Local2c2eIntegralCalculator::getIntegral (:38-51)
static double localIntegral(int t1, int t2, double R, const double* D1, const double* D2, const double* r1, const double* r2) {
  double sum = 0;
  for (const LocalTerm& t : gLocalTerms[t1][t2]) {
    double rs = r1[t.p1] + r2[t.p2];
    double d1 = t.p1 < ss0 ? D1[t.p1] : 0;
    double d2 = t.p2 < ss0 ? D2[t.p2] : 0;
    sum += t.f * mmInteraction(t.m1, t.m2, R, d1, d2, rs * rs);
  }
  return sum;
}


// ---------------------------------------------------------------------------------------------
// Local matrix (Local2c2eMatrix.cpp:30-923).  Every entry the reference fills (the ones with a
// non-empty term list, which is also exactly ZeroLocal2c2eIntegrals' non-zero set) equals the
// direct getIntegral() value, except (x_y|x_y) which is set to 1/2[(xx|xx)-(xx|yy)]
// (buildPPMatrix :139-141, buildPPMatrixSym :169-171).
// ---------------------------------------------------------------------------------------------
static void localMatrix(int l1, int l2, double R, const double* D1, const double* D2, const double* r1,
                        const double* r2, double* mat /* d1 x d2 row-major */) {
  int d1 = l1 == 0 ? 1 : l1 == 1 ? 10 : 40, d2 = l2 == 0 ? 1 : l2 == 1 ? 10 : 40;
  for (int i = 0; i < d1; ++i)
    for (int j = 0; j < d2; ++j)
      mat[i * d2 + j] = gLocalTerms[i][j].empty() ? 0.0 : localIntegral(i, j, R, D1, D2, r1, r2);
  if (l1 > 0 && l2 > 0) mat[4 * d2 + 4] = 0.5 * (mat[2 * d2 + 2] - mat[2 * d2 + 5]);
}

// ---------------------------------------------------------------------------------------------
// Rotation to the global frame.
//   coefficients r(global orbital, local orbital): IE/OrbitalRotation.cpp:77-154
//   term lists: IE/Global2c2eTerms.cpp:20-90 ; evaluation: IE/Global2c2eMatrix.cpp:46-57
// ---------------------------------------------------------------------------------------------
static void rotationCoefficients(const double Rab[3], double r[9][9]) {
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) r[i][j] = 0;
  double X = Rab[0], Y = Rab[1], Z = Rab[2];
  double R = std::sqrt(X * X + Y * Y + Z * Z);
  double pp = 0;
  if (std::sqrt(X * X + Y * Y) != 0) pp = std::sqrt(X * X + Y * Y);
  double ca, sa;
  if (pp != 0) {
    ca = X / pp;
    sa = Y / pp;
  } else {
    ca = 1;
    sa = 0;
  }
  double cb = Z / R, sb = pp / R;
  double s2b = sb * sb, s2a = sa * sa, c2b = cb * cb, c2a = ca * ca;
  (void)s2a;
  double saca = sa * ca, sbcb = sb * cb, dc2am1 = 2 * c2a - 1;
  const double r3 = std::sqrt(3.0);
  r[s_][s_] = 1;
  r[x_][x_] = ca * cb;
  r[y_][x_] = sa * cb;
  r[z_][x_] = -sb;
  r[x_][y_] = -sa;
  r[y_][y_] = ca;
  r[x_][z_] = ca * sb;
  r[y_][z_] = sa * sb;
  r[z_][z_] = cb;
  r[x2y2_][x2y2_] = dc2am1 * c2b + 0.5 * dc2am1 * s2b;
  r[x2y2_][xz_] = dc2am1 * sbcb;
  r[x2y2_][z2_] = 0.5 * r3 * dc2am1 * s2b;
  r[x2y2_][yz_] = -2 * saca * sb;
  r[x2y2_][xy_] = -2 * saca * cb;
  r[xz_][x2y2_] = -ca * sbcb;
  r[xz_][xz_] = ca * (2 * c2b - 1);
  r[xz_][z2_] = r3 * ca * sbcb;
  r[xz_][yz_] = -sa * cb;
  r[xz_][xy_] = sa * sb;
  r[z2_][x2y2_] = 0.5 * r3 * s2b;
  r[z2_][xz_] = -r3 * sbcb;
  r[z2_][z2_] = c2b - 0.5 * s2b;
  r[yz_][x2y2_] = -sa * sbcb;
  r[yz_][xz_] = sa * (2 * c2b - 1);
  r[yz_][z2_] = r3 * sa * sbcb;
  r[yz_][yz_] = ca * cb;
  r[yz_][xy_] = -ca * sb;
  r[xy_][x2y2_] = 2 * saca * c2b + saca * s2b;
  r[xy_][xz_] = 2 * saca * sbcb;
  r[xy_][z2_] = r3 * saca * s2b;
  r[xy_][yz_] = dc2am1 * sb;
  r[xy_][xy_] = dc2am1 * cb;
}

struct RotTerm {
  unsigned char o[8];  // global i0..i3, local i4..i7
  unsigned char p2, p3;
};
static std::vector<RotTerm> gRotTerms[40][40];
static void buildRotTerms() {
  static bool done = false;
  if (done) return;
  buildLocalTerms();
  auto compat = [](int a, int b) { return orbL(a) == orbL(b); };
  int i[8];
  for (i[0] = 0; i[0] < 9; i[0]++)
    for (i[1] = i[0]; i[1] < 9; i[1]++)
      for (i[2] = 0; i[2] < 9; i[2]++)
        for (i[3] = i[2]; i[3] < 9; i[3]++)
          for (i[4] = 0; i[4] < 9; i[4]++)
            for (i[5] = 0; i[5] < 9; i[5]++)
              for (i[6] = 0; i[6] < 9; i[6]++)
                for (i[7] = 0; i[7] < 9; i[7]++) {
                  int pr[4];
                  bool ok = true;
                  for (int k = 0; k < 4 && ok; ++k) {
                    if (!compat(i[k], i[k + 4])) ok = false;
                    pr[k] = gPairIndex[i[2 * k]][i[2 * k + 1]];
                    if (pr[k] == 100) ok = false;
                  }
                  if (!ok) continue;
                  if (gLocalTerms[pr[2]][pr[3]].empty()) continue;
                  for (int k = 0; k < 4 && ok; ++k) {  // rotations that vanish identically
                    if (i[k] == z_ && i[k + 4] == y_) ok = false;
                    if (i[k] == z2_ && (i[k + 4] == yz_ || i[k + 4] == xy_)) ok = false;
                  }
                  if (!ok) continue;
                  RotTerm t;
                  for (int k = 0; k < 8; ++k) t.o[k] = (unsigned char)i[k];
                  t.p2 = (unsigned char)pr[2];
                  t.p3 = (unsigned char)pr[3];
                  gRotTerms[pr[0]][pr[1]].push_back(t);
                }
  done = true;
}

// Global2c2eMatrix::calculate<Zero> (IE/Global2c2eMatrix.cpp:119-129): global d1 x d2 block
// (row-major) for atoms with parameters (D1,r1),(D2,r2), Rab = R_B - R_A.
static void globalBlock(int l1, int l2, const double Rab[3], const double* D1, const double* D2, const double* r1,
                        const double* r2, double* out) {
  buildRotTerms();
  int d1 = l1 == 0 ? 1 : l1 == 1 ? 10 : 40, d2 = l2 == 0 ? 1 : l2 == 1 ? 10 : 40;
  double R = std::sqrt(Rab[0] * Rab[0] + Rab[1] * Rab[1] + Rab[2] * Rab[2]);
  std::vector<double> loc(d1 * d2);
  localMatrix(l1, l2, R, D1, D2, r1, r2, loc.data());
  double rot[9][9];
  rotationCoefficients(Rab, rot);
  for (int a = 0; a < d1; ++a)
    for (int b = 0; b < d2; ++b) {
      double v = 0;
      for (const RotTerm& t : gRotTerms[a][b])
        v += rot[t.o[0]][t.o[4]] * rot[t.o[1]][t.o[5]] * rot[t.o[2]][t.o[6]] * rot[t.o[3]][t.o[7]] * loc[t.p2 * d2 + t.p3];
      out[a * d2 + b] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// STO-6G overlap (IE/GTOOverlapMatrixBlock.cpp:49-165, IE/AtomPairOverlap.cpp:17-58).
// Obara-Saika recursion per Cartesian direction, six Cartesian d GTFs folded into five
// spherical d (xy-type normalisation), AO order s | x y z | x2y2 xz z2 yz xy.
// ---------------------------------------------------------------------------------------------
static void shellOverlap(const Shell& A, const Shell& B, const double Rab[3], double out[5][5] /* nAO_A x nAO_B */) {
  static const int mom[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1},
                                 {0, 1, 1}, {1, 1, 0}, {0, 0, 2}, {0, 2, 0}, {2, 0, 0}};
  int startA = A.l == 0 ? 0 : A.l == 1 ? 1 : 4, startB = B.l == 0 ? 0 : B.l == 1 ? 1 : 4;
  int nA = A.l == 0 ? 1 : A.l == 1 ? 3 : 6, nB = B.l == 0 ? 1 : B.l == 1 ? 3 : 6;
  double res[6][6] = {};
  const double pi = 4.0 * std::atan(1.0);
  double x = Rab[0], y = Rab[1], z = Rab[2];
  double rr[3] = {x, y, z};
  for (size_t ia = 0; ia < A.e.size(); ++ia)
    for (size_t ib = 0; ib < B.e.size(); ++ib) {
      double ea = A.e[ia], eb = B.e[ib], es = ea + eb;
      double fp = eb / es, fac = -fp * ea;
      double KD = std::exp((x * x + y * y + z * z) * fac) * (A.c[ia] * B.c[ib]);
      double S[3][3][3];
      for (int d = 0; d < 3; ++d) {
        double p = rr[d] * fp, dist2 = p - rr[d];
        S[d][0][0] = std::sqrt(pi / es);
        for (int k = 0; k <= A.l; ++k)
          for (int l = 0; l <= B.l; ++l) {
            if (k == 0 && l == 0) continue;
            double v = 0;
            if (l == 0) {
              if (k > 1) v = S[d][k - 2][l] * ((k - 1.0) / (2 * es));
              v += S[d][k - 1][l] * p;
            } else {
              if (l > 1) v = S[d][k][l - 2] * ((l - 1.0) / (2 * es));
              if (k > 0) v += S[d][k - 1][l - 1] * (k / (2 * es));
              v += S[d][k][l - 1] * dist2;
            }
            S[d][k][l] = v;
          }
      }
      for (int k = 0; k < nA; ++k)
        for (int l = 0; l < nB; ++l) {
          const int* ma = mom[k + startA];
          const int* mb = mom[l + startB];
          res[k][l] += KD * (S[0][ma[0]][mb[0]] * S[1][ma[1]][mb[1]] * S[2][ma[2]][mb[2]]);
        }
    }
  const double r3 = std::sqrt(3.0);
  if (nA == 6)
    for (int l = 0; l < nB; ++l) {
      res[3][l] = (res[3][l] - 0.5 * res[4][l] - 0.5 * res[5][l]) / r3;
      res[4][l] = 0.5 * (res[5][l] - res[4][l]);
    }
  if (nB == 6)
    for (int k = 0; k < nA; ++k) {
      res[k][3] = (res[k][3] - 0.5 * res[k][4] - 0.5 * res[k][5]) / r3;
      res[k][4] = 0.5 * (res[k][5] - res[k][4]);
    }
  // Cartesian slot -> AO position inside the shell (d: xz yz xy z2 x2y2 -> x2y2 xz z2 yz xy)
  static const int dPos[5] = {1, 3, 4, 2, 0};
  int naoA = 2 * A.l + 1, naoB = 2 * B.l + 1;
  for (int k = 0; k < naoA; ++k)
    for (int l = 0; l < naoB; ++l) {
      int pk = A.l == 2 ? dPos[k] : k, pl = B.l == 2 ? dPos[l] : l;
      out[pk][pl] = res[k][l];
    }
}

static void atomPairOverlap(const AtomPar& A, const AtomPar& B, const double Rab[3], double* blk /* nA x nB */) {
  for (int i = 0; i < A.nao * B.nao; ++i) blk[i] = 0;
  for (const Shell& sa : A.shells)
    for (const Shell& sb : B.shells) {
      int offA = sa.l == 0 ? 0 : sa.l == 1 ? 1 : 4, offB = sb.l == 0 ? 0 : sb.l == 1 ? 1 : 4;
      if (offA >= A.nao || offB >= B.nao) continue;
      double o[5][5];
      shellOverlap(sa, sb, Rab, o);
      for (int k = 0; k < 2 * sa.l + 1; ++k)
        for (int l = 0; l < 2 * sb.l + 1; ++l) blk[(offA + k) * B.nao + offB + l] = o[k][l];
    }
}

// ---------------------------------------------------------------------------------------------
// A molecule and its density-independent quantities.
// ---------------------------------------------------------------------------------------------
struct Mol {
  Method* m = nullptr;
  int nat = 0, nao = 0;
  std::vector<int> Z, first;
  std::vector<double> xyz;  // bohr, row-major nat x 3
  std::vector<const AtomPar*> par;
  double nElectronsNeutral = 0;
  // density-independent parts
  std::vector<std::vector<double>> pairBlocks;  // i<j pair -> d_i x d_j global 2c2e block
  std::vector<double> S, H;                     // n x n col-major, lower triangle + diagonal
  double Erep = 0;
  size_t pairId(int i, int j) const { return (size_t)i * nat + j; }
};

static int dimOf(int nao) { return nao == 1 ? 1 : nao == 4 ? 10 : 40; }
static int lOf(int nao) { return nao == 1 ? 0 : nao == 4 ? 1 : 2; }

static void setupMol(Mol& mol, Method& m, int nat, const int* Z, const double* xyz) {
  mol.m = &m;
  mol.nat = nat;
  mol.Z.assign(Z, Z + nat);
  mol.xyz.assign(xyz, xyz + 3 * nat);
  mol.par.resize(nat);
  mol.first.resize(nat);
  mol.nao = 0;
  mol.nElectronsNeutral = 0;
  for (int i = 0; i < nat; ++i) {
    mol.par[i] = &atomPar(m, Z[i]);
    mol.first[i] = mol.nao;
    mol.nao += mol.par[i]->nao;
    mol.nElectronsNeutral += std::lround(mol.par[i]->core);
  }
  if (m.raw.diatomic)
    for (int i = 0; i < nat; ++i)
      for (int j = 0; j <= i; ++j)
        if (!m.raw.pairs.count({std::min(Z[i], Z[j]), std::max(Z[i], Z[j])}))
          throw std::runtime_error("parameters do not exist for element pair " + std::to_string(Z[i]) + "-" + std::to_string(Z[j]));
}

// TwoCenterIntegralContainer::update (IE/TwoCenterIntegralContainer.cpp:52-75)
static void computeTwoCenter(Mol& mol) {
  buildRotTerms();
  int nat = mol.nat;
  mol.pairBlocks.assign((size_t)nat * nat, {});
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < nat; ++i)
    for (int j = i + 1; j < nat; ++j) {
      const AtomPar& A = *mol.par[i];
      const AtomPar& B = *mol.par[j];
      double Rab[3] = {mol.xyz[3 * j] - mol.xyz[3 * i], mol.xyz[3 * j + 1] - mol.xyz[3 * i + 1], mol.xyz[3 * j + 2] - mol.xyz[3 * i + 2]};
      std::vector<double> blk(dimOf(A.nao) * dimOf(B.nao));
      globalBlock(A.l, B.l, Rab, A.D, B.D, A.rho, B.rho, blk.data());
      mol.pairBlocks[mol.pairId(i, j)] = std::move(blk);
    }
}

static double pairGet(const Mol& mol, int i, int j, int pa, int pb) {
  // (pair pa on atom i | pair pb on atom j), i<j
  if (pa == 100 || pb == 100) return 0;
  return mol.pairBlocks[mol.pairId(i, j)][pa * dimOf(mol.par[j]->nao) + pb];
}

// OverlapMatrix::calculateOverlap (IE/OverlapMatrix.cpp:32-64)
static void computeOverlap(Mol& mol) {
  int n = mol.nao;
  mol.S.assign((size_t)n * n, 0.0);
  for (int k = 0; k < n; ++k) mol.S[(size_t)k * n + k] = 1.0;
#pragma omp parallel for schedule(dynamic)
  for (int i = 1; i < mol.nat; ++i)
    for (int j = 0; j < i; ++j) {
      const AtomPar& A = *mol.par[i];
      const AtomPar& B = *mol.par[j];
      double Rab[3] = {mol.xyz[3 * j] - mol.xyz[3 * i], mol.xyz[3 * j + 1] - mol.xyz[3 * i + 1], mol.xyz[3 * j + 2] - mol.xyz[3 * i + 2]};
      double blk[81];
      atomPairOverlap(A, B, Rab, blk);
      for (int a = 0; a < A.nao; ++a)
        for (int b = 0; b < B.nao; ++b) mol.S[(size_t)(mol.first[j] + b) * n + mol.first[i] + a] = blk[a * B.nao + b];
    }
}

// OneElectronMatrix::calculate (NU/OneElectronMatrix.cpp:43-134)
static void computeHcore(Mol& mol) {
  int n = mol.nao;
  mol.H.assign((size_t)n * n, 0.0);
  auto H = [&](int r, int c) -> double& { return mol.H[(size_t)c * n + r]; };
#pragma omp parallel for schedule(dynamic)
  for (int a = 0; a < mol.nat; ++a) {
    const AtomPar& A = *mol.par[a];
    int f = mol.first[a];
    for (int i = 0; i < A.nao; ++i) H(f + i, f + i) = i < 1 ? A.U[0] : i < 4 ? A.U[1] : A.U[2];
    for (int b = 0; b < mol.nat; ++b) {
      if (b == a) continue;
      const AtomPar& B = *mol.par[b];
      if (B.pcoreSpecified) {
        // VuvB (IE/VuvB.cpp:17-31): -Z_B (mu nu | s_B s_B) with rho_B = pcore_B, D_B = 0
        double Rab[3] = {mol.xyz[3 * b] - mol.xyz[3 * a], mol.xyz[3 * b + 1] - mol.xyz[3 * a + 1], mol.xyz[3 * b + 2] - mol.xyz[3 * a + 2]};
        double Dc[5] = {0, 0, 0, 0, 0}, rc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        rc[ss0] = B.pcore;
        double v[40];
        globalBlock(A.l, 0, Rab, A.D, Dc, A.rho, rc, v);
        for (int i = 0; i < A.nao; ++i)
          for (int j = 0; j <= i; ++j) {
            int p = gPairIndex[i][j];
            H(f + i, f + j) += p == 100 ? 0.0 : -B.core * v[p];
          }
      } else {
        for (int i = 0; i < A.nao; ++i)
          for (int j = 0; j <= i; ++j) {
            int p = gPairIndex[i][j];
            double g = a < b ? pairGet(mol, a, b, p, 0) : pairGet(mol, b, a, 0, p);
            H(f + i, f + j) += -B.core * g;
          }
      }
    }
  }
  // off-diagonal blocks 1/2 (beta_mu + beta_nu) S_mu,nu  (:113-134)
  for (int i = 1; i < mol.nat; ++i)
    for (int j = 0; j < i; ++j) {
      const AtomPar& A = *mol.par[i];
      const AtomPar& B = *mol.par[j];
      for (int bj = 0; bj < B.nao; ++bj)
        for (int ai = 0; ai < A.nao; ++ai) {
          double ba = ai < 1 ? A.beta[0] : ai < 4 ? A.beta[1] : A.beta[2];
          double bb = bj < 1 ? B.beta[0] : bj < 4 ? B.beta[1] : B.beta[2];
          int r = mol.first[i] + ai, c = mol.first[j] + bj;
          H(r, c) = 0.5 * (ba + bb) * mol.S[(size_t)c * n + r];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Core-core repulsion.
// ---------------------------------------------------------------------------------------------
static double pm6PairRepulsion(const Method& m, const AtomPar& A, const AtomPar& B, double R) {
  // N/Pm6/PM6PairwiseRepulsion.h:76-175, constants PM6PairwiseRepulsion.cpp:17-31
  const double cAdd = 0.0000207389327971913490822259;
  const double expCoef = 0.00001244878365801758178693335;
  const double expCC = 3.1644797213016;
  const double distSiO = 5.480205761238679760340424;
  const double facCC = 9.28, facSiO = -0.0007;
  auto pr = m.raw.pairs.at({std::min(A.Z, B.Z), std::max(A.Z, B.Z)});
  auto isHX = [](int a, int b) { return a == 1 && (b == 8 || b == 7 || b == 6); };
  double alphaAB = (isHX(A.Z, B.Z) || isHX(B.Z, A.Z)) ? pr.first * kAngPerBohr * kAngPerBohr : pr.first * kAngPerBohr;
  double xAB = pr.second;
  double ps = A.pcore + B.pcore;
  double ssss = 1.0 / std::sqrt(R * R + ps * ps);
  double R6 = R * R * R * R * R * R;
  double standard = 1.0 + 2 * xAB * std::exp(-alphaAB * (R + expCoef * R6));
  double par;
  if (isHX(A.Z, B.Z) || isHX(B.Z, A.Z)) par = 1.0 + 2 * xAB * std::exp(-alphaAB * (R * R));
  else if (A.Z == 6 && B.Z == 6) par = standard + facCC * std::exp(-expCC * R);
  else if ((A.Z == 14 && B.Z == 8) || (A.Z == 8 && B.Z == 14)) par = standard + facSiO * std::exp(-(R - distSiO) * (R - distSiO));
  else par = standard;
  double base = A.core * B.core * ssss * par;
  double v = (std::pow((double)A.Z, 1.0 / 3.0) + std::pow((double)B.Z, 1.0 / 3.0)) / R;
  double v2 = v * v, v6 = v2 * v2 * v2;
  double additional = cAdd * v6 * v6 / gEvPerHartree;
  auto gauss1 = [&](const AtomPar& P) {
    if (P.gauss.empty()) return 0.0;
    const auto& g = P.gauss[0];  // only the first Gaussian (:127-131)
    return g[0] * std::exp(-g[1] * (R - g[2]) * (R - g[2]));
  };
  double gauss = (gauss1(A) + gauss1(B)) / R * (A.core * B.core / gEvPerHartree);
  return base + additional + gauss;
}
static double mndoPairRepulsion(const AtomPar& A, const AtomPar& B, double R, bool gaussians) {
  // N/Mndo/MNDOPairwiseRepulsion.h:62-125 ; N/Am1/AM1PairwiseRepulsion.h:64-121
  double ps = A.pcore + B.pcore;
  double ssss = 1.0 / std::sqrt(R * R + ps * ps);
  double par;
  if (A.Z == 1 && (B.Z == 7 || B.Z == 8))
    par = 1.0 + std::exp(-A.alpha * R) + R * kAngPerBohr * std::exp(-B.alpha * R);
  else if (B.Z == 1 && (A.Z == 7 || A.Z == 8))
    par = 1.0 + R * kAngPerBohr * std::exp(-A.alpha * R) + std::exp(-B.alpha * R);
  else
    par = 1.0 + std::exp(-A.alpha * R) + std::exp(-B.alpha * R);
  double e = A.core * B.core * ssss * par;
  if (gaussians) {
    auto gs = [&](const AtomPar& P) {
      double s = 0;
      for (const auto& g : P.gauss) s += g[0] * std::exp(-g[1] * (R - g[2]) * (R - g[2]));
      return s;
    };
    e += (gs(A) + gs(B)) / R * (A.core * B.core / gEvPerHartree);
  }
  return e;
}
static double computeRepulsion(Mol& mol) {
  double e = 0;
  const std::string& nm = mol.m->raw.name;
#pragma omp parallel for reduction(+ : e) schedule(dynamic)
  for (int i = 0; i < mol.nat; ++i)
    for (int j = i + 1; j < mol.nat; ++j) {
      double dx = mol.xyz[3 * j] - mol.xyz[3 * i], dy = mol.xyz[3 * j + 1] - mol.xyz[3 * i + 1], dz = mol.xyz[3 * j + 2] - mol.xyz[3 * i + 2];
      double R = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (nm == "PM6") e += pm6PairRepulsion(*mol.m, *mol.par[i], *mol.par[j], R);
      else e += mndoPairRepulsion(*mol.par[i], *mol.par[j], R, nm != "MNDO");
    }
  return e;
}

// ---------------------------------------------------------------------------------------------
// G(P) (NU/TwoElectronMatrix.cpp:63-166), restricted.  P is full symmetric, col-major; G gets the
// lower triangle + diagonal only.  Serial, as in the reference.
// ---------------------------------------------------------------------------------------------
static void computeG(const Mol& mol, const double* P, double* G) {
  int n = mol.nao;
  std::fill(G, G + (size_t)n * n, 0.0);
  auto Pm = [&](int r, int c) { return P[(size_t)c * n + r]; };
  auto Gm = [&](int r, int c) -> double& { return G[(size_t)c * n + r]; };
  for (int a = 0; a < mol.nat; ++a) {
    const AtomPar& A = *mol.par[a];
    int s = mol.first[a], na = A.nao;
    for (int i = 0; i < na; ++i)
      for (int j = 0; j <= i; ++j)
        for (int k = 0; k < na; ++k)
          for (int l = 0; l < na; ++l)
            Gm(s + i, s + j) += Pm(s + k, s + l) * (oneC(A, i, j, k, l) - 0.5 * oneC(A, i, k, j, l));
  }
  for (int a = 0; a < mol.nat; ++a)
    for (int b = a + 1; b < mol.nat; ++b) {
      int sA = mol.first[a], sB = mol.first[b], nA = mol.par[a]->nao, nB = mol.par[b]->nao;
      for (int i = 0; i < nA; ++i)
        for (int j = 0; j <= i; ++j)
          for (int k = 0; k < nB; ++k)
            for (int l = 0; l <= k; ++l) {
              int mu = sA + i, nu = sA + j, la = sB + k, si = sB + l;
              double I = pairGet(mol, a, b, gPairIndex[i][j], gPairIndex[k][l]);
              int mA = mu == nu ? 1 : 2, mB = la == si ? 1 : 2;
              Gm(mu, nu) += Pm(la, si) * I * mB;
              Gm(la, si) += Pm(mu, nu) * I * mA;
              Gm(la, mu) += -0.5 * Pm(si, nu) * I;
              if (mu > nu) {
                Gm(la, nu) += -0.5 * Pm(si, mu) * I;
                if (la > si) Gm(si, nu) += -0.5 * Pm(la, mu) * I;
              }
              if (la > si) Gm(si, mu) += -0.5 * Pm(la, nu) * I;
            }
    }
}

// Spin-polarised G (TwoElectronMatrix::calculateSameAtomBlock / calculateDifferentAtomsBlock,
// the spinPolarized_ branches, NU/TwoElectronMatrix.cpp:82-108,142-160): Coulomb with the total
// density P = Pa + Pb, exchange with the same-spin density and no factor 1/2.  Lower triangles
// of Ga and Gb are filled, as in the restricted computeG.
static void computeGUnrestricted(const Mol& mol, const double* Pa, const double* Pb, double* Ga, double* Gb) {
  int n = mol.nao;
  std::fill(Ga, Ga + (size_t)n * n, 0.0);
  std::fill(Gb, Gb + (size_t)n * n, 0.0);
  auto Pm = [&](int r, int c) { return Pa[(size_t)c * n + r] + Pb[(size_t)c * n + r]; };
  auto Pam = [&](int r, int c) { return Pa[(size_t)c * n + r]; };
  auto Pbm = [&](int r, int c) { return Pb[(size_t)c * n + r]; };
  auto Gam = [&](int r, int c) -> double& { return Ga[(size_t)c * n + r]; };
  auto Gbm = [&](int r, int c) -> double& { return Gb[(size_t)c * n + r]; };
  for (int a = 0; a < mol.nat; ++a) {
    const AtomPar& A = *mol.par[a];
    int s = mol.first[a], na = A.nao;
    for (int i = 0; i < na; ++i)
      for (int j = 0; j <= i; ++j)
        for (int k = 0; k < na; ++k)
          for (int l = 0; l < na; ++l) {
            Gam(s + i, s + j) += Pm(s + k, s + l) * oneC(A, i, j, k, l) - Pam(s + k, s + l) * oneC(A, i, k, j, l);
            Gbm(s + i, s + j) += Pm(s + k, s + l) * oneC(A, i, j, k, l) - Pbm(s + k, s + l) * oneC(A, i, k, j, l);
          }
  }
  for (int a = 0; a < mol.nat; ++a)
    for (int b = a + 1; b < mol.nat; ++b) {
      int sA = mol.first[a], sB = mol.first[b], nA = mol.par[a]->nao, nB = mol.par[b]->nao;
      for (int i = 0; i < nA; ++i)
        for (int j = 0; j <= i; ++j)
          for (int k = 0; k < nB; ++k)
            for (int l = 0; l <= k; ++l) {
              int mu = sA + i, nu = sA + j, la = sB + k, si = sB + l;
              double I = pairGet(mol, a, b, gPairIndex[i][j], gPairIndex[k][l]);
              int mA = mu == nu ? 1 : 2, mB = la == si ? 1 : 2;
              for (int spin = 0; spin < 2; ++spin) {
                double* Gs = spin == 0 ? Ga : Gb;
                const double* Pss = spin == 0 ? Pa : Pb;
                auto G = [&](int r, int c) -> double& { return Gs[(size_t)c * n + r]; };
                auto Ps = [&](int r, int c) { return Pss[(size_t)c * n + r]; };
                G(mu, nu) += Pm(la, si) * I * mB;
                G(la, si) += Pm(mu, nu) * I * mA;
                G(la, mu) -= Ps(si, nu) * I;
                if (mu > nu) {
                  G(la, nu) -= Ps(si, mu) * I;
                  if (la > si) G(si, nu) -= Ps(la, mu) * I;
                }
                if (la > si) G(si, mu) -= Ps(la, nu) * I;
              }
            }
    }
}

// AO Fock matrix of a CIS pseudo-density (CISMatrixAOFockBuilder::buildFock,
// N/TimeDependent/LinearResponse/CISMatrixAOFockBuilder.cpp:171-330, blocks assembled at :48-170):
//   F = cJ J(DJ) - cK K(DK^T),  J(Y)_xy = sum_cd (xy|cd) Y_cd,  K(Y)_xy = sum_cd (xc|yd) Y_cd,
// in the NDDO approximation (Coulomb on the atom blocks from the one- and two-centre integrals,
// exchange on the atom blocks from the one-centre and on the pair blocks from the two-centre
// integrals).  The reference's cases: restricted singlet cJ = 2 c1, DJ = DK = D; restricted
// triplet cJ = 0, DK = D; unrestricted (per spin s) cJ = c1, DJ = D^alpha + D^beta, DK = D^s; with
// cK = c2 throughout.  D is not symmetric; F is full (both triangles), column-major.
static void computeCisFock(const Mol& mol, const double* DJ, const double* DK, double cJ, double cK, double* F) {
  int n = mol.nao;
  std::fill(F, F + (size_t)n * n, 0.0);
  auto Dj = [&](int r, int c) { return DJ[(size_t)c * n + r]; };
  auto Dk = [&](int r, int c) { return DK[(size_t)c * n + r]; };
  auto Fm = [&](int r, int c) -> double& { return F[(size_t)c * n + r]; };
  for (int a = 0; a < mol.nat; ++a) {
    const AtomPar& A = *mol.par[a];
    int s = mol.first[a], na = A.nao;
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < na; ++j)
        for (int k = 0; k < na; ++k)
          for (int l = 0; l < na; ++l)
            Fm(s + i, s + j) += cJ * oneC(A, i, j, k, l) * Dj(s + k, s + l) - cK * oneC(A, i, k, j, l) * Dk(s + l, s + k);
  }
  for (int a = 0; a < mol.nat; ++a)
    for (int b = a + 1; b < mol.nat; ++b) {
      int sA = mol.first[a], sB = mol.first[b], nA = mol.par[a]->nao, nB = mol.par[b]->nao;
      for (int i = 0; i < nA; ++i)
        for (int j = 0; j < nA; ++j)
          for (int k = 0; k < nB; ++k)
            for (int l = 0; l < nB; ++l) {
              const double I = pairGet(mol, a, b, gPairIndex[i][j], gPairIndex[k][l]);  // (ij|kl), i,j on A
              Fm(sA + i, sA + j) += cJ * I * Dj(sB + k, sB + l);
              Fm(sB + k, sB + l) += cJ * I * Dj(sA + i, sA + j);
              Fm(sA + i, sB + k) -= cK * I * Dk(sB + l, sA + j);
              Fm(sB + k, sA + i) -= cK * I * Dk(sA + j, sB + l);
            }
    }
}

// NDDOElectronicEnergyCalculator::unrestrictedEnergy (NU/NDDOElectronicEnergyCalculator.cpp:69-89)
static double electronicEnergyUnrestricted(int n, const double* Pa, const double* Pb, const double* H, const double* Ga,
                                           const double* Gb) {
  double e = 0;
  for (int i = 0; i < n; ++i) {
    const size_t d = (size_t)i * n + i;
    e += 0.5 * ((Pa[d] + Pb[d]) * 2 * H[d] + Pa[d] * Ga[d] + Pb[d] * Gb[d]);
    for (int j = 0; j < i; ++j) {
      const size_t o = (size_t)j * n + i;
      e += (Pa[o] + Pb[o]) * 2 * H[o] + Pa[o] * Ga[o] + Pb[o] * Gb[o];
    }
  }
  return e;
}

// NDDOElectronicEnergyCalculator::restrictedEnergy (NU/NDDOElectronicEnergyCalculator.cpp:48-67)
static double electronicEnergy(int n, const double* P, const double* H, const double* G) {
  double e = 0;
  for (int i = 0; i < n; ++i) {
    e += 0.5 * P[(size_t)i * n + i] * (G[(size_t)i * n + i] + 2 * H[(size_t)i * n + i]);
    for (int j = 0; j < i; ++j) e += P[(size_t)j * n + i] * (G[(size_t)j * n + i] + 2 * H[(size_t)j * n + i]);
  }
  return e;
}

// NDDODensityGuess::calculateGuess (NU/NDDODensityGuess.cpp:31-66)
static void densityGuess(const Mol& mol, double nElectrons, double* P) {
  int n = mol.nao;
  for (size_t k = 0; k < (size_t)n * n; ++k) P[k] = mol.S[k] * nElectrons / (2.0 * n);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) P[(size_t)j * n + i] = P[(size_t)i * n + j];
  int idx = 0;
  for (int a = 0; a < mol.nat; ++a)
    for (int j = 0; j < mol.par[a]->nao; ++j, ++idx) P[(size_t)idx * n + idx] = mol.par[a]->core / mol.par[a]->nao;
}

// ---------------------------------------------------------------------------------------------
// Symmetric eigensolver: Householder tridiagonalisation followed by the implicit QL algorithm
// with Wilkinson shifts and eigenvector accumulation (Wilkinson & Reinsch, Handbook for
// Automatic Computation vol. II, procedures tred2 / tql2; Golub & Van Loan ch. 8).
// a: n x n symmetric, row-major, overwritten by the eigenvectors (column k = eigenvector k);
// w: eigenvalues in ascending order.
// ---------------------------------------------------------------------------------------------
static void symmetricEigen(int n, std::vector<double>& a, std::vector<double>& w) {
  std::vector<double> d(n), e(n);
  auto A = [&](int r, int c) -> double& { return a[(size_t)r * n + c]; };
  // --- reduction to tridiagonal form, working on rows from the bottom up ---
  for (int i = n - 1; i > 0; --i) {
    int l = i - 1;
    double h = 0, scale = 0;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += std::abs(A(i, k));
      if (scale == 0.0) {
        e[i] = A(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          A(i, k) /= scale;
          h += A(i, k) * A(i, k);
        }
        double f = A(i, l);
        double g = f >= 0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        A(i, l) = f - g;
        f = 0;
        for (int j = 0; j <= l; ++j) {
          A(j, i) = A(i, j) / h;
          double gg = 0;
          for (int k = 0; k <= j; ++k) gg += A(j, k) * A(i, k);
          for (int k = j + 1; k <= l; ++k) gg += A(k, j) * A(i, k);
          e[j] = gg / h;
          f += e[j] * A(i, j);
        }
        double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          double fj = A(i, j);
          double gj = e[j] - hh * fj;
          e[j] = gj;
          for (int k = 0; k <= j; ++k) A(j, k) -= (fj * e[k] + gj * A(i, k));
        }
      }
    } else {
      e[i] = A(i, l);
    }
    d[i] = h;
  }
  d[0] = 0;
  e[0] = 0;
  // --- accumulate the transformation ---
  for (int i = 0; i < n; ++i) {
    int l = i - 1;
    if (d[i] != 0.0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0;
        for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
        for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
      }
    }
    d[i] = A(i, i);
    A(i, i) = 1;
    for (int j = 0; j <= l; ++j) A(j, i) = A(i, j) = 0;
  }
  // --- implicit QL on the tridiagonal matrix (d diagonal, e sub-diagonal) ---
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  if (n > 0) e[n - 1] = 0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::abs(d[m]) + std::abs(d[m + 1]);
        if (std::abs(e[m]) <= 1e-300 || std::abs(e[m]) + dd == dd) break;
      }
      if (m != l) {
        if (iter++ == 60) throw std::runtime_error("QL iteration did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::abs(r) : -std::abs(r)));
        double s = 1, c = 1, p = 0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < n; ++k) {
            double fz = A(k, i + 1);
            A(k, i + 1) = s * A(k, i) + c * fz;
            A(k, i) = c * A(k, i) - s * fz;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0;
      }
    } while (m != l);
  }
  // --- sort ascending ---
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  std::vector<double> v(a);
  w.resize(n);
  for (int k = 0; k < n; ++k) {
    w[k] = d[idx[k]];
    for (int r = 0; r < n; ++r) a[(size_t)r * n + k] = v[(size_t)r * n + idx[k]];
  }
}

// ---------------------------------------------------------------------------------------------
// Restricted SCF driver (the absent Utils::ScfMethod loop; call sites PM6MethodWrapper.cpp:87-89,
// NDDOMethodWrapper.cpp:94-98).  Fock = H + G (NU/FockMatrix.cpp:108-122); Pulay DIIS on the
// commutator FP - PF (orthogonal basis, S = I); aufbau; P = 2 C_occ C_occ^T; converged when
// |dE| < dE_tol and RMSD(dP) < rmsd_tol; after convergence G is rebuilt from the final P
// (FockMatrix::finalize :147-150) and E = E_el + E_rep.
// ---------------------------------------------------------------------------------------------
struct ScfOut {
  double Etot = 0, Eel = 0, Erep = 0;
  std::vector<double> P, eps, C, F;
  int iterations = 0;
  bool converged = false;
};

static void solveSmall(int m, std::vector<double> A, std::vector<double> b, std::vector<double>& x) {
  // Gaussian elimination with partial pivoting (row-major m x m)
  for (int k = 0; k < m; ++k) {
    int piv = k;
    for (int r = k + 1; r < m; ++r)
      if (std::abs(A[r * m + k]) > std::abs(A[piv * m + k])) piv = r;
    if (piv != k) {
      for (int c = 0; c < m; ++c) std::swap(A[k * m + c], A[piv * m + c]);
      std::swap(b[k], b[piv]);
    }
    double dkk = A[k * m + k];
    if (dkk == 0.0) throw std::runtime_error("singular DIIS system");
    for (int r = k + 1; r < m; ++r) {
      double f = A[r * m + k] / dkk;
      for (int c = k; c < m; ++c) A[r * m + c] -= f * A[k * m + c];
      b[r] -= f * b[k];
    }
  }
  x.assign(m, 0.0);
  for (int k = m - 1; k >= 0; --k) {
    double s = b[k];
    for (int c = k + 1; c < m; ++c) s -= A[k * m + c] * x[c];
    x[k] = s / A[k * m + k];
  }
}

static void runScf(Mol& mol, int charge, double dEtol, double rmsdTol, int maxit, int mixer, int diisMax,
                   const double* Pguess, ScfOut& out) {
  int n = mol.nao;
  size_t nn = (size_t)n * n;
  double nEl = mol.nElectronsNeutral - charge;
  if (nEl < 0 || std::lround(nEl) % 2 != 0) throw std::runtime_error("restricted SCF needs an even electron count");
  int nocc = (int)std::lround(nEl) / 2;
  if (nocc > n) throw std::runtime_error("more electrons than orbitals");
  std::vector<double> P(nn), Pold(nn), G(nn), F(nn), Fsym(nn), FP(nn), err(nn);
  if (Pguess) std::copy(Pguess, Pguess + nn, P.begin());
  else densityGuess(mol, nEl, P.data());
  std::vector<std::vector<double>> fHist, eHist;
  std::vector<double> eps, Cm;
  double Eold = 0;
  bool conv = false;
  int it = 0;
  for (it = 1; it <= maxit; ++it) {
    computeG(mol, P.data(), G.data());
    for (size_t k = 0; k < nn; ++k) F[k] = mol.H[k] + G[k];
    double E = electronicEnergy(n, P.data(), mol.H.data(), G.data());
    // full symmetric Fock from the lower triangle
    for (int c = 0; c < n; ++c)
      for (int r = c; r < n; ++r) Fsym[(size_t)c * n + r] = Fsym[(size_t)r * n + c] = F[(size_t)c * n + r];
    if (mixer == 1 && diisMax > 1) {
      // error = F P - P F (symmetric inputs, so err = FP - (FP)^T)
      for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) {
          double s = 0;
          for (int k = 0; k < n; ++k) s += Fsym[(size_t)k * n + r] * P[(size_t)c * n + k];
          FP[(size_t)c * n + r] = s;
        }
      for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) err[(size_t)c * n + r] = FP[(size_t)c * n + r] - FP[(size_t)r * n + c];
      fHist.push_back(Fsym);
      eHist.push_back(err);
      if ((int)fHist.size() > diisMax) {
        fHist.erase(fHist.begin());
        eHist.erase(eHist.begin());
      }
      int m = (int)fHist.size();
      if (m >= 2) {
        int dim = m + 1;
        std::vector<double> B((size_t)dim * dim, 0.0), rhs(dim, 0.0), x;
        for (int i = 0; i < m; ++i)
          for (int j = 0; j <= i; ++j) {
            double s = 0;
            for (size_t k = 0; k < nn; ++k) s += eHist[i][k] * eHist[j][k];
            B[i * dim + j] = B[j * dim + i] = s;
          }
        for (int i = 0; i < m; ++i) B[i * dim + m] = B[m * dim + i] = -1.0;
        rhs[m] = -1.0;
        bool ok = true;
        try {
          solveSmall(dim, B, rhs, x);
        } catch (...) {
          ok = false;
        }
        if (ok) {
          std::fill(Fsym.begin(), Fsym.end(), 0.0);
          for (int i = 0; i < m; ++i)
            for (size_t k = 0; k < nn; ++k) Fsym[k] += x[i] * fHist[i][k];
        }
      }
    }
    std::vector<double> a(nn);
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) a[(size_t)r * n + c] = Fsym[(size_t)c * n + r];
    symmetricEigen(n, a, eps);
    Cm.assign(nn, 0.0);  // col-major: C[k*n + r] = coefficient of AO r in MO k
    for (int k = 0; k < n; ++k)
      for (int r = 0; r < n; ++r) Cm[(size_t)k * n + r] = a[(size_t)r * n + k];
    Pold = P;
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < n; ++r) {
        double s = 0;
        for (int k = 0; k < nocc; ++k) s += Cm[(size_t)k * n + r] * Cm[(size_t)k * n + c];
        P[(size_t)c * n + r] = 2.0 * s;
      }
    double rmsd = 0;
    for (size_t k = 0; k < nn; ++k) rmsd += (P[k] - Pold[k]) * (P[k] - Pold[k]);
    rmsd = std::sqrt(rmsd / (double)nn);
    if (it > 1 && std::abs(E - Eold) < dEtol && rmsd < rmsdTol) {
      conv = true;
      break;
    }
    Eold = E;
  }
  computeG(mol, P.data(), G.data());
  for (size_t k = 0; k < nn; ++k) F[k] = mol.H[k] + G[k];
  out.Eel = electronicEnergy(n, P.data(), mol.H.data(), G.data());
  out.Erep = mol.Erep;
  out.Etot = out.Eel + out.Erep;
  out.P = P;
  out.eps = eps;
  out.C = Cm;
  out.F = F;
  out.iterations = std::min(it, maxit);
  out.converged = conv;
}

}  // namespace orc

// =============================================================================================
// C API (ctypes).  All arrays are caller-owned.  Matrices are n x n column-major.  Returns 0 on
// success, -1 on error (message via orc_last_error).
// =============================================================================================
namespace {
thread_local std::string tError;
std::map<std::string, orc::Method*> gMethods;
template <class Fn>
int guarded(Fn&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    tError = e.what();
    return -1;
  }
}
}  // namespace

extern "C" {
const char* orc_last_error() { return tError.c_str(); }
void orc_set_ev_per_hartree(double v) {
  orc::gEvPerHartree = v;
  for (auto& kv : gMethods) kv.second->atoms.clear();
}
double orc_get_ev_per_hartree() { return orc::gEvPerHartree; }
void orc_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}
int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void* orc_load(const char* path) {
  void* h = nullptr;
  if (guarded([&] {
        auto it = gMethods.find(path);
        if (it != gMethods.end()) {
          h = it->second;
          return;
        }
        auto* m = new orc::Method();
        m->raw = orc::loadParams(path);
        gMethods[path] = m;
        h = m;
      }) != 0)
    return nullptr;
  return h;
}

int orc_atom_params(void* h, int Z, double* D5, double* rho8, double* oneC58, double* misc /* nao,l,core,pcore,pcoreSpec,U3,beta3,alpha */) {
  return guarded([&] {
    const orc::AtomPar& a = orc::atomPar(*static_cast<orc::Method*>(h), Z);
    for (int i = 0; i < 5; ++i) D5[i] = a.D[i];
    for (int i = 0; i < 8; ++i) rho8[i] = a.rho[i];
    for (int i = 0; i < 58; ++i) oneC58[i] = i < a.n1c ? a.oneC[i] : 0.0;
    double m[13] = {(double)a.nao, (double)a.l, a.core, a.pcore, a.pcoreSpecified ? 1.0 : 0.0, a.U[0], a.U[1], a.U[2],
                    a.beta[0], a.beta[1], a.beta[2], a.alpha, (double)a.n1c};
    for (int i = 0; i < 13; ++i) misc[i] = m[i];
  });
}

double orc_one_center(void* h, int Z, int i, int j, int k, int l) {
  double v = 0;
  guarded([&] {
    const orc::AtomPar& a = orc::atomPar(*static_cast<orc::Method*>(h), Z);
    v = orc::oneC(a, i, j, k, l);
  });
  return v;
}

// Parameter functions on explicit inputs, as the reference's unit tests call them.
int orc_charge_separations(int ns, int np, int nd, double zs, double zp, double zd, double* D5) {
  return guarded([&] { orc::chargeSeparations(ns, np, nd, zs, zp, zd, nd > 0, D5); });
}
int orc_klopman(int order, const double* in12, double* rho8) {
  return guarded([&] { orc::klopman(order, in12, rho8); });
}
// Slater-Condon parameters of element Z from internal exponents zint (s, p, d) with F0sd / G2sd
// (Hartree, 0 = computed), and the n1c one-centre integrals from them and g5 (see oneCenterIntegrals).
int orc_one_center_explicit(int Z, int n1c, const double* zint3, double f0sd, double g2sd, const double* g5, double* sc17,
                            double* out58) {
  return guarded([&] {
    double sc[17] = {};
    if (n1c == 58) orc::slaterCondon(Z, zint3, f0sd, g2sd, sc);
    orc::oneCenterIntegrals(n1c, g5, sc, out58);
    for (int i = 0; i < 17; ++i) sc17[i] = sc[i];
  });
}

double orc_local_integral(int t1, int t2, double R, const double* D1, const double* D2, const double* r1, const double* r2) {
  orc::buildLocalTerms();
  return orc::localIntegral(t1, t2, R, D1, D2, r1, r2);
}
int orc_local_matrix(int l1, int l2, double R, const double* D1, const double* D2, const double* r1, const double* r2, double* out) {
  return guarded([&] {
    orc::buildLocalTerms();
    orc::localMatrix(l1, l2, R, D1, D2, r1, r2, out);
  });
}
int orc_rotation(const double* Rab, double* r81) {
  double r[9][9];
  orc::rotationCoefficients(Rab, r);
  for (int i = 0; i < 81; ++i) r81[i] = r[i / 9][i % 9];
  return 0;
}
int orc_global_block(int l1, int l2, const double* Rab, const double* D1, const double* D2, const double* r1, const double* r2, double* out) {
  return guarded([&] { orc::globalBlock(l1, l2, Rab, D1, D2, r1, r2, out); });
}

int orc_nao(void* h, int nat, const int* Z) {
  int n = -1;
  if (guarded([&] {
        n = 0;
        for (int i = 0; i < nat; ++i) n += orc::atomPar(*static_cast<orc::Method*>(h), Z[i]).nao;
      }) != 0)
    return -1;
  return n;
}

// Density-independent quantities: S, H (lower triangle, col-major), E_rep.
int orc_density_independent(void* h, int nat, const int* Z, const double* xyz, double* S, double* H, double* Erep) {
  return guarded([&] {
    orc::Mol mol;
    orc::setupMol(mol, *static_cast<orc::Method*>(h), nat, Z, xyz);
    orc::computeTwoCenter(mol);
    orc::computeOverlap(mol);
    orc::computeHcore(mol);
    size_t nn = (size_t)mol.nao * mol.nao;
    if (S) std::copy(mol.S.begin(), mol.S.end(), S);
    if (H) std::copy(mol.H.begin(), mol.H.end(), H);
    (void)nn;
    if (Erep) *Erep = orc::computeRepulsion(mol);
  });
}

// G(P) for a full symmetric P (lower triangle of G returned), plus the electronic energy.
int orc_gmatrix(void* h, int nat, const int* Z, const double* xyz, const double* P, double* G, double* Eel) {
  return guarded([&] {
    orc::Mol mol;
    orc::setupMol(mol, *static_cast<orc::Method*>(h), nat, Z, xyz);
    orc::computeTwoCenter(mol);
    orc::computeOverlap(mol);
    orc::computeHcore(mol);
    orc::computeG(mol, P, G);
    if (Eel) *Eel = orc::electronicEnergy(mol.nao, P, mol.H.data(), G);
  });
}

// Unrestricted G^alpha, G^beta (lower triangles) for full symmetric Pa, Pb, plus the electronic energy.
int orc_gmatrix_unrestricted(void* h, int nat, const int* Z, const double* xyz, const double* Pa, const double* Pb,
                             double* Ga, double* Gb, double* Eel) {
  return guarded([&] {
    orc::Mol mol;
    orc::setupMol(mol, *static_cast<orc::Method*>(h), nat, Z, xyz);
    orc::computeTwoCenter(mol);
    orc::computeOverlap(mol);
    orc::computeHcore(mol);
    orc::computeGUnrestricted(mol, Pa, Pb, Ga, Gb);
    if (Eel) *Eel = orc::electronicEnergyUnrestricted(mol.nao, Pa, Pb, mol.H.data(), Ga, Gb);
  });
}

// nvec CIS AO Fock matrices (n x n each, consecutive) from nvec pairs (DJ_k, DK_k).
int orc_cis_fock(void* h, int nat, const int* Z, const double* xyz, int nvec, const double* DJ, const double* DK, double cJ,
                 double cK, double* F) {
  return guarded([&] {
    orc::Mol mol;
    orc::setupMol(mol, *static_cast<orc::Method*>(h), nat, Z, xyz);
    orc::computeTwoCenter(mol);
    const size_t nn = (size_t)mol.nao * mol.nao;
#pragma omp parallel for schedule(dynamic)
    for (int v = 0; v < nvec; ++v) orc::computeCisFock(mol, DJ + v * nn, DK + v * nn, cJ, cK, F + v * nn);
  });
}

int orc_density_guess(void* h, int nat, const int* Z, const double* xyz, int charge, double* P) {
  return guarded([&] {
    orc::Mol mol;
    orc::setupMol(mol, *static_cast<orc::Method*>(h), nat, Z, xyz);
    orc::computeOverlap(mol);
    orc::densityGuess(mol, mol.nElectronsNeutral - charge, P);
  });
}

int orc_eigh(int n, const double* A /* col-major symmetric */, double* w, double* V /* col-major */) {
  return guarded([&] {
    std::vector<double> a((size_t)n * n), ww;
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) a[(size_t)r * n + c] = A[(size_t)c * n + r];
    orc::symmetricEigen(n, a, ww);
    for (int k = 0; k < n; ++k) {
      w[k] = ww[k];
      for (int r = 0; r < n; ++r) V[(size_t)k * n + r] = a[(size_t)r * n + k];
    }
  });
}

// Full restricted SCF.  energies[3] = {E_total, E_el, E_rep}; timing[2] = {setup s, scf s}.
int orc_scf(void* h, int nat, const int* Z, const double* xyz, int charge, double dE, double rmsd, int maxit,
            int mixer, int diisMax, const double* Pguess, double* energies, double* P, double* eps, double* C,
            int* iterations, int* converged) {
  return guarded([&] {
    orc::Mol mol;
    orc::setupMol(mol, *static_cast<orc::Method*>(h), nat, Z, xyz);
    orc::computeTwoCenter(mol);
    orc::computeOverlap(mol);
    orc::computeHcore(mol);
    mol.Erep = orc::computeRepulsion(mol);
    orc::ScfOut out;
    orc::runScf(mol, charge, dE, rmsd, maxit, mixer, diisMax, Pguess, out);
    energies[0] = out.Etot;
    energies[1] = out.Eel;
    energies[2] = out.Erep;
    if (P) std::copy(out.P.begin(), out.P.end(), P);
    if (eps) std::copy(out.eps.begin(), out.eps.end(), eps);
    if (C) std::copy(out.C.begin(), out.C.end(), C);
    *iterations = out.iterations;
    *converged = out.converged ? 1 : 0;
  });
}
}  // extern "C"
