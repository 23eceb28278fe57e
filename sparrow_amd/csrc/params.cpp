// Parameter loading and runtime element data.
//
// Reference behaviour (src/Sparrow/Sparrow/Implementations/Nddo/Utils/ParameterUtils):
//   processAtomicParameters ...... RawParameterProcessor.cpp:67-99 (eV->Eh, 1/A->1/bohr)
//   charge separations D ......... ChargeSeparationParameter.cpp:24-51 (n = min(ns,np) for sp)
//   Klopman-Ohno rho ............. KlopmanParameter.cpp:26-92, RawParameterProcessor.cpp:126-159
//                                  (hpp >= 0.1 eV clamp for sp elements: REPRODUCE_PM6_ERRORS)
//   Slater-Condon parameters ..... SlaterCondonParameters.cpp:43-82, ../IntegralsEvaluationUtils/
//                                  oneCenterSlaterIntegral.cpp:38-109
//   one-centre integrals ......... RawParameterProcessor.cpp:42-65,
//                                  ../IntegralsEvaluationUtils/oneCenterTwoElectronIntegrals.cpp:35-87
//   element data ................. PrincipalQuantumNumbers.cpp:25-167
//   diatomic exponent units ...... RawParameterProcessor.cpp:101-111
#include "params.hpp"

#include <array>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace sparrow_amd {

namespace {

int principalS(int Z) { return Z <= 2 ? 1 : Z <= 9 ? 2 : Z <= 17 ? 3 : Z <= 35 ? 4 : Z <= 53 ? 5 : Z <= 83 ? 6 : 7; }
int principalP(int Z) { return Z <= 10 ? 2 : Z <= 18 ? 3 : Z <= 36 ? 4 : Z <= 54 ? 5 : Z <= 83 ? 6 : 7; }
int principalD(int Z) { return Z <= 32 ? 3 : Z <= 50 ? 4 : Z <= 83 ? 5 : 6; }

int aoCount(int Z, bool spd) {
  if (Z == 1) return 1;
  if (!spd) return 4;
  // spd basis: d shells for Al-Cl, Sc-Cu, As, Br, Y-Ag, Sb, I, La-Au
  static const int dRanges[][2] = {{13, 17}, {21, 29}, {33, 33}, {35, 35}, {39, 47}, {51, 51}, {53, 53}, {57, 79}};
  for (const auto& r : dRanges)
    if (Z >= r[0] && Z <= r[1]) return 9;
  return Z <= 83 ? 4 : 0;
}

double coreChargeOf(int Z) {
  if (Z == 10 || Z == 18 || Z == 36 || Z == 54) return 6;
  static const int shells[] = {86, 78, 68, 54, 46, 36, 28, 18, 10, 2};
  for (int c : shells)
    if (Z > c) return Z - c;
  return Z;
}

double factorial(int n) {
  double r = 1;
  for (int k = 2; k <= n; ++k) r *= k;
  return r;
}

// radial overlap-type factor A(n1,n2,z1,z2,L) of the charge-separation formulas
double sepFactor(int n1, int n2, double z1, double z2, int L) {
  return std::pow(2 * z1, n1 + 0.5) * std::pow(2 * z2, n2 + 0.5) * std::pow(z1 + z2, -1.0 * (n1 + n2 + L + 1)) /
         std::sqrt(factorial(2 * n1) * factorial(2 * n2)) * factorial(n1 + n2 + L);
}

// Newton iteration for the Klopman-Ohno additive term of a dipole (order 1) or quadrupole
// (order 2) charge distribution with separation D: solve E(rho) = shift from rho0 = 0.001.
double klopmanRoot(int order, double D, double shift) {
  double r = 0.001, prev = 0;
  while (std::abs(r - prev) > 1e-8) {
    double r2 = r * r, d2 = D * D;
    double f, df;
    if (order == 1) {
      double q = std::sqrt(r2 + d2);
      f = 1.0 / r - 1.0 / q - shift;
      df = -1.0 / r2 + r / (q * q * q);
    } else {
      double q1 = std::sqrt(r2 + d2), q2 = std::sqrt(r2 + 2 * d2);
      f = 1.0 / r - 2.0 / q1 + 1.0 / q2 - shift;
      df = -1.0 / r2 + 2 * r / (q1 * q1 * q1) - r / (q2 * q2 * q2);
    }
    prev = r;
    r = r - f / df;
  }
  return r;
}

// Radial Slater integral R^l(ab,cd) over Slater orbitals with principal numbers n[] and
// exponents e[] (closed form for integer n).
double slaterRadial(int l, const int n[4], const double e[4]) {
  int n1 = n[0] + n[1], n2 = n[2] + n[3];
  double e1 = e[0] + e[1], e2 = e[2] + e[3];
  double norm = std::pow(2 * e[2], n[2] + 0.5) * std::pow(2 * e[3], n[3] + 0.5) / std::sqrt(factorial(2 * n[2]) * factorial(2 * n[3])) *
                std::pow(2 * e[0], n[0] + 0.5) * std::pow(2 * e[1], n[1] + 0.5) / std::sqrt(factorial(2 * n[0]) * factorial(2 * n[1]));
  double c = norm * factorial(n2 + l) / std::pow(e2, n2 + l + 1);
  double t1 = factorial(n1 - l - 1) / std::pow(e1, n1 - l);
  double t2 = 0, t3 = 0;
  for (int k = 1; k <= n2 + l + 1; ++k)
    t2 += std::pow(e2, n2 + l - k + 1) / factorial(n2 + l - k + 1) * factorial(n1 + n2 - k) / std::pow(e1 + e2, n1 + n2 - k + 1);
  for (int k = 1; k <= n2 - l; ++k)
    t3 += std::pow(e2, n2 + l - k + 1) * factorial(n2 - l - 1) * factorial(n1 + n2 - k) /
          (factorial(n2 + l) * factorial(n2 - l - k) * std::pow(e1 + e2, n1 + n2 - k + 1));
  return c * (t1 - t2 + t3);
}

}  // namespace

ParameterSet ParameterSet::load(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open NDDO parameter table " + path);
  ParameterSet ps;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream in(line);
    std::string tag;
    in >> tag;
    if (tag == "method") {
      std::string kw1, basis, kw2;
      int diat = 0;
      in >> ps.name_ >> kw1 >> basis >> kw2 >> diat;
      ps.spd_ = basis == "spd";
      ps.diatomic_ = diat != 0;
      static const std::pair<const char*, MethodKind> names[] = {
          {"MNDO", MethodKind::MNDO}, {"AM1", MethodKind::AM1}, {"PM3", MethodKind::PM3}, {"RM1", MethodKind::RM1}, {"PM6", MethodKind::PM6}};
      bool known = false;
      for (auto& nm : names)
        if (ps.name_ == nm.first) {
          ps.kind_ = nm.second;
          known = true;
        }
      if (!known) throw std::runtime_error("unknown NDDO method " + ps.name_);
    } else if (tag == "atom") {
      RawElement e;
      in >> e.Z;
      for (double* v : {e.U, e.beta, e.zeta, e.zetaInt}) in >> v[0] >> v[1] >> v[2];
      in >> e.gss >> e.gpp >> e.gsp >> e.gp2 >> e.hsp >> e.pcore >> e.f0sd >> e.g2sd >> e.alpha;
      int ng = 0;
      in >> ng;
      for (int k = 0; k < ng; ++k) {
        std::array<double, 3> g;
        in >> g[0] >> g[1] >> g[2];
        e.gauss.push_back(g);
      }
      ps.raw_[e.Z] = e;
    } else if (tag == "pair") {
      int a, b;
      double ex, fac;
      in >> a >> b >> ex >> fac;
      ps.pairs_[{std::min(a, b), std::max(a, b)}] = {ex, fac};
    } else if (tag == "gto") {
      int Z, l, np;
      in >> Z >> l >> np;
      auto it = ps.raw_.find(Z);
      if (it == ps.raw_.end() || np != 6 || it->second.nShell >= 3) throw std::runtime_error("malformed gto record in " + path);
      RawElement& e = it->second;
      e.shellL[e.nShell] = l;
      for (int k = 0; k < np; ++k) in >> e.shE[e.nShell][k] >> e.shC[e.nShell][k];
      ++e.nShell;
    }
  }
  if (ps.name_.empty()) throw std::runtime_error("parameter table without a method line: " + path);
  return ps;
}

bool ParameterSet::hasPair(int Z1, int Z2) const { return pairs_.count({std::min(Z1, Z2), std::max(Z1, Z2)}) != 0; }

DiatomicDev ParameterSet::diatomicParams(int Z1, int Z2) const {
  DiatomicDev d{};
  auto it = pairs_.find({std::min(Z1, Z2), std::max(Z1, Z2)});
  if (it == pairs_.end()) return d;
  bool xh = (Z1 == 1 && (Z2 == 6 || Z2 == 7 || Z2 == 8)) || (Z2 == 1 && (Z1 == 6 || Z1 == 7 || Z1 == 8));
  // X-H exponents are tabulated in A^-2, all others in A^-1
  d.alpha = xh ? it->second.first * kAngstromPerBohr * kAngstromPerBohr : it->second.first * kAngstromPerBohr;
  d.x = it->second.second;
  d.kind = xh ? 1 : (Z1 == 6 && Z2 == 6) ? 2 : ((Z1 == 14 && Z2 == 8) || (Z1 == 8 && Z2 == 14)) ? 3 : 0;
  return d;
}

const ElementDev& ParameterSet::element(int Z, const ModelTables& tables) {
  auto cached = runtime_.find(Z);
  if (cached != runtime_.end()) return cached->second;
  auto rit = raw_.find(Z);
  if (rit == raw_.end()) throw std::runtime_error("NDDO parameters do not exist for element Z=" + std::to_string(Z));
  const RawElement& r = rit->second;
  const double eh = 1.0 / kEvPerHartree;
  ElementDev e{};
  e.Z = Z;
  e.w1c = (int)std::distance(raw_.begin(), rit);
  e.nao = aoCount(Z, spd_);
  e.l = e.nao == 1 ? 0 : e.nao == 4 ? 1 : 2;
  e.core = coreChargeOf(Z);
  e.zcbrt = std::pow((double)Z, 1.0 / 3.0);
  for (int k = 0; k < 3; ++k) {
    e.U[k] = r.U[k] * eh;
    e.beta[k] = r.beta[k] * eh;
  }
  e.alpha = r.alpha * kAngstromPerBohr;

  // Slater-Condon parameters (only d elements compute a new set; the processor keeps the last one)
  enum { F0ss, F0pp, F0dd, F0sp, F0sd, F0pd, F2pp, F2dd, F2pd, F4dd, G1sp, G1pd, G2sd, G3pd, R1sppd, R2sdpp, R2sddd };
  double sc[17];
  for (int k = 0; k < 17; ++k) sc[k] = lastSlaterCondon_[k];
  if (e.nao == 9) {
    const int nq[3] = {principalS(Z), principalP(Z), principalD(Z)};
    // (l, shells a b c d) of the 17 parameters in sc_t order
    static const int def[17][5] = {{0, 0, 0, 0, 0}, {0, 1, 1, 1, 1}, {0, 2, 2, 2, 2}, {0, 0, 0, 1, 1}, {0, 0, 0, 2, 2}, {0, 1, 1, 2, 2},
                                   {2, 1, 1, 1, 1}, {2, 2, 2, 2, 2}, {2, 1, 1, 2, 2}, {4, 2, 2, 2, 2}, {1, 0, 1, 1, 0}, {1, 1, 2, 2, 1},
                                   {2, 0, 2, 2, 0}, {3, 1, 2, 2, 1}, {1, 0, 1, 1, 2}, {2, 0, 2, 1, 1}, {2, 0, 2, 2, 2}};
    for (int k = 0; k < 17; ++k) {
      int n[4];
      double x[4];
      for (int q = 0; q < 4; ++q) {
        n[q] = nq[def[k][q + 1]];
        x[q] = r.zetaInt[def[k][q + 1]];
      }
      sc[k] = slaterRadial(def[k][0], n, x);
    }
    if (r.f0sd != 0) sc[F0sd] = r.f0sd * eh;
    if (r.g2sd != 0) sc[G2sd] = r.g2sd * eh;
    for (int k = 0; k < 17; ++k) lastSlaterCondon_[k] = sc[k];
  }

  // charge separations
  for (double& d : e.D) d = 0;
  if (e.nao >= 4) {
    int ns = principalS(Z), np = principalP(Z), n = std::min(ns, np);
    e.D[kSp1] = sepFactor(n, n, r.zeta[0], r.zeta[1], 1) / std::sqrt(3.0);
    e.D[kPp2] = std::sqrt(sepFactor(n, n, r.zeta[1], r.zeta[1], 2)) / std::sqrt(5.0);
    if (e.nao == 9) {
      int nd = principalD(Z);
      e.D[kPd1] = sepFactor(np, nd, r.zeta[1], r.zeta[2], 1) / std::sqrt(5.0);
      e.D[kSd2] = std::sqrt(sepFactor(ns, nd, r.zeta[0], r.zeta[2], 2)) / std::pow(15.0, 0.25);
      e.D[kDd2] = std::sqrt(sepFactor(nd, nd, r.zeta[2], r.zeta[2], 2)) / std::sqrt(7.0);
    }
  }
  // Klopman-Ohno additive terms
  for (double& v : e.rho) v = 0;
  e.rho[kSs0] = 1.0 / (2.0 * r.gss * eh);
  if (e.nao >= 4) {
    double hpp = 0.5 * (r.gpp - r.gp2) * eh;
    if (e.nao == 4) hpp = std::max(0.1 * eh, hpp);
    e.rho[kPp0] = e.rho[kSs0];
    e.rho[kSp1] = klopmanRoot(1, e.D[kSp1], 4 * r.hsp * eh);
    e.rho[kPp2] = klopmanRoot(2, e.D[kPp2], 8 * hpp);
    if (e.nao == 9) {
      e.rho[kDd0] = 1.0 / (2.0 * sc[F0dd]);
      e.rho[kPd1] = klopmanRoot(1, e.D[kPd1], 16.0 / 15.0 * sc[G1pd]);
      e.rho[kSd2] = klopmanRoot(2, e.D[kSd2], 8.0 / 5.0 * sc[G2sd]);
      e.rho[kDd2] = klopmanRoot(2, e.D[kDd2], 24.0 / 49.0 * sc[F2dd]);
    }
  }
  e.pcoreSpecified = r.pcore != 0;
  e.pcore = r.pcore != 0 ? r.pcore : e.rho[kSs0];

  // one-centre two-electron integrals
  e.n1c = e.nao == 1 ? 1 : e.nao == 4 ? 6 : 58;
  bool given[58] = {};
  auto put = [&](int i, int j, int k, int l, double v) {
    int idx = tables.oneCIndex[((i * 9 + j) * 9 + k) * 9 + l];
    e.oneC[idx] = v;
    given[idx] = true;
  };
  if (r.gss != 0) put(0, 0, 0, 0, r.gss * eh);
  if (r.gsp != 0) put(0, 0, 1, 1, r.gsp * eh);
  if (r.gpp != 0) put(1, 1, 1, 1, r.gpp * eh);
  if (r.gp2 != 0) put(1, 1, 2, 2, r.gp2 * eh);
  if (r.hsp != 0) put(0, 1, 0, 1, r.hsp * eh);
  if (e.n1c >= 6 && given[3] && given[4]) put(1, 2, 1, 2, 0.5 * (e.oneC[3] - e.oneC[4]));  // <pp'|pp'>
  for (int k = 0; k < e.n1c; ++k)
    if (!given[k]) {
      double v = 0;
      for (int c = 0; c < 3; ++c) v += tables.oneCCoef[k][c] * sc[tables.oneCTerm[k][c]];
      e.oneC[k] = v;
    }

  // core-core Gaussians: a [eV*A] -> [Eh*bohr], b [A^-2] -> [bohr^-2], c [A] -> [bohr]
  e.nGauss = (int)std::min<size_t>(r.gauss.size(), 4);
  for (int k = 0; k < e.nGauss; ++k) {
    e.gauss[k][0] = r.gauss[k][0] * kBohrPerAngstrom * eh;
    e.gauss[k][1] = r.gauss[k][1] * kAngstromPerBohr * kAngstromPerBohr;
    e.gauss[k][2] = r.gauss[k][2] * kBohrPerAngstrom;
  }
  // STO-6G shells
  e.nShell = r.nShell;
  for (int s = 0; s < r.nShell; ++s) {
    // the overlap kernels take shell s as the one of angular momentum s (s, p, d in that order)
    if (r.shellL[s] != s)
      throw std::runtime_error("STO-6G shells of element Z=" + std::to_string(Z) + " are not listed as s, p, d");
    e.shellL[s] = r.shellL[s];
    for (int k = 0; k < 6; ++k) {
      e.shE[s][k] = r.shE[s][k];
      e.shC[s][k] = r.shC[s][k];
    }
  }
  return runtime_[Z] = e;
}

std::vector<double> ParameterSet::oneCenterExchangeTable(const ModelTables& tables) {
  std::vector<double> w(raw_.size() * (size_t)kOneCenterBlock, 0.0);
  for (const auto& kv : raw_) {
    const ElementDev& e = element(kv.first, tables);
    const int na = e.nao, na2 = na * na;
    double* blk = w.data() + (size_t)e.w1c * kOneCenterBlock;
    auto val = [&](int i, int j, int k, int l) {
      const int c = tables.oneCIndex[((i * 9 + j) * 9 + k) * 9 + l];
      return c < e.n1c ? e.oneC[c] : 0.0;
    };
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < na; ++j)
        for (int k = 0; k < na; ++k)
          for (int l = 0; l < na; ++l) blk[(k * na + l) * na2 + i * na + j] = val(i, k, j, l);
  }
  return w;
}

std::vector<double> ParameterSet::oneCenterTable(const ModelTables& tables) {
  std::vector<double> w(raw_.size() * (size_t)kOneCenterBlock, 0.0);
  for (const auto& kv : raw_) {
    const ElementDev& e = element(kv.first, tables);
    const int na = e.nao, na2 = na * na;
    double* blk = w.data() + (size_t)e.w1c * kOneCenterBlock;
    auto val = [&](int i, int j, int k, int l) {
      const int c = tables.oneCIndex[((i * 9 + j) * 9 + k) * 9 + l];
      return c < e.n1c ? e.oneC[c] : 0.0;
    };
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < na; ++j)
        for (int k = 0; k < na; ++k)
          for (int l = 0; l < na; ++l) blk[(k * na + l) * na2 + i * na + j] = val(i, j, k, l) - 0.5 * val(i, k, j, l);
  }
  return w;
}

}  // namespace sparrow_amd
