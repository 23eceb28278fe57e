// Host construction of the static NDDO model tables that the kernels read:
//   * charge-distribution (orbital pair) indices and the block structure of the pair rotation;
//   * the multipole expansion of every local two-centre integral, flattened per (lA, lB) class
//     into "multipole-multipole instances" (point-charge sums) and "entries" (linear combinations
//     of instances);
//   * the one-centre two-electron integral index table (9^4 -> 58 unique integrals).
//
// Reference behaviour restated here (file:line into src/Sparrow/Sparrow/Implementations/Nddo/Utils):
//   point charges of each multipole ............. IntegralsEvaluationUtils/ChargesInMultipoles.cpp:27-125
//   multipoles of each charge distribution ...... IntegralsEvaluationUtils/Local2c2eIntegralCalculator.cpp:142-292
//   term selection (equal m, PM6 M20xM20 error) . IntegralsEvaluationUtils/Local2c2eIntegralCalculator.cpp:53-140
//   point-charge interaction ..................... IntegralsEvaluationUtils/MultipoleMultipoleTerm.h:57-64
//   one-centre index/expression table ............ IntegralsEvaluationUtils/OneCenterTwoElectronCalculator.cpp:31-348
#include "tables.hpp"

#include <array>
#include <cmath>
#include <map>
#include <stdexcept>
#include <tuple>

namespace sparrow_amd {

namespace {

enum Orb { S = 0, X, Y, Z, X2Y2, XZ, Z2, YZ, XY };
enum Pole { M00 = 0, Qxx, Qyy, Qzz, M1m1, M10, M11, M2m2, M2m1, M20, M21, M22, Qzx };

const int kPoleM[13] = {0, 0, 0, 0, -1, 0, 1, -2, -1, 0, 1, 2, 2};
const int kPoleL[13] = {0, 2, 2, 2, 1, 1, 1, 2, 2, 2, 2, 2, 2};

struct PointCharge {
  double x, y, z, q;
};

std::vector<PointCharge> chargesOf(int pole) {
  const double r2 = std::sqrt(2.0);
  switch (pole) {
    case M00: return {{0, 0, 0, 1}};
    case Qxx: return {{-2, 0, 0, .25}, {2, 0, 0, .25}, {0, 0, 0, -.5}};
    case Qyy: return {{0, -2, 0, .25}, {0, 2, 0, .25}, {0, 0, 0, -.5}};
    case Qzz: return {{0, 0, -2, .25}, {0, 0, 2, .25}, {0, 0, 0, -.5}};
    case M1m1: return {{0, -1, 0, -.5}, {0, 1, 0, .5}};
    case M10: return {{0, 0, -1, -.5}, {0, 0, 1, .5}};
    case M11: return {{-1, 0, 0, -.5}, {1, 0, 0, .5}};
    case M2m2: return {{1, 1, 0, .25}, {-1, -1, 0, .25}, {1, -1, 0, -.25}, {-1, 1, 0, -.25}};
    case M2m1: return {{0, 1, 1, .25}, {0, -1, -1, .25}, {0, 1, -1, -.25}, {0, -1, 1, -.25}};
    case M21: return {{1, 0, 1, .25}, {-1, 0, -1, .25}, {1, 0, -1, -.25}, {-1, 0, 1, -.25}};
    case M20:
      return {{0, 0, r2, .25}, {0, 0, -r2, .25}, {r2, 0, 0, -.375}, {-r2, 0, 0, -.375}, {0, r2, 0, .125}, {0, -r2, 0, .125}};
    case M22: return {{r2, 0, 0, .25}, {-r2, 0, 0, .25}, {0, r2, 0, -.25}, {0, -r2, 0, -.25}};
    case Qzx: return {{0, 0, r2, .25}, {0, 0, -r2, .25}, {r2, 0, 0, -.25}, {-r2, 0, 0, -.25}};
  }
  return {};
}

int orbL(int o) { return o == S ? 0 : o <= Z ? 1 : 2; }

struct Weighted {
  double f;
  int pole;
};

// Multipole content of each charge distribution.  `dBasis` selects the convention used when
// either side of the integral involves a d-containing distribution (index >= 10).
std::vector<Weighted> polesOf(int t, bool dBasis) {
  const double k = 1 / std::sqrt(3.0);
  switch (t) {
    case 0: return {{1, M00}};
    case 1: return {{1, M11}};
    case 2: return dBasis ? std::vector<Weighted>{{1, M00}, {-2.0 / 3.0, M20}, {1, M22}} : std::vector<Weighted>{{1, M00}, {1, Qxx}};
    case 3: return {{1, M1m1}};
    case 4: return {{1, M2m2}};
    case 5: return dBasis ? std::vector<Weighted>{{1, M00}, {-2.0 / 3.0, M20}, {-1, M22}} : std::vector<Weighted>{{1, M00}, {1, Qyy}};
    case 6: return {{1, M10}};
    case 7: return {{1, M21}};
    case 8: return {{1, M2m1}};
    case 9: return dBasis ? std::vector<Weighted>{{1, M00}, {4.0 / 3.0, M20}} : std::vector<Weighted>{{1, M00}, {1, Qzz}};
    case 10: return {{2 * k, M20}};
    case 11: return {{1, M21}};
    case 12: return {{1, M2m1}};
    case 13: return {{1, M22}};
    case 14: return {{1, M2m2}};
    case 15: return {{-k, M11}};
    case 16: return {{1, M10}};
    case 17: return {{1, M11}};
    case 18: return {{1, M1m1}};
    case 19: return {{-k, M1m1}};
    case 20: return {{1, M10}};
    case 21: return {{-1, M1m1}};
    case 22: return {{1, M11}};
    case 23: return {{2 * k, M10}};
    case 24: return {{1, M11}};
    case 25: return {{1, M1m1}};
    case 26: return {{1, M00}, {4.0 / 3.0, M20}};
    case 27: return {{k, M21}};
    case 28: return {{k, M2m1}};
    case 29: return {{-2 * k, M22}};
    case 30: return {{-2 * k, M2m2}};
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

// (l of orbital 1, l of orbital 2, l of multipole) -> MultipolePair index
int poleType(int l1, int l2, int l) {
  if (l1 > l2) std::swap(l1, l2);
  if (l1 == 0 && l2 == 1 && l == 1) return kSp1;
  if (l1 == 1 && l2 == 2 && l == 1) return kPd1;
  if (l1 == 1 && l2 == 1 && l == 2) return kPp2;
  if (l1 == 0 && l2 == 2 && l == 2) return kSd2;
  if (l1 == 2 && l2 == 2 && l == 2) return kDd2;
  if (l1 == 0 && l2 == 0 && l == 0) return kSs0;
  if (l1 == 1 && l2 == 1 && l == 0) return kPp0;
  if (l1 == 2 && l2 == 2 && l == 0) return kDd0;
  throw std::logic_error("no multipole pair type");
}

}  // namespace

const int kPairOrbitals[40][2] = {
    {S, S}, {S, X}, {X, X}, {S, Y}, {X, Y}, {Y, Y}, {S, Z}, {X, Z}, {Y, Z}, {Z, Z},
    {S, Z2}, {S, XZ}, {S, YZ}, {S, X2Y2}, {S, XY}, {X, Z2}, {X, XZ}, {X, X2Y2}, {X, XY},
    {Y, Z2}, {Y, YZ}, {Y, X2Y2}, {Y, XY}, {Z, Z2}, {Z, XZ}, {Z, YZ}, {Z2, Z2}, {Z2, XZ},
    {Z2, YZ}, {Z2, X2Y2}, {Z2, XY}, {XZ, XZ}, {XZ, YZ}, {XZ, X2Y2}, {XZ, XY}, {YZ, YZ},
    {YZ, X2Y2}, {YZ, XY}, {X2Y2, X2Y2}, {XY, XY}};

ModelTables buildModelTables() {
  ModelTables t;
  // ---- orbital pair index (TwoElectronIntegralIndexes.cpp:17-117) ----
  for (auto& r : t.pairIndex)
    for (auto& c : r) c = kNoPair;
  for (int p = 0; p < 40; ++p) {
    t.pairIndex[kPairOrbitals[p][0]][kPairOrbitals[p][1]] = p;
    t.pairIndex[kPairOrbitals[p][1]][kPairOrbitals[p][0]] = p;
    t.pairOrb[p][0] = kPairOrbitals[p][0];
    t.pairOrb[p][1] = kPairOrbitals[p][1];
  }
  // ---- rotation blocks: a global pair only mixes with local pairs of the same (l1,l2) ----
  {
    int blockOf[40];
    auto cls = [](int p) {
      int a = orbL(kPairOrbitals[p][0]), b = orbL(kPairOrbitals[p][1]);
      if (a > b) std::swap(a, b);
      return a * 3 + b;  // 0 ss, 1 sp, 2 sd, 4 pp, 5 pd, 8 dd
    };
    int order[6] = {0, 1, 4, 2, 5, 8};
    int pos = 0;
    for (int bi = 0; bi < 6; ++bi) {
      t.blockStart[bi] = pos;
      for (int p = 0; p < 40; ++p)
        if (cls(p) == order[bi]) {
          t.blockMembers[pos++] = p;
          blockOf[p] = bi;
        }
      t.blockLen[bi] = pos - t.blockStart[bi];
    }
    for (int p = 0; p < 40; ++p) t.pairBlock[p] = blockOf[p];
  }
  // ---- local-integral classes ----
  for (int lA = 0; lA < 3; ++lA)
    for (int lB = 0; lB < 3; ++lB) {
      int dA = lA == 0 ? 1 : lA == 1 ? 10 : 40, dB = lB == 0 ? 1 : lB == 1 ? 10 : 40;
      LocalClassDev c{};
      c.dA = dA;
      c.dB = dB;
      c.inst0 = (int)t.instances.size();
      c.entry0 = (int)t.entries.size();
      std::map<std::tuple<int, int, int, int>, int> instId;  // (m1, m2, p1, p2) -> instance
      auto instance = [&](int m1, int m2, int p1, int p2) {
        auto key = std::make_tuple(m1, m2, p1, p2);
        auto it = instId.find(key);
        if (it != instId.end()) return it->second;
        MMInstanceDev in{};
        in.p1 = p1;
        in.p2 = p2;
        in.charge0 = (int)t.charges.size();
        // merge point-charge pairs at identical separations (exact: same geometry)
        std::vector<ChargePairDev> cp;
        for (const auto& a : chargesOf(m1))
          for (const auto& b : chargesOf(m2)) {
            bool merged = false;
            for (auto& e : cp)
              if (e.x1 == a.x && e.y1 == a.y && e.z1 == a.z && e.x2 == b.x && e.y2 == b.y && e.z2 == b.z) {
                e.q += a.q * b.q;
                merged = true;
              }
            if (!merged) cp.push_back({a.q * b.q, a.x, a.y, a.z, b.x, b.y, b.z});
          }
        for (auto& e : cp)
          if (e.q != 0.0) t.charges.push_back(e);
        in.nCharge = (int)t.charges.size() - in.charge0;
        t.instances.push_back(in);
        int id = (int)t.instances.size() - 1 - c.inst0;
        instId[key] = id;
        return id;
      };
      for (int i = 0; i < dA; ++i)
        for (int j = 0; j < dB; ++j) {
          bool dBasis = !(i < 10 && j < 10);
          auto pa = polesOf(i, dBasis), pb = polesOf(j, dBasis);
          int l1a = orbL(kPairOrbitals[i][0]), l1b = orbL(kPairOrbitals[i][1]);
          int l2a = orbL(kPairOrbitals[j][0]), l2b = orbL(kPairOrbitals[j][1]);
          std::vector<EntryTermDev> terms;
          for (const auto& a : pa)
            for (const auto& b : pb) {
              int p1 = poleType(l1a, l1b, kPoleL[a.pole]), p2 = poleType(l2a, l2b, kPoleL[b.pole]);
              if (dBasis && a.pole == M20 && b.pole == M20) {
                // the reference reproduces the published PM6 error [M20,M20] = [Qzx,Qzx] - 1/4 [M22,M22]
                terms.push_back({a.f * b.f, instance(Qzx, Qzx, p1, p2), 0});
                terms.push_back({a.f * b.f * -0.25, instance(M22, M22, p1, p2), 0});
              } else if (kPoleM[a.pole] == kPoleM[b.pole]) {
                terms.push_back({a.f * b.f, instance(a.pole, b.pole, p1, p2), 0});
              }
            }
          if (terms.empty()) continue;
          EntryDev e{};
          e.i = i;
          e.j = j;
          e.term0 = (int)t.entryTerms.size();
          e.nTerm = (int)terms.size();
          for (auto& x : terms) t.entryTerms.push_back(x);
          t.entries.push_back(e);
        }
      c.nInst = (int)t.instances.size() - c.inst0;
      c.nEntry = (int)t.entries.size() - c.entry0;
      t.classes[lA * 3 + lB] = c;
    }
  // ---- one-centre integral table ----
  build1cTable(t);
  return t;
}

// One-centre (ij|kl) table: 58 unique integrals, each c1*SC[e1]+c2*SC[e2]+c3*SC[e3] over the
// Slater-Condon parameter set (order F0ss F0pp F0dd F0sp F0sd F0pd F2pp F2dd F2pd F4dd G1sp G1pd
// G2sd G3pd R1sppd R2sdpp R2sddd).
void build1cTable(ModelTables& t) {
  enum { F0ss, F0pp, F0dd, F0sp, F0sd, F0pd, F2pp, F2dd, F2pd, F4dd, G1sp, G1pd, G2sd, G3pd, R1sppd, R2sdpp, R2sddd };
  for (auto& a : t.oneCIndex) a = kNoPair;
  auto at = [&](int i, int j, int k, int l) -> uint8_t& { return t.oneCIndex[((i * 9 + j) * 9 + k) * 9 + l]; };
  const double r3 = std::sqrt(3.0), s12 = std::sqrt(12.0), s27 = std::sqrt(27.0), s245 = std::sqrt(245.0),
               s125 = std::sqrt(125.0), s45 = std::sqrt(45.0), s75 = std::sqrt(75.0), s300 = std::sqrt(300.0);
  struct Row {
    std::vector<std::array<int, 4>> where;
    double c[3];
    int e[3];
  };
  const std::vector<Row> rows = {
      {{{S, S, S, S}}, {1, 0, 0}, {F0ss, 0, 0}},
      {{{S, S, Z, Z}, {S, S, X, X}, {S, S, Y, Y}}, {1, 0, 0}, {F0sp, 0, 0}},
      {{{S, Z, S, Z}, {S, X, S, X}, {S, Y, S, Y}}, {1.0 / 3.0, 0, 0}, {G1sp, 0, 0}},
      {{{Z, Z, Z, Z}, {X, X, X, X}, {Y, Y, Y, Y}}, {1, 4.0 / 25.0, 0}, {F0pp, F2pp, 0}},
      {{{X, X, Z, Z}, {Y, Y, Z, Z}, {X, X, Y, Y}}, {1, -2.0 / 25.0, 0}, {F0pp, F2pp, 0}},
      {{{X, Z, X, Z}, {Y, Z, Y, Z}, {X, Y, X, Y}}, {3.0 / 25.0, 0, 0}, {F2pp, 0, 0}},
      {{{X, Z2, X, Z2}, {Y, Z2, Y, Z2}}, {1.0 / 15.0, 18.0 / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{X, Z2, X, X2Y2}, {X, Z2, Y, XY}, {X, XY, Y, Z2}}, {-r3 / 15.0, -s27 / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{X, X2Y2, Y, XY}}, {1.0 / 5.0, -21.0 / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{X, XY, Y, X2Y2}}, {-1.0 / 5.0, 21.0 / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{X, X, Z2, Z2}, {Y, Y, Z2, Z2}}, {1, -2.0 / 35.0, 0}, {F0pd, F2pd, 0}},
      {{{X, X, Z2, X2Y2}, {X, Y, Z2, XY}}, {-s12 / 35.0, 0, 0}, {F2pd, 0, 0}},
      {{{Y, Z2, Y, X2Y2}}, {r3 / 15.0, s27 / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{Z, Z2, Z, Z2}}, {4.0 / 15.0, 27.0 / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{Z, X2Y2, Z, X2Y2}, {Z, XY, Z, XY}, {X, YZ, X, YZ}, {Y, XZ, Y, XZ}, {X, XZ, Z, X2Y2}, {X, YZ, Z, XY}, {Y, XZ, Z, XY}, {X, YZ, Y, XZ}},
       {3.0 / 49.0, 0, 0}, {G3pd, 0, 0}},
      {{{Z, XZ, Z, XZ}, {Z, YZ, Z, YZ}, {X, XY, X, XY}, {X, X2Y2, X, X2Y2}, {X, XZ, X, XZ}, {Y, XY, Y, XY}, {Y, YZ, Y, YZ}, {Y, X2Y2, Y, X2Y2}},
       {1.0 / 5.0, 24.0 / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{X, XZ, Z, Z2}, {Y, YZ, Z, Z2}}, {s12 / 15.0, -std::sqrt(243.0) / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{X, X2Y2, Z, XZ}, {X, XY, Z, YZ}, {Y, XY, Z, XZ}, {X, XZ, Y, YZ}}, {1.0 / 5.0, -6.0 / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{Y, YZ, Z, X2Y2}}, {-3.0 / 49.0, 0, 0}, {G3pd, 0, 0}},
      {{{Y, Z2, Z, YZ}, {X, Z2, Z, XZ}}, {-r3 / 15.0, std::sqrt(432.0) / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{Y, X2Y2, Z, YZ}}, {-1.0 / 5.0, 6.0 / 245.0, 0}, {G1pd, G3pd, 0}},
      {{{Z, Z, Z2, Z2}}, {1, 4.0 / 35.0, 0}, {F0pd, F2pd, 0}},
      {{{X, Z, Z2, XZ}, {Y, Z, Z2, YZ}}, {r3 / 35.0, 0, 0}, {F2pd, 0, 0}},
      {{{Z, Z, X2Y2, X2Y2}, {Z, Z, XY, XY}, {X, X, YZ, YZ}, {Y, Y, XZ, XZ}}, {1, -4.0 / 35.0, 0}, {F0pd, F2pd, 0}},
      {{{X, Z, X2Y2, XZ}, {X, Z, XY, YZ}, {Y, Z, XY, XZ}, {X, Y, XZ, YZ}}, {3.0 / 35.0, 0, 0}, {F2pd, 0, 0}},
      {{{Y, Z, X2Y2, YZ}}, {-3.0 / 35.0, 0, 0}, {F2pd, 0, 0}},
      {{{Z, Z, XZ, XZ}, {Z, Z, YZ, YZ}, {X, X, XY, XY}, {X, X, X2Y2, X2Y2}, {X, X, XZ, XZ}, {Y, Y, XY, XY}, {Y, Y, YZ, YZ}, {Y, Y, X2Y2, X2Y2}},
       {1, 2.0 / 35.0, 0}, {F0pd, F2pd, 0}},
      {{{Y, Y, Z2, X2Y2}}, {s12 / 35.0, 0, 0}, {F2pd, 0, 0}},
      {{{S, Z2, S, Z2}, {S, X2Y2, S, X2Y2}, {S, XY, S, XY}, {S, XZ, S, XZ}, {S, YZ, S, YZ}}, {1.0 / 5.0, 0, 0}, {G2sd, 0, 0}},
      {{{S, S, Z2, Z2}, {S, S, X2Y2, X2Y2}, {S, S, XY, XY}, {S, S, XZ, XZ}, {S, S, YZ, YZ}}, {1, 0, 0}, {F0sd, 0, 0}},
      {{{Z2, Z2, Z2, Z2}, {X2Y2, X2Y2, X2Y2, X2Y2}, {XY, XY, XY, XY}, {XZ, XZ, XZ, XZ}, {YZ, YZ, YZ, YZ}}, {1, 4.0 / 49.0, 36.0 / 441.0}, {F0dd, F2dd, F4dd}},
      {{{Z2, X2Y2, Z2, X2Y2}, {Z2, XY, Z2, XY}}, {4.0 / 49.0, 15.0 / 441.0, 0}, {F2dd, F4dd, 0}},
      {{{Z2, XZ, Z2, XZ}, {Z2, YZ, Z2, YZ}}, {1.0 / 49.0, 30.0 / 441.0, 0}, {F2dd, F4dd, 0}},
      {{{Z2, Z2, X2Y2, X2Y2}, {Z2, Z2, XY, XY}}, {1, -4.0 / 49.0, 6.0 / 441.0}, {F0dd, F2dd, F4dd}},
      {{{Z2, XZ, X2Y2, XZ}, {Z2, XZ, XY, YZ}, {Z2, YZ, XY, XZ}}, {r3 / 49.0, -s75 / 441.0, 0}, {F2dd, F4dd, 0}},
      {{{Z2, YZ, X2Y2, YZ}}, {-r3 / 49.0, s75 / 441.0, 0}, {F2dd, F4dd, 0}},
      {{{Z2, Z2, XZ, XZ}, {Z2, Z2, YZ, YZ}}, {1, 2.0 / 49.0, -24.0 / 441.0}, {F0dd, F2dd, F4dd}},
      {{{Z2, X2Y2, XZ, XZ}, {Z2, XY, XZ, YZ}}, {-s12 / 49.0, s300 / 441.0, 0}, {F2dd, F4dd, 0}},
      {{{Z2, X2Y2, YZ, YZ}}, {s12 / 49.0, -s300 / 441.0, 0}, {F2dd, F4dd, 0}},
      {{{X2Y2, XY, X2Y2, XY}}, {35.0 / 441.0, 0, 0}, {F4dd, 0, 0}},
      {{{XY, XZ, XY, XZ}, {XY, YZ, XY, YZ}, {XZ, YZ, XZ, YZ}, {X2Y2, XZ, X2Y2, XZ}, {X2Y2, YZ, X2Y2, YZ}}, {3.0 / 49.0, 20.0 / 441.0, 0}, {F2dd, F4dd, 0}},
      {{{X2Y2, X2Y2, XY, XY}}, {1, 4.0 / 49.0, -34.0 / 441.0}, {F0dd, F2dd, F4dd}},
      {{{X2Y2, XZ, XY, YZ}}, {3.0 / 49.0, -15.0 / 441.0, 0}, {F2dd, F4dd, 0}},
      {{{X2Y2, YZ, XY, XZ}}, {-3.0 / 49.0, 15.0 / 441.0, 0}, {F2dd, F4dd, 0}},
      {{{XY, XY, XZ, XZ}, {XY, XY, YZ, YZ}, {XZ, XZ, YZ, YZ}, {X2Y2, X2Y2, XZ, XZ}, {X2Y2, X2Y2, YZ, YZ}}, {1, -2.0 / 49.0, -4.0 / 441.0}, {F0dd, F2dd, F4dd}},
      {{{S, Z2, Z2, Z2}}, {2.0 / s245, 0, 0}, {R2sddd, 0, 0}},
      {{{S, X2Y2, Z2, X2Y2}, {S, XY, Z2, XY}, {S, Z2, X2Y2, X2Y2}, {S, Z2, XY, XY}}, {-2.0 / s245, 0, 0}, {R2sddd, 0, 0}},
      {{{S, XZ, Z2, XZ}, {S, YZ, Z2, YZ}, {S, Z2, XZ, XZ}, {S, Z2, YZ, YZ}}, {1.0 / s245, 0, 0}, {R2sddd, 0, 0}},
      {{{S, XZ, X2Y2, XZ}, {S, XZ, XY, YZ}, {S, YZ, XY, XZ}, {S, XY, XZ, YZ}, {S, X2Y2, XZ, XZ}}, {r3 / s245, 0, 0}, {R2sddd, 0, 0}},
      {{{S, YZ, X2Y2, YZ}, {S, X2Y2, YZ, YZ}}, {-r3 / s245, 0, 0}, {R2sddd, 0, 0}},
      {{{S, Z2, X, X}, {S, Z2, Y, Y}}, {-1.0 / s125, 0, 0}, {R2sdpp, 0, 0}},
      {{{S, Y, Y, X2Y2}}, {-r3 / s45, 0, 0}, {R1sppd, 0, 0}},
      {{{S, X2Y2, Y, Y}}, {-r3 / s125, 0, 0}, {R2sdpp, 0, 0}},
      {{{S, Z, Z, Z2}}, {2.0 / s45, 0, 0}, {R1sppd, 0, 0}},
      {{{S, Z, X, XZ}, {S, X, Z, XZ}, {S, Z, Y, YZ}, {S, Y, Z, YZ}, {S, X, X, X2Y2}, {S, X, Y, XY}, {S, Y, X, XY}}, {r3 / s45, 0, 0}, {R1sppd, 0, 0}},
      {{{S, Z2, Z, Z}}, {2.0 / s125, 0, 0}, {R2sdpp, 0, 0}},
      {{{S, XZ, X, Z}, {S, YZ, Y, Z}, {S, X2Y2, X, X}, {S, XY, X, Y}}, {r3 / s125, 0, 0}, {R2sdpp, 0, 0}},
      {{{S, X, X, Z2}, {S, Y, Y, Z2}}, {-1.0 / s45, 0, 0}, {R1sppd, 0, 0}},
  };
  if (rows.size() != 58) throw std::logic_error("one-centre table must have 58 rows");
  for (size_t k = 0; k < rows.size(); ++k) {
    for (const auto& w : rows[k].where) at(w[0], w[1], w[2], w[3]) = (uint8_t)k;
    for (int c = 0; c < 3; ++c) {
      t.oneCCoef[k][c] = rows[k].c[c];
      t.oneCTerm[k][c] = rows[k].e[c];
    }
  }
  // (ij|kl) = (ji|kl) = (ij|lk) = (kl|ij) ...  (expansion in the reference's scan order)
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j)
      for (int k = 0; k < 9; ++k)
        for (int l = 0; l < 9; ++l) {
          uint8_t v = at(i, j, k, l);
          if (v == kNoPair) continue;
          at(j, i, k, l) = v;
          at(i, j, l, k) = v;
          at(j, i, l, k) = v;
          at(k, l, i, j) = v;
          at(k, l, j, i) = v;
          at(l, k, i, j) = v;
          at(l, k, j, i) = v;
        }
}

}  // namespace sparrow_amd
