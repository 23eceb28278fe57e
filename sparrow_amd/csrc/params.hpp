// NDDO parameter sets: loading of the flat parameter tables (sparrow_amd/params/<method>.txt)
// and conversion of raw parameters into runtime element data (ElementDev).
#pragma once
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "model.hpp"
#include "tables.hpp"

namespace sparrow_amd {

// CODATA 2014 (scine_utilities Utils::Constants); eV/Hartree pinned by the reference's molden
// orbital energies (Tests/AppTests/molden_input_reference.molden) to 1e-9 Eh.
constexpr double kAngstromPerBohr = 0.52917721067;
constexpr double kBohrPerAngstrom = 1.0 / 0.52917721067;
constexpr double kEvPerHartree = 27.21138602;

enum class MethodKind { MNDO = 0, AM1 = 1, PM3 = 2, RM1 = 3, PM6 = 4 };

struct RawElement {
  int Z = 0;
  double U[3]{}, beta[3]{}, zeta[3]{}, zetaInt[3]{};
  double gss = 0, gpp = 0, gsp = 0, gp2 = 0, hsp = 0, pcore = 0, f0sd = 0, g2sd = 0, alpha = 0;
  std::vector<std::array<double, 3>> gauss;
  int nShell = 0;
  int shellL[3]{};
  double shE[3][6]{}, shC[3][6]{};
};

class ParameterSet {
 public:
  static ParameterSet load(const std::string& path);
  MethodKind kind() const { return kind_; }
  const std::string& name() const { return name_; }
  bool spd() const { return spd_; }
  bool diatomic() const { return diatomic_; }
  bool hasElement(int Z) const { return raw_.count(Z) != 0; }
  std::vector<int> elements() const {
    std::vector<int> z;
    for (const auto& kv : raw_) z.push_back(kv.first);
    return z;
  }
  bool hasPair(int Z1, int Z2) const;
  // Runtime data of element Z (computed once, cached).
  const ElementDev& element(int Z, const ModelTables& tables);
  DiatomicDev diatomicParams(int Z1, int Z2) const;
  // One-centre Fock coefficients W(ij,kl) = (ij|kl) - 1/2 (ik|jl) of every element, a dense
  // nao^2 x nao^2 block at offset kOneCenterBlock * w1c, stored by column: W(ij,kl) at
  // kl nao^2 + ij (ij = i nao + j, kl = k nao + l), so that consecutive ij are adjacent.
  std::vector<double> oneCenterTable(const ModelTables& tables);
  // One-centre exchange integrals K(ij,kl) = (ik|jl) in the same layout (for the spin-polarised
  // G: the same-spin exchange on an atom's diagonal block).
  std::vector<double> oneCenterExchangeTable(const ModelTables& tables);

 private:
  MethodKind kind_ = MethodKind::PM6;
  std::string name_;
  bool spd_ = false, diatomic_ = false;
  std::map<int, RawElement> raw_;
  std::map<std::pair<int, int>, std::pair<double, double>> pairs_;
  std::map<int, ElementDev> runtime_;
  double lastSlaterCondon_[17]{};
};

}  // namespace sparrow_amd
