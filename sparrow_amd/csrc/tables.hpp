// Host-built static model tables (see tables.cpp).
#pragma once
#include <cstdint>
#include <vector>

#include "model.hpp"

namespace sparrow_amd {

extern const int kPairOrbitals[40][2];

struct ModelTables {
  uint8_t pairIndex[9][9];
  uint8_t pairOrb[40][2];
  // rotation block structure: blocks of global/local pairs with the same (l1,l2)
  int pairBlock[40];
  int blockStart[6], blockLen[6];
  int blockMembers[40];
  // local integral classes (index lA*3+lB)
  LocalClassDev classes[9];
  std::vector<MMInstanceDev> instances;
  std::vector<ChargePairDev> charges;
  std::vector<EntryDev> entries;
  std::vector<EntryTermDev> entryTerms;
  // one-centre integrals
  uint8_t oneCIndex[9 * 9 * 9 * 9];
  double oneCCoef[58][3];
  int oneCTerm[58][3];
};

ModelTables buildModelTables();
void build1cTable(ModelTables& t);

}  // namespace sparrow_amd
