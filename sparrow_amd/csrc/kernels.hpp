// Device-side table handles and kernel launch wrappers.
#pragma once
#include <hip/hip_runtime.h>

#include "model.hpp"

namespace sparrow_amd {

// Read-only model tables in device memory.
struct KTables {
  const uint8_t* pairIndex;     // [9][9]
  const uint8_t* pairOrb;       // [40][2]
  const int* pairBlock;         // [40]
  const int* blockStart;        // [6]
  const int* blockLen;          // [6]
  const int* blockMembers;      // [40]
  const LocalClassDev* classes; // [9]
  const MMInstanceDev* inst;
  const ChargePairDev* charges;
  const EntryDev* entries;
  const EntryTermDev* eterms;
  const uint8_t* oneCIndex;     // [9^4]
};

// Batch handles in device memory.
struct KBatch {
  const ElementDev* elems;
  const DiatomicDev* diatomic;  // [nElem * nElem]
  int nElem;
  int methodKind;               // MethodKind
  const double* oneCW;          // one-centre coefficient table of the parameter set (ElementDev::w1c)
  const AtomDev* atoms;
  const PairDev* pairs;
  const MolDev* mols;
  int natoms, npairs, nmols;
  const int* classPairs;        // pair ids grouped by class
  int classStart[10];           // ranges into classPairs
  // Work lists (subsets of the batch): classPairs/classStart list the pairs, atomList the atoms
  // and molList the molecules a launch covers (all of them for the set-up kernels of newly
  // admitted molecules; for the Fock build the molecules still iterating or still needing their
  // final energy)
  const int* atomList;          // atom ids: those without d shells first
  int nAtomList;
  int nAtomListSp;              // atomList[0, nAtomListSp) have no d shells, the rest have
  const int* molList;           // molecule ids
  int nMolList;
};

// ---- integrals.hip ----
// Two-centre two-electron blocks + VuvB vectors for all pairs of the batch.
void launchTwoCenter(const KTables& t, const KBatch& b, double* ints, double* vuvb, hipStream_t s);
// Overlap S (identity on the atom blocks), the off-diagonal S and H blocks of the lower triangle and
// the per-pair core repulsion.
void launchPairOverlap(const KBatch& b, double* S, double* H, double* erepPair, hipStream_t s);
// Diagonal H blocks (U + sum of VuvB).
void launchAtomHcore(const KTables& t, const KBatch& b, const double* vuvb, double* H, hipStream_t s);
// Upper triangles of S and H (H may be null) from their lower triangles (after launchPairOverlap
// and launchAtomHcore).
void launchMirrorLower(const KBatch& b, double* S, double* H, hipStream_t s);
// Work lists from per-slot lists kept in device memory: segment q (segs[3q] = source offset,
// segs[3q+1] = destination offset, segs[3q+2] = length) copies src[...] to dst[...].
void launchGatherSegments(const int* src, int* dst, const int* segs, int nsegs, hipStream_t s);
// Per-molecule sum of per-pair values (used for the repulsion energy).
void launchMolPairSum(const KBatch& b, const double* perPair, double* perMol, hipStream_t s);
// Standard density guess P0 = S*Nel/(2n) off the atom blocks, Z/nAO on the diagonal.
void launchDensityGuess(const KBatch& b, const double* S, const double* nElectrons, double* P, hipStream_t s);
// Nuclear gradient (Hartree/bohr, 3 per atom, atom order of the batch) at density P (n x n per
// molecule at matOff): the pair energies at fixed P with the second atom displaced by +-h
// (bohr) along each axis into epair (6 per pair), then their central differences summed per
// atom into grad.  Ps (may be null): the spin density P^a - P^b of a spin-polarised P = P^a + P^b
// (its exchange term is added).
void launchGradients(const KTables& t, const KBatch& b, const double* P, const double* Ps, double h, double* epair,
                     double* grad, hipStream_t s);

// ---- fock.hip ----
// F = H + G(P); per-pair / per-atom partial energies; J partial vectors.
void launchFockPairs(const KTables& t, const KBatch& b, const double* ints, const double* P, const double* H, double* F,
                     double* jvec, double* ePair, hipStream_t s);
void launchFockAtoms(const KTables& t, const KBatch& b, const double* P, const double* H, const double* jvec, double* F,
                     double* eAtom, hipStream_t s);
// Unrestricted F^alpha, F^beta (and E_el into e) of one molecule (n AOs, aoAtom: atom of every
// AO) from the restricted Fock matrices FT = F(P), Fa1 = F(Pa), Fb1 = F(Pb); WK: the one-centre
// exchange table.
void launchUnrestrictedFock(const KBatch& b, const double* WK, int n, const int* aoAtom, const double* H, const double* FT,
                            const double* Fa1, const double* Fb1, const double* P, const double* Pa, const double* Pb,
                            double* Fa, double* Fb, double* e, hipStream_t s);
// out = a + b, nn elements
void launchMatAdd(long long nn, const double* a, const double* b, double* out, hipStream_t s);
void launchEnergyReduce(const KBatch& b, const double* eAtom, const double* ePair, double* eMol, hipStream_t s);

// ---- cis.hip ----
// CIS AO Fock matrices F_v = cJ J(DJ_v) - cK K(DK_v^T) of nv pseudo-densities (n x n each,
// matStride apart, one molecule); jvec: Coulomb partial vectors (vecStride per vector).
void launchCisFock(const KTables& t, const KBatch& b, const double* ints, const double* WK, int nv, const double* DJ,
                   const double* DK, long long matStride, double cJ, double cK, double* jvec, long long vecStride,
                   double* F, hipStream_t s);
// S = dE .* X column by column (nconf x nv)
void launchCisDiag(long long nconf, int nv, const double* dE, const double* X, double* S, hipStream_t s);

}  // namespace sparrow_amd
