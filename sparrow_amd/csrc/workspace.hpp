// Device-resident lock-step SCF over a changing set of molecules.
//
// A Workspace owns fixed-capacity pools in one block of device memory (atoms, atom pairs,
// two-centre integrals, per-pair vectors, n x n matrices, DIIS history, length-n vectors and the
// split-BF16 copies of the mixed-precision SP2 iterates).  A molecule admitted into it gets a
// slot (its MolDev index) and one range in each pool; all kernels address a molecule through
// those offsets, so molecules of any size can come and go between iterations.  The set-up of
// newly admitted molecules (integrals, overlap, H, repulsion, guess) runs at the start of the
// next iteration; a molecule leaves the iteration lists once it has converged (or reached the
// iteration cap) and its final energy is known.
//
// Engine::solve runs one Workspace over a fixed batch; the serving loop keeps one alive and
// admits requests as capacity frees up.
#pragma once
#include <array>
#include <map>
#include <vector>

#include "engine.hpp"
#include "util.hpp"

namespace sparrow_amd {

// First-fit allocator of ranges in [0, capacity).
class RangeAlloc {
 public:
  void reset(long long capacity);
  bool alloc(long long len, long long* off);
  void release(long long off, long long len);
  long long capacity() const { return cap_; }
  long long used() const { return used_; }

 private:
  std::map<long long, long long> free_;  // start -> length
  long long cap_ = 0, used_ = 0;
};

// Pool capacities (elements) and the number of molecule slots.
struct WorkspaceCaps {
  int mols = 0;
  long long atoms = 0, pairs = 0, ints = 0, pvec = 0, mat = 0, vecN = 0, hist = 0, lowp = 0;
};

// What one molecule takes from each pool.
struct MolDemand {
  long long atoms = 0, pairs = 0, ints = 0, pvec = 0, mat = 0, vecN = 0, hist = 0, lowp = 0;
  int n = 0, nocc = 0;
};

class Workspace {
 public:
  // Device bytes a workspace with these capacities needs.
  static size_t bytesFor(const WorkspaceCaps& caps, int diisSlots);
  // Capacities that hold exactly these molecules.
  static WorkspaceCaps capsFor(const std::vector<MolDemand>& mols);

  Workspace(Engine& engine, ParameterSet& ps, const ScfSettings& st, const WorkspaceCaps& caps, char* deviceBase,
            size_t deviceBytes);

  // Demand of a molecule; throws std::runtime_error for invalid input (empty, unknown element or
  // element pair, odd electron count).
  static MolDemand demandOf(ParameterSet& ps, const ModelTables& tables, const MoleculeInput& in, int diisSlots, bool mixed);
  MolDemand demand(const MoleculeInput& in) { return demandOf(ps_, eng_.hostTables(), in, slotsDiis_, mixed_); }
  // Admits a molecule if every pool has room (set-up runs in the next iterate()); returns its
  // slot or -1.  The density guess pointer, if any, is copied.
  int admit(const MoleculeInput& in);
  // Admits with the guess taken from device memory (a cached density of the same size).
  int admitWarm(const MoleculeInput& in, const double* devGuess);
  // Set-up of newly admitted molecules, then one SCF iteration of every active molecule.
  void iterate();
  bool busy() const { return nActive_ > 0 || !pending_.empty(); }
  bool finished(int slot) const { return slots_[slot].used && slots_[slot].finished; }
  std::vector<int> finishedSlots() const;
  // Results of finished slots (orbital energies / orbitals from one batched eigensolve of
  // their last extrapolated Fock matrices when asked; densities when asked), then the slots are
  // released.
  std::vector<MoleculeResult> collect(const std::vector<int>& slots, bool wantEps, bool wantOrbitals,
                                      bool wantDensity = true);
  // Device address of a finished slot's density (valid until collect releases the slot).
  const double* densityOf(int slot) const;
  int nao(int slot) const { return slots_[slot].md.n; }
  const SolveCounters& counters() const { return counters_; }
  int active() const { return nActive_; }
  PhaseTimes times;

 private:
  struct Slot {
    bool used = false, setupDone = false, finished = false;
    bool inFock = false, inIter = false;
    MolDev md{};
    MolDemand dm;
    long long intOff = 0, pvecOff = 0, lpOff = 0;
    std::array<int, 9> classCnt{};                 // pairs per (l_A, l_B) class
    int nAtomsSp = 0;                              // atoms without d shells
    double nel = 0;
    std::vector<double> guess;                     // host warm-start density (n x n) if given
    const double* devGuess = nullptr;              // device warm-start density if given
    ScfMolDev scf{};
    int iterations = 0;
  };

  int admitImpl(const MoleculeInput& in, const double* devGuess);
  void runSetup();
  void buildFockLists();
  void buildIterLists();
  // Work lists of the given slots (pairs by class, atoms without d shells first) gathered on the
  // device from the per-slot lists into dClassPairs_ / dAtomList_; fills b's list fields
  void gatherLists(const std::vector<int>& slots, KBatch& b);
  void eigenDensity(const std::vector<int>& ids);

  Engine& eng_;
  ParameterSet& ps_;
  ScfSettings st_;
  WorkspaceCaps caps_;
  int slotsDiis_ = 8;
  bool sp2_ = true, mixed_ = false;
  std::map<int, int> elemIndex_;
  std::vector<Slot> slots_;
  std::vector<int> pending_;
  int nActive_ = 0;
  RangeAlloc aAtoms_, aPairs_, aInts_, aPvec_, aMat_, aVecN_, aHist_, aLowp_;
  SolveCounters counters_;
  int lastSp2Steps_ = 0;
  bool fockDirty_ = true, iterDirty_ = true;
  KBatch b_{};
  int tileX_ = 32;

  // device memory (carved from the block given to the constructor)
  ElementDev* dElems_ = nullptr;
  DiatomicDev* dDiat_ = nullptr;
  AtomDev* dAtoms_ = nullptr;
  PairDev* dPairs_ = nullptr;
  MolDev* dMols_ = nullptr;
  MolDev* dMolsAct_ = nullptr;
  int *dClassPairs_ = nullptr, *dAtomList_ = nullptr, *dMolList_ = nullptr;
  // per-slot lists written at admission: a slot's pair ids sorted by class at [pair0, pair0 +
  // npairs), its atoms (without d shells first) at [atom0, atom0 + natoms); gather segments
  int *dSlotPairs_ = nullptr, *dSlotAtoms_ = nullptr, *dSegs_ = nullptr;
  PinnedBuf<int> segHost_;
  double *dInts_ = nullptr, *dVuvb_ = nullptr, *dJvec_ = nullptr;
  double *dS_ = nullptr, *dH_ = nullptr, *dF_ = nullptr, *dP_ = nullptr, *dPold_ = nullptr, *dTmp_ = nullptr;
  double *dFx_ = nullptr, *dA_ = nullptr, *dZ_ = nullptr, *dWork_ = nullptr;
  unsigned char* dPiv_ = nullptr;
  double *dHistF_ = nullptr, *dHistE_ = nullptr, *dHistP_ = nullptr;
  double *dEd_ = nullptr, *dEe_ = nullptr, *dEtau_ = nullptr, *dEv_ = nullptr, *dEw_ = nullptr, *dEval_ = nullptr;
  double *dLo_ = nullptr, *dHi_ = nullptr;
  double *dLz_ = nullptr, *dLzCoef_ = nullptr;  // Lanczos vectors and coefficients of the SP2 bounds
  // SP2 spectral bounds: 0 Gershgorin discs only, 1 narrowed by Lanczos for every molecule of 32 AOs
  // or more, 2 for molecules of kLanczosMinN AOs or more (environment SPARROW_AMD_SP2_BOUNDS =
  // gershgorin | lanczos | auto)
  int lanczosMode_ = 2;
  static constexpr int kLanczosMinN = 400;
  // tridiagonalisation on the many-workgroup path: density fallbacks from this order (by size only),
  // final orbital energies of a few molecules from kWideSytrdMinN
  static constexpr int kWideEigenMinN = 768;
  static constexpr int kWideSytrdMinN = 192;
  double *dErepPair_ = nullptr, *dEPair_ = nullptr, *dEAtom_ = nullptr;
  double *dEMol_ = nullptr, *dErep_ = nullptr, *dRmsd_ = nullptr, *dNel_ = nullptr;
  double *dDots_ = nullptr, *dCoef_ = nullptr, *dBmat_ = nullptr, *dEHist_ = nullptr, *dPf_ = nullptr, *dMmat_ = nullptr;
  ScfMolDev* dScf_ = nullptr;
  EigJob *dEigOcc_ = nullptr, *dEigSub_ = nullptr;
  GemmJob* dGemmSub_ = nullptr;
  Sp2StepJob *dStepX0_ = nullptr, *dStepX1_ = nullptr;
  int *dTilesSub_ = nullptr, *dTiles32_ = nullptr, *dTilesS_ = nullptr, *dTiles64_ = nullptr;
  double* dPart_[2] = {nullptr, nullptr};
  Sp2Dev* dSp2_[2] = {nullptr, nullptr};
  double* dRed_ = nullptr;
  unsigned short* dLp_[4] = {nullptr, nullptr, nullptr, nullptr};

  // host copies of the per-slot job templates
  std::vector<EigJob> eigOcc_, eigAll_;
  std::vector<GemmJob> gP_;
  std::vector<Sp2StepJob> gX0_, gX1_;

  // iteration lists (molecules not yet converged)
  struct IterLists {
    std::vector<MolDev> mols;
    std::vector<int> t32, tS, t64;
    int nt32 = 0, ntS = 0, maxN = 0, maxNocc = 0;
    int k32 = 0, nt32Small = 0;  // molecules with 32-wide SP2 tiles (listed first) and their 32 x 32 tiles
    int nt64 = 0;                // 64 x 64 tiles (all, not only the upper ones) of the others
    int ntS32 = 0, ntS64 = 0, ntS128 = 0;  // SP2 tiles of the molecules with 32-, 64- and 128-wide tiles
  } IL_;
  std::vector<int> fockMols_;
  PinnedBuf<Sp2Dev> sp2State_;
  PinnedBuf<ScfMolDev> scfHost_;

  // Device work of an iteration between two host checks, replayed from a hipGraph while the
  // work lists and launch counts stay the same: segment A = Fock build, mixer, SP2 start and the
  // first chunk of SP2 steps (ending in the copy of the SP2 states); segment B = P from X, RMSD,
  // SCF check (ending in the copy of the SCF states).  A launch sequence is captured the second
  // time its key is seen, so sets that change every iteration run eagerly; up to 16 captured
  // sequences are kept per segment (the oldest is dropped).
  struct GraphSlot {
    std::map<long long, hipGraphExec_t> exec;
    std::vector<long long> order;  // capture order
    std::map<long long, int> seen;
  };
  template <class Launches>
  void runSegment(GraphSlot& g, long long key, Launches&& launches);
  void launchFockAndMixer(int k);
  GraphSlot segA_, segB_;
  long long listVersion_ = 0;
  bool useGraphs_ = false;

 public:
  ~Workspace();
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  // hipGraph replay of the iteration segments (ScfSettings::graphs; off while kernels are profiled)
  void setUseGraphs(bool on) { useGraphs_ = on; }
};

}  // namespace sparrow_amd
