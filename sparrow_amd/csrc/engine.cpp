// SCF driver for batches of molecules on one GPU.
//
// The SCF loop restates the absent scine_utilities Utils::ScfMethod as it is driven by the
// reference (N/Pm6/Wrapper/PM6MethodWrapper.cpp:87-89, N/NDDOMethodWrapper.cpp:68-103):
//   per geometry: S, H, two-centre integrals, E_rep, standard guess (or a given density);
//   per iteration: F = H + G(P), E_el(P); Pulay DIIS on FP - PF (orthogonal NDDO basis);
//                  diagonalise; aufbau; P = 2 C_occ C_occ^T;
//   converged when |E_k - E_{k-1}| < self_consistence_criterion and RMSD(P_k - P_{k-1}) <
//   density_rmsd_criterion; then F and E_el are rebuilt from the converged P
//   (NU/FockMatrix.cpp:147-150) and E = E_el + E_rep.
// All molecules of a batch advance in lock step; each molecule's reported state is the one of
// its own convergence iteration, so its result does not depend on the other molecules.
#include "engine.hpp"

#include "util.hpp"
#include "workspace.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <stdexcept>

namespace sparrow_amd {



DeviceArena::~DeviceArena() {
  if (base_) (void)hipFree(base_);
}

void DeviceArena::reserve(size_t bytes) {
  if (bytes <= cap_) return;
  if (base_) HIPCHK(hipFree(base_));
  base_ = nullptr;
  cap_ = 0;
  size_t want = std::max(bytes, (size_t)64 << 20);
  HIPCHK(hipMalloc(&base_, want));
  cap_ = want;
}

KernelProfiler::~KernelProfiler() {
  for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  for (auto& p : pending_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
}

hipEvent_t KernelProfiler::get() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  HIPCHK(hipEventCreate(&e));
  return e;
}

void KernelProfiler::begin(const char* name, hipStream_t s) {
  current_ = name;
  curStart_ = get();
  HIPCHK(hipEventRecord(curStart_, s));
}

void KernelProfiler::end(hipStream_t s) {
  hipEvent_t b = get();
  HIPCHK(hipEventRecord(b, s));
  pending_.push_back({current_, curStart_, b});
}

void KernelProfiler::collect() {
  for (auto& p : pending_) {
    HIPCHK(hipEventSynchronize(p.b));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
    auto& a = acc_[p.name];
    a.first += ms;
    a.second += 1;
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  pending_.clear();
}

void KernelProfiler::reset() {
  collect();
  acc_.clear();
}

bool KernelProfiler::query(const std::string& name, double* totalMs, long long* count) const {
  auto it = acc_.find(name);
  if (it == acc_.end()) {
    *totalMs = 0;
    *count = 0;
    return false;
  }
  *totalMs = it->second.first;
  *count = it->second.second;
  return true;
}

std::vector<std::string> KernelProfiler::names() const {
  std::vector<std::string> v;
  for (auto& kv : acc_) v.push_back(kv.first);
  return v;
}

namespace {

template <class T>
T* uploadTable(std::vector<void*>& allocs, const T* host, size_t count) {
  void* p = nullptr;
  size_t bytes = std::max<size_t>(count * sizeof(T), 16);
  HIPCHK(hipMalloc(&p, bytes));
  if (count) HIPCHK(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
  allocs.push_back(p);
  return static_cast<T*>(p);
}


// Host-side layout of a batch.
struct Layout {
  std::vector<ElementDev> elems;
  std::map<int, int> elemIndex;
  std::vector<DiatomicDev> diat;
  std::vector<AtomDev> atoms;
  std::vector<PairDev> pairs;
  std::vector<MolDev> mols;
  std::vector<int> classPairs;
  int classStart[10] = {};
  long long nInts = 0, nVec = 0, nMat = 0, nVecN = 0, nHist = 0;
  int maxN = 0, maxNocc = 0;
  std::vector<double> nel;
};

int dimOf(int l) { return l == 0 ? 1 : l == 1 ? 10 : 40; }

// closedShell: reject odd electron counts (every restricted path); the spin-polarised Fock build
// takes any electron count (its densities come from the caller)
Layout buildLayout(ParameterSet& ps, const ModelTables& t, const std::vector<MoleculeInput>& mols, int slots,
                   bool closedShell = true) {
  Layout L;
  for (const auto& m : mols) {
    if (m.Z.empty()) throw std::runtime_error("empty molecular structure");
    if (m.xyz.size() != 3 * m.Z.size()) throw std::runtime_error("position array does not match the number of atoms");
    for (int Z : m.Z)
      if (!L.elemIndex.count(Z)) {
        const ElementDev& e = ps.element(Z, t);
        L.elemIndex[Z] = (int)L.elems.size();
        L.elems.push_back(e);
      }
  }
  const int ne = (int)L.elems.size();
  L.diat.assign((size_t)ne * ne, DiatomicDev{});
  for (int a = 0; a < ne; ++a)
    for (int b = 0; b < ne; ++b) {
      int za = L.elems[a].Z, zb = L.elems[b].Z;
      if (ps.diatomic()) L.diat[(size_t)a * ne + b] = ps.diatomicParams(za, zb);
    }
  std::vector<std::vector<int>> byClass(9);
  for (size_t mi = 0; mi < mols.size(); ++mi) {
    const auto& in = mols[mi];
    MolDev md{};
    md.id = (int)mi;
    md.atom0 = (int)L.atoms.size();
    md.natoms = (int)in.Z.size();
    md.pair0 = (int)L.pairs.size();
    int ao = 0;
    double coreSum = 0;
    for (size_t k = 0; k < in.Z.size(); ++k) {
      AtomDev a{};
      a.elem = L.elemIndex[in.Z[k]];
      a.mol = (int)mi;
      a.ao = ao;
      a.nao = L.elems[a.elem].nao;
      a.x = in.xyz[3 * k];
      a.y = in.xyz[3 * k + 1];
      a.z = in.xyz[3 * k + 2];
      ao += a.nao;
      coreSum += std::lround(L.elems[a.elem].core);
      L.atoms.push_back(a);
    }
    if (ps.diatomic())
      for (size_t i = 0; i < in.Z.size(); ++i)
        for (size_t j = 0; j <= i; ++j)
          if (!ps.hasPair(in.Z[i], in.Z[j]))
            throw std::runtime_error("NDDO parameters do not exist for element pair Z=" + std::to_string(in.Z[i]) + "-" +
                                     std::to_string(in.Z[j]));
    md.n = ao;
    double nel = coreSum - in.charge;
    long long ne2 = std::llround(nel);
    if (ne2 < 0 || (closedShell && ne2 % 2 != 0))
      throw std::runtime_error("restricted closed-shell SCF needs an even, non-negative electron count");
    md.nocc = (int)(ne2 / 2);
    if (md.nocc > md.n) throw std::runtime_error("more electrons than spin orbitals");
    L.nel.push_back(nel);
    const long long vecBase = L.nVec;
    for (int i = 0; i < md.natoms; ++i)
      for (int j = i + 1; j < md.natoms; ++j) {
        PairDev p{};
        p.a = md.atom0 + i;
        p.b = md.atom0 + j;
        p.mol = (int)mi;
        int lA = L.elems[L.atoms[p.a].elem].l, lB = L.elems[L.atoms[p.b].elem].l;
        p.cls = lA * 3 + lB;
        p.intOff = L.nInts;
        p.vecOff = L.nVec;
        p.aoA = L.atoms[p.a].ao;
        p.aoB = L.atoms[p.b].ao;
        p.n = ao;
        p.matOff = L.nMat;  // this molecule's matrices start at the current end (set below)
        L.nInts += (long long)dimOf(lA) * dimOf(lB);
        L.nVec += dimOf(lA) + dimOf(lB);
        byClass[p.cls].push_back((int)L.pairs.size());
        L.pairs.push_back(p);
      }
    md.npairs = (int)L.pairs.size() - md.pair0;
    {
      std::vector<int> d(md.natoms);
      for (int k = 0; k < md.natoms; ++k) d[k] = dimOf(L.elems[L.atoms[md.atom0 + k].elem].l);
      setCoulombPartialOffsets(L.atoms.data() + md.atom0, L.pairs.data() + md.pair0, md.natoms, d.data(), vecBase);
    }
    md.matOff = L.nMat;
    md.vecOff = L.nVecN;
    md.histOff = L.nHist;
    L.nMat += (long long)md.n * md.n;
    L.nVecN += md.n;
    L.nHist += (long long)slots * md.n * md.n;
    L.maxN = std::max(L.maxN, md.n);
    L.maxNocc = std::max(L.maxNocc, md.nocc);
    L.mols.push_back(md);
  }
  for (int c = 0; c < 9; ++c) {
    L.classStart[c] = (int)L.classPairs.size();
    for (int p : byClass[c]) L.classPairs.push_back(p);
  }
  L.classStart[9] = (int)L.classPairs.size();
  return L;
}


template <class T>
size_t sizeOf(size_t count) {
  return ((std::max<size_t>(count * sizeof(T), 8) + 255) & ~size_t(255)) + 256;
}

}  // namespace

Engine::Engine(int device, const std::string& paramDir) : device_(device), paramDir_(paramDir) {
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreate(&stream_));
  tables_ = buildModelTables();
  ktab_.pairIndex = uploadTable(tableAllocs_, &tables_.pairIndex[0][0], 81);
  ktab_.pairOrb = uploadTable(tableAllocs_, &tables_.pairOrb[0][0], 80);
  ktab_.pairBlock = uploadTable(tableAllocs_, tables_.pairBlock, 40);
  ktab_.blockStart = uploadTable(tableAllocs_, tables_.blockStart, 6);
  ktab_.blockLen = uploadTable(tableAllocs_, tables_.blockLen, 6);
  ktab_.blockMembers = uploadTable(tableAllocs_, tables_.blockMembers, 40);
  ktab_.classes = uploadTable(tableAllocs_, tables_.classes, 9);
  ktab_.inst = uploadTable(tableAllocs_, tables_.instances.data(), tables_.instances.size());
  ktab_.charges = uploadTable(tableAllocs_, tables_.charges.data(), tables_.charges.size());
  ktab_.entries = uploadTable(tableAllocs_, tables_.entries.data(), tables_.entries.size());
  ktab_.eterms = uploadTable(tableAllocs_, tables_.entryTerms.data(), tables_.entryTerms.size());
  ktab_.oneCIndex = uploadTable(tableAllocs_, tables_.oneCIndex, 9 * 9 * 9 * 9);
  scratch_ = scratchAddress(device);  // device address of the allocation-free scratch array
}

Engine::~Engine() {
  for (void* p : tableAllocs_) (void)hipFree(p);
  if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::sp2Tile(int n, bool batched) const {
  const auto& t = batched ? sp2TilesBatched_ : sp2TilesSingle_;
  for (const auto& e : t)
    if (n <= e.first) return e.second;
  return t.back().second;
}

void Engine::setSp2Tiles(bool batched, const std::vector<std::pair<int, int>>& table) {
  if (table.empty()) throw std::runtime_error("empty SP2 tile table");
  for (size_t i = 0; i < table.size(); ++i) {
    const int t = table[i].second;
    if (t != 32 && t != 64 && t != 128) throw std::runtime_error("SP2 tile edges are 32, 64 or 128");
    if (i > 0 && table[i].first <= table[i - 1].first) throw std::runtime_error("SP2 tile table: nmax must ascend");
  }
  (batched ? sp2TilesBatched_ : sp2TilesSingle_) = table;
}

const double* Engine::oneCenterTable(ParameterSet& ps) {
  auto it = oneCW_.find(&ps);
  if (it != oneCW_.end()) return it->second;
  const std::vector<double> w = ps.oneCenterTable(tables_);
  return oneCW_[&ps] = uploadTable(tableAllocs_, w.data(), w.size());
}

const double* Engine::oneCenterExchangeTable(ParameterSet& ps) {
  auto it = oneCK_.find(&ps);
  if (it != oneCK_.end()) return it->second;
  const std::vector<double> w = ps.oneCenterExchangeTable(tables_);
  return oneCK_[&ps] = uploadTable(tableAllocs_, w.data(), w.size());
}

ParameterSet& Engine::method(const std::string& name) {
  std::string key;
  for (char c : name) key += (char)std::toupper((unsigned char)c);
  auto it = methods_.find(key);
  if (it != methods_.end()) return it->second;
  static const char* known[] = {"MNDO", "AM1", "PM3", "RM1", "PM6"};
  bool ok = false;
  for (const char* k : known) ok |= key == k;
  if (!ok) throw std::runtime_error("This calculator does not provide the requested method: " + name);
  std::string lower;
  for (char c : key) lower += (char)std::tolower((unsigned char)c);
  return methods_[key] = ParameterSet::load(paramDir_ + "/" + lower + ".txt");
}

std::vector<MoleculeResult> Engine::solve(const std::string& methodName, const std::vector<MoleculeInput>& inputs,
                                          const ScfSettings& st, bool wantOrbitals, bool wantEps, PhaseTimes* times,
                                          bool wantDensity) {
  HIPCHK(hipSetDevice(device_));
  const double t0 = nowSeconds();
  ParameterSet& ps = method(methodName);
  const int slots = std::max(2, st.diisSize);
  const bool mixed = st.precision == 1;
  std::vector<MolDemand> dm;
  for (const auto& in : inputs) dm.push_back(Workspace::demandOf(ps, tables_, in, slots, mixed));
  const WorkspaceCaps caps = Workspace::capsFor(dm);
  arena_.reserve(Workspace::bytesFor(caps, slots));
  Workspace ws(*this, ps, st, caps, arena_.base(), arena_.capacity());
  std::vector<int> ids;
  for (const auto& in : inputs) {
    const int slot = ws.admit(in);
    if (slot < 0) throw std::runtime_error("internal: batch does not fit its own workspace");
    ids.push_back(slot);
  }
  while (ws.busy()) ws.iterate();
  std::vector<MoleculeResult> res = ws.collect(ids, wantEps, wantOrbitals, wantDensity);
  counters_ = ws.counters();
  HIPCHK(hipStreamSynchronize(stream_));
  if (prof_.enabled()) prof_.collect();
  if (times) {
    *times = ws.times;
    times->total = nowSeconds() - t0;
  }
  return res;
}

void Engine::densityIndependent(const std::string& methodName, const MoleculeInput& mol, double* S, double* H, double* Erep) {
  HIPCHK(hipSetDevice(device_));
  ParameterSet& ps = method(methodName);
  Layout L = buildLayout(ps, tables_, {mol}, 2, /*closedShell*/ false);  // S, H, E_rep: any electron count
  const long long nMat = std::max<long long>(L.nMat, 1);
  size_t need = sizeOf<ElementDev>(L.elems.size()) + sizeOf<DiatomicDev>(L.diat.size()) + sizeOf<AtomDev>(L.atoms.size()) +
                sizeOf<PairDev>(L.pairs.size()) + sizeOf<MolDev>(1) + sizeOf<int>(L.classPairs.size()) +
                sizeOf<double>(L.nInts) + sizeOf<double>(L.nVec) + 2 * sizeOf<double>(nMat) + sizeOf<double>(L.pairs.size()) +
                sizeOf<double>(1) + sizeOf<int>(L.atoms.size()) + sizeOf<int>(1);
  arena_.reserve(need);
  Carver cv{arena_.base(), 0, arena_.capacity()};
  ElementDev* dElems = cv.take<ElementDev>(L.elems.size());
  DiatomicDev* dDiat = cv.take<DiatomicDev>(L.diat.size());
  AtomDev* dAtoms = cv.take<AtomDev>(L.atoms.size());
  PairDev* dPairs = cv.take<PairDev>(L.pairs.size());
  MolDev* dMols = cv.take<MolDev>(1);
  int* dClassPairs = cv.take<int>(L.classPairs.size());
  int* dAtomList = cv.take<int>(L.atoms.size());
  int* dMolList = cv.take<int>(1);
  double* dInts = cv.take<double>(L.nInts);
  double* dVuvb = cv.take<double>(L.nVec);
  double* dS = cv.take<double>(nMat);
  double* dH = cv.take<double>(nMat);
  double* dErepPair = cv.take<double>(L.pairs.size());
  double* dErep = cv.take<double>(1);
  hipStream_t s = stream_;
  std::vector<int> atomList;  // atoms without d shells first
  for (int pass = 0; pass < 2; ++pass)
    for (size_t a = 0; a < L.atoms.size(); ++a)
      if ((L.elems[L.atoms[a].elem].l == 2) == (pass == 1)) atomList.push_back((int)a);
  int nAtomListSp = 0;
  for (const AtomDev& a : L.atoms) nAtomListSp += L.elems[a.elem].l == 2 ? 0 : 1;
  const int molList = 0;
  HIPCHK(hipMemcpyAsync(dAtomList, atomList.data(), atomList.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dMolList, &molList, sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dElems, L.elems.data(), L.elems.size() * sizeof(ElementDev), hipMemcpyHostToDevice, s));
  if (!L.diat.empty()) HIPCHK(hipMemcpyAsync(dDiat, L.diat.data(), L.diat.size() * sizeof(DiatomicDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dAtoms, L.atoms.data(), L.atoms.size() * sizeof(AtomDev), hipMemcpyHostToDevice, s));
  if (!L.pairs.empty()) HIPCHK(hipMemcpyAsync(dPairs, L.pairs.data(), L.pairs.size() * sizeof(PairDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dMols, L.mols.data(), sizeof(MolDev), hipMemcpyHostToDevice, s));
  if (!L.classPairs.empty())
    HIPCHK(hipMemcpyAsync(dClassPairs, L.classPairs.data(), L.classPairs.size() * sizeof(int), hipMemcpyHostToDevice, s));
  KBatch b{};
  b.elems = dElems;
  b.diatomic = dDiat;
  b.nElem = (int)L.elems.size();
  b.methodKind = (int)ps.kind();
  b.oneCW = oneCenterTable(ps);
  b.atoms = dAtoms;
  b.pairs = dPairs;
  b.mols = dMols;
  b.natoms = (int)L.atoms.size();
  b.npairs = (int)L.pairs.size();
  b.nmols = 1;
  b.classPairs = dClassPairs;
  for (int c = 0; c < 10; ++c) b.classStart[c] = L.classStart[c];
  b.atomList = dAtomList;
  b.nAtomList = (int)atomList.size();
  b.nAtomListSp = nAtomListSp;
  b.molList = dMolList;
  b.nMolList = 1;
  launchTwoCenter(ktab_, b, dInts, dVuvb, s);
  launchPairOverlap(b, dS, dH, dErepPair, s);
  launchAtomHcore(ktab_, b, dVuvb, dH, s);
  launchMirrorLower(b, dS, dH, s);
  double erep = 0;
  if (b.npairs > 0) {
    launchMolPairSum(b, dErepPair, dErep, s);
    HIPCHK(hipMemcpyAsync(&erep, dErep, sizeof(double), hipMemcpyDeviceToHost, s));
  }
  const size_t nn = (size_t)L.mols[0].n * L.mols[0].n;
  if (S) HIPCHK(hipMemcpyAsync(S, dS, nn * sizeof(double), hipMemcpyDeviceToHost, s));
  if (H) HIPCHK(hipMemcpyAsync(H, dH, nn * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (Erep) *Erep = erep;
}

const std::vector<double>& Engine::oneCenterHost(ParameterSet& ps) {
  auto it = oneCWHost_.find(&ps);
  if (it == oneCWHost_.end()) it = oneCWHost_.emplace(&ps, ps.oneCenterTable(tables_)).first;
  return it->second;
}

void Engine::fock(const std::string& methodName, const MoleculeInput& mol, const double* P, double* F, double* Eel) {
  fockImpl(methodName, mol, P, F, Eel, nullptr, 0);
}

void Engine::fockNoAlloc(const std::string& methodName, const MoleculeInput& mol, const double* P, double* F,
                         double* Eel) {
  fockImpl(methodName, mol, P, F, Eel, scratch_, kNoAllocScratchBytes);
}

void Engine::fockImpl(const std::string& methodName, const MoleculeInput& mol, const double* P, double* F, double* Eel,
                      char* base, size_t cap) {
  HIPCHK(hipSetDevice(device_));
  ParameterSet& ps = method(methodName);
  Layout L = buildLayout(ps, tables_, {mol}, 2);
  const bool scratch = base != nullptr;
  const long long nMat = std::max<long long>(L.nMat, 1);
  const int ne = (int)L.elems.size();
  // scratch mode: the one-centre blocks of the present elements, renumbered 0..ne-1
  std::vector<double> oneC;
  if (scratch) {
    const std::vector<double>& all = oneCenterHost(ps);
    oneC.resize((size_t)ne * kOneCenterBlock);
    for (int e = 0; e < ne; ++e) {
      std::copy(all.begin() + (size_t)L.elems[e].w1c * kOneCenterBlock,
                all.begin() + (size_t)(L.elems[e].w1c + 1) * kOneCenterBlock, oneC.begin() + (size_t)e * kOneCenterBlock);
      L.elems[e].w1c = e;
    }
  }
  struct Buffers {
    ElementDev* elems;
    DiatomicDev* diat;
    AtomDev* atoms;
    PairDev* pairs;
    MolDev* mols;
    int *classPairs, *atomList, *molList;
    double *oneC, *ints, *vuvb, *jvec, *S, *H, *P, *F, *erepPair, *ePair, *eAtom, *eMol;
  };
  auto carve = [&](Carver& cv) {
    Buffers b{};
    b.elems = cv.take<ElementDev>(L.elems.size());
    b.diat = cv.take<DiatomicDev>(L.diat.size());
    b.atoms = cv.take<AtomDev>(L.atoms.size());
    b.pairs = cv.take<PairDev>(L.pairs.size());
    b.mols = cv.take<MolDev>(1);
    b.classPairs = cv.take<int>(L.classPairs.size());
    b.atomList = cv.take<int>(L.atoms.size());
    b.molList = cv.take<int>(1);
    b.oneC = scratch ? cv.take<double>(oneC.size()) : nullptr;
    b.ints = cv.take<double>(L.nInts);
    b.vuvb = cv.take<double>(L.nVec);
    b.jvec = cv.take<double>(L.nVec);
    for (double** m : {&b.S, &b.H, &b.P, &b.F}) *m = cv.take<double>(nMat);
    b.erepPair = cv.take<double>(L.pairs.size());
    b.ePair = cv.take<double>(L.pairs.size());
    b.eAtom = cv.take<double>(L.atoms.size());
    b.eMol = cv.take<double>(1);
    return b;
  };
  Carver sizing{nullptr, 0, 0};
  carve(sizing);
  const size_t need = sizing.off + 256;
  if (!scratch) {
    arena_.reserve(need);
    base = arena_.base();
    cap = arena_.capacity();
  } else if (need > cap) {
    throw std::runtime_error("sa_fock_noalloc: this molecule needs " + std::to_string(need) +
                             " bytes of device scratch, the static scratch array holds " + std::to_string(cap) +
                             " (use sa_fock)");
  }
  Carver cv{base, 0, cap};
  const Buffers d = carve(cv);
  hipStream_t s = stream_;
  const size_t nn = (size_t)L.mols[0].n * L.mols[0].n;
  HIPCHK(hipMemcpyAsync(d.elems, L.elems.data(), L.elems.size() * sizeof(ElementDev), hipMemcpyHostToDevice, s));
  if (!L.diat.empty()) HIPCHK(hipMemcpyAsync(d.diat, L.diat.data(), L.diat.size() * sizeof(DiatomicDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.atoms, L.atoms.data(), L.atoms.size() * sizeof(AtomDev), hipMemcpyHostToDevice, s));
  if (!L.pairs.empty()) HIPCHK(hipMemcpyAsync(d.pairs, L.pairs.data(), L.pairs.size() * sizeof(PairDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.mols, L.mols.data(), sizeof(MolDev), hipMemcpyHostToDevice, s));
  if (!L.classPairs.empty())
    HIPCHK(hipMemcpyAsync(d.classPairs, L.classPairs.data(), L.classPairs.size() * sizeof(int), hipMemcpyHostToDevice, s));
  if (scratch) HIPCHK(hipMemcpyAsync(d.oneC, oneC.data(), oneC.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.P, P, nn * sizeof(double), hipMemcpyHostToDevice, s));
  std::vector<int> atomList;  // atoms without d shells first
  for (int pass = 0; pass < 2; ++pass)
    for (size_t a = 0; a < L.atoms.size(); ++a)
      if ((L.elems[L.atoms[a].elem].l == 2) == (pass == 1)) atomList.push_back((int)a);
  int nAtomListSp = 0;
  for (const AtomDev& a : L.atoms) nAtomListSp += L.elems[a.elem].l == 2 ? 0 : 1;
  const int molList = 0;
  HIPCHK(hipMemcpyAsync(d.atomList, atomList.data(), atomList.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.molList, &molList, sizeof(int), hipMemcpyHostToDevice, s));
  KBatch b{};
  b.atomList = d.atomList;
  b.nAtomList = (int)atomList.size();
  b.nAtomListSp = nAtomListSp;
  b.molList = d.molList;
  b.nMolList = 1;
  b.elems = d.elems;
  b.diatomic = d.diat;
  b.nElem = ne;
  b.methodKind = (int)ps.kind();
  b.oneCW = scratch ? d.oneC : oneCenterTable(ps);
  b.atoms = d.atoms;
  b.pairs = d.pairs;
  b.mols = d.mols;
  b.natoms = (int)L.atoms.size();
  b.npairs = (int)L.pairs.size();
  b.nmols = 1;
  b.classPairs = d.classPairs;
  for (int c = 0; c < 10; ++c) b.classStart[c] = L.classStart[c];
  launchTwoCenter(ktab_, b, d.ints, d.vuvb, s);
  launchPairOverlap(b, d.S, d.H, d.erepPair, s);
  launchAtomHcore(ktab_, b, d.vuvb, d.H, s);
  launchMirrorLower(b, d.S, d.H, s);
  launchFockPairs(ktab_, b, d.ints, d.P, d.H, d.F, d.jvec, d.ePair, s);
  launchFockAtoms(ktab_, b, d.P, d.H, d.jvec, d.F, d.eAtom, s);
  launchEnergyReduce(b, d.eAtom, d.ePair, d.eMol, s);
  double e = 0;
  HIPCHK(hipMemcpyAsync(&e, d.eMol, sizeof(double), hipMemcpyDeviceToHost, s));
  if (F) HIPCHK(hipMemcpyAsync(F, d.F, nn * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (Eel) *Eel = e;
}

template <class Launch>
void Engine::timed(const char* name, Launch&& launch) {
  if (prof_.enabled()) prof_.begin(name, stream_);
  launch();
  if (prof_.enabled()) prof_.end(stream_);
}

void Engine::molBuild(MolState& M, ParameterSet& ps, const MoleculeInput& in, bool own) {
  HIPCHK(hipSetDevice(device_));
  Layout L = buildLayout(ps, tables_, {in}, 2, /*closedShell*/ false);
  const long long nMat = std::max<long long>(L.nMat, 1);
  const int n = L.mols[0].n;
  M.ps = &ps;
  M.n = n;
  M.natoms = (int)L.atoms.size();
  M.nel = L.nel[0];
  struct Ptr {
    ElementDev* elems;
    DiatomicDev* diat;
    AtomDev* atoms;
    PairDev* pairs;
    MolDev* mols;
    int *classPairs, *atomList, *molList;
  } d{};
  auto carve = [&](Carver& cv) {
    d.elems = cv.take<ElementDev>(L.elems.size());
    d.diat = cv.take<DiatomicDev>(L.diat.size());
    d.atoms = cv.take<AtomDev>(L.atoms.size());
    d.pairs = cv.take<PairDev>(L.pairs.size());
    d.mols = cv.take<MolDev>(1);
    d.classPairs = cv.take<int>(L.classPairs.size());
    d.atomList = cv.take<int>(L.atoms.size());
    d.molList = cv.take<int>(1);
    M.aoAtom = cv.take<int>(std::max(n, 1));
    M.ints = cv.take<double>(L.nInts);
    M.vuvb = cv.take<double>(L.nVec);
    M.jvec = cv.take<double>(L.nVec);
    for (double** m : {&M.S, &M.H, &M.guess, &M.P, &M.F, &M.Pb, &M.PT, &M.FT, &M.Fa1, &M.Fb1, &M.Fa, &M.Fb})
      *m = cv.take<double>(nMat);
    M.erepPair = cv.take<double>(L.pairs.size());
    M.ePair = cv.take<double>(L.pairs.size());
    M.eAtom = cv.take<double>(L.atoms.size());
    M.eMol = cv.take<double>(1);
    M.nelDev = cv.take<double>(1);
    M.erepDev = cv.take<double>(1);
    M.E = cv.take<double>(1);
  };
  Carver sizing{nullptr, 0, 0};
  carve(sizing);
  const size_t need = sizing.off + 256;
  char* base;
  if (own) {
    HIPCHK(hipMalloc(&M.own, need));
    base = M.own;
  } else {
    arena_.reserve(need);
    base = arena_.base();
  }
  Carver cv{base, 0, need};
  carve(cv);
  hipStream_t s = stream_;
  std::vector<int> aoAtom(std::max(n, 1), 0);
  for (size_t a = 0; a < L.atoms.size(); ++a)
    for (int k = 0; k < L.atoms[a].nao; ++k) aoAtom[L.atoms[a].ao + k] = (int)a;
  std::vector<int> atomList;  // atoms without d shells first
  for (int pass = 0; pass < 2; ++pass)
    for (size_t a = 0; a < L.atoms.size(); ++a)
      if ((L.elems[L.atoms[a].elem].l == 2) == (pass == 1)) atomList.push_back((int)a);
  int nAtomListSp = 0;
  for (const AtomDev& a : L.atoms) nAtomListSp += L.elems[a.elem].l == 2 ? 0 : 1;
  const int molList = 0;
  HIPCHK(hipMemcpyAsync(d.elems, L.elems.data(), L.elems.size() * sizeof(ElementDev), hipMemcpyHostToDevice, s));
  if (!L.diat.empty()) HIPCHK(hipMemcpyAsync(d.diat, L.diat.data(), L.diat.size() * sizeof(DiatomicDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.atoms, L.atoms.data(), L.atoms.size() * sizeof(AtomDev), hipMemcpyHostToDevice, s));
  if (!L.pairs.empty()) HIPCHK(hipMemcpyAsync(d.pairs, L.pairs.data(), L.pairs.size() * sizeof(PairDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.mols, L.mols.data(), sizeof(MolDev), hipMemcpyHostToDevice, s));
  if (!L.classPairs.empty())
    HIPCHK(hipMemcpyAsync(d.classPairs, L.classPairs.data(), L.classPairs.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(M.aoAtom, aoAtom.data(), aoAtom.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.atomList, atomList.data(), atomList.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.molList, &molList, sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(M.nelDev, &M.nel, sizeof(double), hipMemcpyHostToDevice, s));
  KBatch& b = M.b;
  b = KBatch{};
  b.atomList = d.atomList;
  b.nAtomList = (int)atomList.size();
  b.nAtomListSp = nAtomListSp;
  b.molList = d.molList;
  b.nMolList = 1;
  b.elems = d.elems;
  b.diatomic = d.diat;
  b.nElem = (int)L.elems.size();
  b.methodKind = (int)ps.kind();
  b.oneCW = oneCenterTable(ps);
  b.atoms = d.atoms;
  b.pairs = d.pairs;
  b.mols = d.mols;
  b.natoms = (int)L.atoms.size();
  b.npairs = (int)L.pairs.size();
  b.nmols = 1;
  b.classPairs = d.classPairs;
  for (int c = 0; c < 10; ++c) b.classStart[c] = L.classStart[c];
  timed("twocenter", [&] { launchTwoCenter(ktab_, b, M.ints, M.vuvb, s); });
  timed("pair_overlap", [&] { launchPairOverlap(b, M.S, M.H, M.erepPair, s); });
  timed("atom_hcore", [&] { launchAtomHcore(ktab_, b, M.vuvb, M.H, s); });
  launchMirrorLower(b, M.S, M.H, s);
  M.erep = 0;
  if (b.npairs > 0) {
    launchMolPairSum(b, M.erepPair, M.erepDev, s);
    HIPCHK(hipMemcpyAsync(&M.erep, M.erepDev, sizeof(double), hipMemcpyDeviceToHost, s));
  }
  launchDensityGuess(b, M.S, M.nelDev, M.guess, s);
  HIPCHK(hipStreamSynchronize(s));  // host vectors above go out of scope
  if (prof_.enabled()) prof_.collect();
}

MolState* Engine::molCreate(const std::string& methodName, const MoleculeInput& mol) {
  ParameterSet& ps = method(methodName);
  auto M = std::make_unique<MolState>();
  try {
    molBuild(*M, ps, mol, /*own*/ true);
  } catch (...) {
    if (M->own) (void)hipFree(M->own);
    throw;
  }
  return M.release();
}

void Engine::molDestroy(MolState* m) {
  if (!m) return;
  HIPCHK(hipSetDevice(device_));
  HIPCHK(hipStreamSynchronize(stream_));
  if (m->own) HIPCHK(hipFree(m->own));
  delete m;
}

void Engine::molDensityIndependent(MolState* M, double* S, double* H, double* Erep) {
  HIPCHK(hipSetDevice(device_));
  const size_t nn = (size_t)M->n * M->n;
  if (S) HIPCHK(hipMemcpyAsync(S, M->S, nn * sizeof(double), hipMemcpyDeviceToHost, stream_));
  if (H) HIPCHK(hipMemcpyAsync(H, M->H, nn * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIPCHK(hipStreamSynchronize(stream_));
  if (Erep) *Erep = M->erep;
}

void Engine::molDensityGuess(MolState* M, double* P) {
  HIPCHK(hipSetDevice(device_));
  HIPCHK(hipMemcpyAsync(P, M->guess, (size_t)M->n * M->n * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIPCHK(hipStreamSynchronize(stream_));
}

void Engine::molFock(MolState* M, const double* P, double* F, double* Eel) {
  HIPCHK(hipSetDevice(device_));
  hipStream_t s = stream_;
  const size_t nn = (size_t)M->n * M->n;
  HIPCHK(hipMemcpyAsync(M->P, P, nn * sizeof(double), hipMemcpyHostToDevice, s));
  timed("fock_pairs", [&] { launchFockPairs(ktab_, M->b, M->ints, M->P, M->H, M->F, M->jvec, M->ePair, s); });
  timed("fock_atoms", [&] { launchFockAtoms(ktab_, M->b, M->P, M->H, M->jvec, M->F, M->eAtom, s); });
  launchEnergyReduce(M->b, M->eAtom, M->ePair, M->eMol, s);
  double e = 0;
  HIPCHK(hipMemcpyAsync(&e, M->eMol, sizeof(double), hipMemcpyDeviceToHost, s));
  if (F) HIPCHK(hipMemcpyAsync(F, M->F, nn * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (prof_.enabled()) prof_.collect();
  if (Eel) *Eel = e;
}

void Engine::molFockUnrestricted(MolState* M, const double* Pa, const double* Pb, double* Fa, double* Fb, double* Eel) {
  HIPCHK(hipSetDevice(device_));
  hipStream_t s = stream_;
  const size_t nn = (size_t)M->n * M->n;
  HIPCHK(hipMemcpyAsync(M->P, Pa, nn * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(M->Pb, Pb, nn * sizeof(double), hipMemcpyHostToDevice, s));
  launchMatAdd((long long)nn, M->P, M->Pb, M->PT, s);
  // restricted builds F(D) = H + G_R(D) of the total and of each spin density
  const double* dens[3] = {M->PT, M->P, M->Pb};
  double* outs[3] = {M->FT, M->Fa1, M->Fb1};
  for (int k = 0; k < 3; ++k) {
    timed("fock_pairs", [&] { launchFockPairs(ktab_, M->b, M->ints, dens[k], M->H, outs[k], M->jvec, M->ePair, s); });
    timed("fock_atoms", [&] { launchFockAtoms(ktab_, M->b, dens[k], M->H, M->jvec, outs[k], M->eAtom, s); });
  }
  launchUnrestrictedFock(M->b, oneCenterExchangeTable(*M->ps), M->n, M->aoAtom, M->H, M->FT, M->Fa1, M->Fb1, M->PT, M->P,
                         M->Pb, M->Fa, M->Fb, M->E, s);
  double e = 0;
  HIPCHK(hipMemcpyAsync(&e, M->E, sizeof(double), hipMemcpyDeviceToHost, s));
  if (Fa) HIPCHK(hipMemcpyAsync(Fa, M->Fa, nn * sizeof(double), hipMemcpyDeviceToHost, s));
  if (Fb) HIPCHK(hipMemcpyAsync(Fb, M->Fb, nn * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (prof_.enabled()) prof_.collect();
  if (Eel) *Eel = e;
}

void Engine::fockUnrestricted(const std::string& methodName, const MoleculeInput& mol, const double* Pa, const double* Pb,
                              double* Fa, double* Fb, double* Eel) {
  MolState M;  // in the arena: set-up and one build
  molBuild(M, method(methodName), mol, /*own*/ false);
  molFockUnrestricted(&M, Pa, Pb, Fa, Fb, Eel);
}

void Engine::gradients(const std::string& methodName, const MoleculeInput& mol, const double* P, double* grad, double h,
                       const double* Ps) {
  HIPCHK(hipSetDevice(device_));
  if (!(h > 0.0 && h < 0.1)) throw std::runtime_error("gradients: step must be in (0, 0.1) bohr");
  ParameterSet& ps = method(methodName);
  Layout L = buildLayout(ps, tables_, {mol}, 2, /*closedShell*/ Ps == nullptr);
  const long long nMat = std::max<long long>(L.nMat, 1);
  const size_t natoms = L.atoms.size(), npairs = L.pairs.size();
  size_t need = sizeOf<ElementDev>(L.elems.size()) + sizeOf<DiatomicDev>(L.diat.size()) + sizeOf<AtomDev>(natoms) +
                sizeOf<PairDev>(npairs) + sizeOf<MolDev>(1) + sizeOf<int>(L.classPairs.size()) + sizeOf<double>(nMat) +
                sizeOf<double>(6 * std::max<size_t>(npairs, 1)) + sizeOf<double>(3 * std::max<size_t>(natoms, 1)) +
                (Ps ? sizeOf<double>(nMat) : 0);
  arena_.reserve(need);
  Carver cv{arena_.base(), 0, arena_.capacity()};
  ElementDev* dElems = cv.take<ElementDev>(L.elems.size());
  DiatomicDev* dDiat = cv.take<DiatomicDev>(L.diat.size());
  AtomDev* dAtoms = cv.take<AtomDev>(natoms);
  PairDev* dPairs = cv.take<PairDev>(npairs);
  MolDev* dMols = cv.take<MolDev>(1);
  int* dClassPairs = cv.take<int>(L.classPairs.size());
  double* dP = cv.take<double>(nMat);
  double* dEpair = cv.take<double>(6 * std::max<size_t>(npairs, 1));
  double* dGrad = cv.take<double>(3 * std::max<size_t>(natoms, 1));
  double* dPs = Ps ? cv.take<double>(nMat) : nullptr;
  hipStream_t s = stream_;
  const size_t nn = (size_t)L.mols[0].n * L.mols[0].n;
  HIPCHK(hipMemcpyAsync(dElems, L.elems.data(), L.elems.size() * sizeof(ElementDev), hipMemcpyHostToDevice, s));
  if (!L.diat.empty()) HIPCHK(hipMemcpyAsync(dDiat, L.diat.data(), L.diat.size() * sizeof(DiatomicDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dAtoms, L.atoms.data(), natoms * sizeof(AtomDev), hipMemcpyHostToDevice, s));
  if (npairs) HIPCHK(hipMemcpyAsync(dPairs, L.pairs.data(), npairs * sizeof(PairDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dMols, L.mols.data(), sizeof(MolDev), hipMemcpyHostToDevice, s));
  if (!L.classPairs.empty())
    HIPCHK(hipMemcpyAsync(dClassPairs, L.classPairs.data(), L.classPairs.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dP, P, nn * sizeof(double), hipMemcpyHostToDevice, s));
  if (Ps) HIPCHK(hipMemcpyAsync(dPs, Ps, nn * sizeof(double), hipMemcpyHostToDevice, s));
  KBatch b{};
  b.elems = dElems;
  b.diatomic = dDiat;
  b.nElem = (int)L.elems.size();
  b.methodKind = (int)ps.kind();
  b.oneCW = oneCenterTable(ps);
  b.atoms = dAtoms;
  b.pairs = dPairs;
  b.mols = dMols;
  b.natoms = (int)natoms;
  b.npairs = (int)npairs;
  b.nmols = 1;
  b.classPairs = dClassPairs;
  for (int c = 0; c < 10; ++c) b.classStart[c] = L.classStart[c];
  launchGradients(ktab_, b, dP, dPs, h, dEpair, dGrad, s);
  HIPCHK(hipMemcpyAsync(grad, dGrad, 3 * natoms * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
}

std::array<double, 4> Engine::workCounts(const std::string& methodName, const std::vector<int>& Z) {
  ParameterSet& ps = method(methodName);
  const ModelTables& t = tables_;
  auto dimOf = [](int l) { return l == 0 ? 1 : l == 1 ? 10 : 40; };
  // per (lA, lB) class: point-charge pairs of the multipole instances (16 flops each: offsets,
  // squared distance, reciprocal square root, weighted sum), local-integral terms (2 each) and
  // the two rotation contractions X = L T^T, G = T X (5 per term: coefficient and multiply-add)
  double cls[9];
  for (int lA = 0; lA < 3; ++lA)
    for (int lB = 0; lB < 3; ++lB) {
      const LocalClassDev& C = t.classes[lA * 3 + lB];
      double f = 0;
      for (int i = 0; i < C.nInst; ++i) f += 16.0 * t.instances[C.inst0 + i].nCharge;
      for (int e = 0; e < C.nEntry; ++e) f += 2.0 * t.entries[C.entry0 + e].nTerm;
      const int DA = dimOf(lA), DB = dimOf(lB);
      double sQ = 0, sP = 0;
      for (int Q = 0; Q < DB; ++Q) sQ += t.blockLen[t.pairBlock[Q]];
      for (int P = 0; P < DA; ++P) sP += t.blockLen[t.pairBlock[P]];
      cls[lA * 3 + lB] = f + 5.0 * (DA * sQ + DB * sP);
    }
  std::vector<int> l;
  long long n = 0, oneCentre = 0;
  for (int z : Z) {
    if (!ps.hasElement(z)) throw std::runtime_error("parameters do not exist for element Z=" + std::to_string(z));
    const ElementDev& e = ps.element(z, tables_);
    l.push_back(e.l);
    n += e.nao;
    oneCentre += e.l == 0 ? 1 : e.l == 1 ? 6 : 58;
  }
  double flops = 0, ints = 0, pairs = 0;
  for (size_t a = 0; a < l.size(); ++a)
    for (size_t b = a + 1; b < l.size(); ++b) {
      flops += cls[l[a] * 3 + l[b]];
      ints += (double)dimOf(l[a]) * dimOf(l[b]);
      pairs += 1;
    }
  // G-build: every integral block and unique one-centre integral read once, the lower triangle
  // of P read and of F written
  const double bytes = 8.0 * (ints + (double)oneCentre) + 8.0 * (double)n * (double)(n + 1);
  return {flops, bytes, pairs, (double)n};
}

void Engine::twoCenterBlock(const std::string& methodName, int ZA, int ZB, const double* Rab, double* out) {
  MoleculeInput m;
  m.Z = {ZA, ZB};
  m.xyz = {0, 0, 0, Rab[0], Rab[1], Rab[2]};
  HIPCHK(hipSetDevice(device_));
  ParameterSet& ps = method(methodName);
  // one pair, no diatomic parameters needed for the two-centre block itself
  Layout L;
  {
    const ElementDev& ea = ps.element(ZA, tables_);
    const ElementDev& eb = ps.element(ZB, tables_);
    L.elems = {ea, eb};
    L.diat.assign(4, DiatomicDev{});
    AtomDev a{0, 0, 0, ea.nao, 0, 0, 0}, c{1, 0, ea.nao, eb.nao, Rab[0], Rab[1], Rab[2]};
    L.atoms = {a, c};
    PairDev p{};
    p.a = 0;
    p.b = 1;
    p.cls = ea.l * 3 + eb.l;
    p.intOff = 0;
    p.vecOff = 0;
    p.aoA = 0;
    p.aoB = ea.nao;
    p.n = ea.nao + eb.nao;
    p.matOff = 0;
    L.pairs = {p};
    const int d2[2] = {dimOf(ea.l), dimOf(eb.l)};
    setCoulombPartialOffsets(L.atoms.data(), L.pairs.data(), 2, d2, 0);
    MolDev md{};
    md.n = ea.nao + eb.nao;
    md.natoms = 2;
    md.npairs = 1;
    L.mols = {md};
    L.classPairs = {0};
    for (int k = 0; k <= 9; ++k) L.classStart[k] = k <= p.cls ? 0 : 1;
  }
  const int dA = dimOf(L.elems[0].l), dB = dimOf(L.elems[1].l);
  arena_.reserve(sizeOf<ElementDev>(2) + sizeOf<DiatomicDev>(4) + sizeOf<AtomDev>(2) + sizeOf<PairDev>(1) + sizeOf<MolDev>(1) +
                 sizeOf<int>(1) + sizeOf<double>(dA * dB) + sizeOf<double>(dA + dB));
  Carver cv{arena_.base(), 0, arena_.capacity()};
  ElementDev* dElems = cv.take<ElementDev>(2);
  DiatomicDev* dDiat = cv.take<DiatomicDev>(4);
  AtomDev* dAtoms = cv.take<AtomDev>(2);
  PairDev* dPairs = cv.take<PairDev>(1);
  MolDev* dMols = cv.take<MolDev>(1);
  int* dCP = cv.take<int>(1);
  double* dInts = cv.take<double>(dA * dB);
  double* dV = cv.take<double>(dA + dB);
  hipStream_t s = stream_;
  HIPCHK(hipMemcpyAsync(dElems, L.elems.data(), 2 * sizeof(ElementDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dDiat, L.diat.data(), 4 * sizeof(DiatomicDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dAtoms, L.atoms.data(), 2 * sizeof(AtomDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dPairs, L.pairs.data(), sizeof(PairDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dMols, L.mols.data(), sizeof(MolDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dCP, L.classPairs.data(), sizeof(int), hipMemcpyHostToDevice, s));
  KBatch b{};
  b.elems = dElems;
  b.diatomic = dDiat;
  b.nElem = 2;
  b.methodKind = (int)ps.kind();
  b.oneCW = oneCenterTable(ps);
  b.atoms = dAtoms;
  b.pairs = dPairs;
  b.mols = dMols;
  b.natoms = 2;
  b.npairs = 1;
  b.nmols = 1;
  b.classPairs = dCP;
  for (int c = 0; c < 10; ++c) b.classStart[c] = L.classStart[c];
  launchTwoCenter(ktab_, b, dInts, dV, s);
  HIPCHK(hipMemcpyAsync(out, dInts, (size_t)dA * dB * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
}

constexpr int kWideSytrd = 192;  // order from which sa_eigh tridiagonalises with many workgroups
constexpr int kGemmOrmtr = 128;  // order from which sa_eigh back-transforms with GEMMs

bool Engine::densityProjector(int n, int nocc, const double* F, double* X, int* stepsOut) {
  HIPCHK(hipSetDevice(device_));
  if (n <= 0) throw std::runtime_error("density projector: empty matrix");
  if (nocc < 0 || nocc > n) throw std::runtime_error("density projector: nocc out of range");
  const long long nn = (long long)n * n;
  const int tile = sp2Tile(n, false);
  const int tn = (n + tile - 1) / tile, ntiles = tn * (tn + 1) / 2;
  const int maxSteps = ScfSettings{}.sp2MaxSteps;
  arena_.reserve(4 * sizeOf<double>(nn) + 2 * sizeOf<double>(n) + 2 * sizeOf<double>(2 * (size_t)ntiles) +
                 sizeOf<Sp2Dev>(2) + sizeOf<MolDev>(1) + sizeOf<Sp2StepJob>(2) + sizeOf<int>(2) + sizeOf<double>(5 * (size_t)n) +
                 sizeOf<double>(sp2LanczosCoefs()));
  Carver cv{arena_.base(), 0, arena_.capacity()};
  double* dF = cv.take<double>(nn);
  double* dX0 = cv.take<double>(nn);
  double* dX1 = cv.take<double>(nn);
  double* dP = cv.take<double>(nn);
  double* dLo = cv.take<double>(n);
  double* dHi = cv.take<double>(n);
  double* dPart[2] = {cv.take<double>(2 * (size_t)ntiles), cv.take<double>(2 * (size_t)ntiles)};
  Sp2Dev* dSt[2] = {cv.take<Sp2Dev>(1), cv.take<Sp2Dev>(1)};
  MolDev* dMol = cv.take<MolDev>(1);
  Sp2StepJob* dJob[2] = {cv.take<Sp2StepJob>(1), cv.take<Sp2StepJob>(1)};
  int* dTiles = cv.take<int>(2);
  double* dLz = cv.take<double>(5 * (size_t)n);
  double* dLzCoef = cv.take<double>(sp2LanczosCoefs());
  MolDev md{};
  md.n = n;
  md.nocc = nocc;
  Sp2StepJob j0{};
  j0.n = n;
  j0.mol = 0;
  j0.tile = tile;
  j0.X = dX0;
  j0.Y = dX1;
  Sp2StepJob j1 = j0;
  j1.X = dX1;
  j1.Y = dX0;
  const int tiles[2] = {0, ntiles};
  const int nt32 = tile == 32 ? ntiles : 0, nt64 = tile == 64 ? ntiles : 0, nt128 = tile == 128 ? ntiles : 0;
  hipStream_t s = stream_;
  HIPCHK(hipMemcpyAsync(dF, F, nn * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dMol, &md, sizeof(MolDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dJob[0], &j0, sizeof(Sp2StepJob), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dJob[1], &j1, sizeof(Sp2StepJob), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dTiles, tiles, sizeof(tiles), hipMemcpyHostToDevice, s));
  // the SCF's start: Gershgorin discs, narrowed by Lanczos for 400 AOs and more
  launchSp2Init(dMol, 1, n, dF, dJob[0], dTiles, ntiles, dLo, dHi, dPart[0], dSt[0], maxSteps, nullptr, s, dLz, dLzCoef,
                400);
  Sp2Dev st{};
  int steps = 0;
  for (;;) {  // a first chunk of 24 steps, then chunks of 8, each followed by the controller's state
    const int chunk = steps == 0 ? 24 : 8;  // steps after the end return at once
    for (int c = 0; c < chunk && steps < maxSteps; ++c, ++steps) {
      const int e = steps % 2;
      launchSp2Step(dJob[e], dTiles, 1, nt32, nt64, nt128, dPart[e], dPart[1 - e], dSt[e], dSt[1 - e], s);
    }
    const int e = steps % 2;
    launchSp2Decide(dJob[e], dTiles, 1, dPart[e], dSt[e], dSt[1 - e], s);
    HIPCHK(hipMemcpyAsync(&st, dSt[1 - e], sizeof(Sp2Dev), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (st.done || steps >= maxSteps) break;
  }
  launchPFromX(dMol, 1, dSt[1 - steps % 2], dX0, dX1, dP, s);  // P = 2 X
  HIPCHK(hipMemcpyAsync(X, dP, nn * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (long long k = 0; k < nn; ++k) X[k] *= 0.5;  // exact
  if (stepsOut) *stepsOut = st.doneStep;
  return st.done && st.converged;
}

void Engine::eigh(int n, int nvec, const double* A, double* w, double* V) {
  HIPCHK(hipSetDevice(device_));
  if (n <= 0) return;
  if (nvec < 0 || nvec > n) throw std::runtime_error("eigh: nvec out of range");
  const long long nn = (long long)n * n;
  const int ngemm = ormtrGemmJobs(n);
  arena_.reserve(2 * sizeOf<double>(nn) + 5 * sizeOf<double>(nn) + sizeOf<unsigned char>(nn) + 8 * sizeOf<double>(2 * n) +
                 sizeOf<EigJob>(1) + sizeOf<GemmJob>(std::max(ngemm, 1)) + sizeOf<int>(1));
  Carver cv{arena_.base(), 0, arena_.capacity()};
  EigJob J{};
  J.n = n;
  J.nvec = nvec;
  J.A = cv.take<double>(nn);
  J.Z = cv.take<double>(nn);
  J.work = cv.take<double>(5 * nn);
  J.ipiv = cv.take<unsigned char>(nn);
  J.d = cv.take<double>(n);
  J.e = cv.take<double>(n);
  J.tau = cv.take<double>(n);
  J.v = cv.take<double>(2 * n);
  J.w = cv.take<double>(2 * n);
  J.wval = cv.take<double>(n);
  EigJob* dJ = cv.take<EigJob>(1);
  GemmJob* dG = cv.take<GemmJob>(std::max(ngemm, 1));
  int* dZero = cv.take<int>(1);
  std::vector<GemmJob> hG(std::max(ngemm, 1));
  hipStream_t s = stream_;
  HIPCHK(hipMemsetAsync(dZero, 0, sizeof(int), s));
  HIPCHK(hipMemcpyAsync(J.A, A, nn * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dJ, &J, sizeof(EigJob), hipMemcpyHostToDevice, s));
  // large matrices: the trailing updates over many workgroups; small ones: one workgroup
  timed("sytrd", [&] { n >= kWideSytrd ? launchSytrdWide(dJ, n, s) : launchSytrd(dJ, 1, s); });
  timed("stebz", [&] { launchStebz(dJ, 1, n, s); });
  timed("stein", [&] { launchStein(dJ, 1, nvec, s); });
  // large matrices: blocks of reflectors applied by matrix-core GEMMs
  timed("ormtr", [&] { n >= kGemmOrmtr ? launchOrmtrGemm(J, dJ, hG.data(), dG, dZero, s) : launchOrmtr(dJ, 1, n, nvec, s); });
  std::vector<double> Z((size_t)n * std::max(nvec, 1));
  HIPCHK(hipMemcpyAsync(w, J.wval, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (nvec > 0) HIPCHK(hipMemcpyAsync(Z.data(), J.Z, (size_t)n * nvec * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (prof_.enabled()) prof_.collect();
  if (V)
    for (int r = 0; r < n; ++r)
      for (int k = 0; k < nvec; ++k) V[(size_t)k * n + r] = Z[(size_t)r * nvec + k];
}

void Engine::gemmNT(int M, int N, int K, const double* A, const double* B, double* C) {
  // C (M x N, col-major) = A (M x K, col-major) * B (N x K, col-major)^T
  HIPCHK(hipSetDevice(device_));
  arena_.reserve(sizeOf<double>((size_t)M * K) + sizeOf<double>((size_t)N * K) + sizeOf<double>((size_t)M * N) +
                 sizeOf<GemmJob>(1) + sizeOf<int>(2));
  Carver cv{arena_.base(), 0, arena_.capacity()};
  double* dA = cv.take<double>((size_t)M * K);
  double* dB = cv.take<double>((size_t)N * K);
  double* dC = cv.take<double>((size_t)M * N);
  GemmJob* dJ = cv.take<GemmJob>(1);
  int* dT = cv.take<int>(2);
  GemmJob g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = dA;
  g.B = dB;
  g.C = dC;
  g.sAi = 1;
  g.sAk = M;
  g.sBj = 1;
  g.sBk = N;
  g.sCi = 1;
  g.sCj = M;
  g.D = dC;
  g.sDi = 1;
  g.sDj = M;
  g.alpha = 1.0;
  g.beta = 0.0;
  const int tile = (M >= 128 && N >= 128) ? 64 : 32;
  int tiles[2] = {0, ((M + tile - 1) / tile) * ((N + tile - 1) / tile)};
  hipStream_t s = stream_;
  HIPCHK(hipMemcpyAsync(dA, A, (size_t)M * K * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dB, B, (size_t)N * K * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dJ, &g, sizeof(GemmJob), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dT, tiles, sizeof(tiles), hipMemcpyHostToDevice, s));
  launchGemm(dJ, dT, 1, tiles[1], tile, s);
  HIPCHK(hipMemcpyAsync(C, dC, (size_t)M * N * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------------------------
// CIS sigma vectors (N/TimeDependent/LinearResponse/CISSigmaVectorEvaluator.cpp:96-165)
// ---------------------------------------------------------------------------------------------
struct CisSpin {
  int nocc = 0, nvir = 0, nconf = 0, off = 0;  // off: start of this spin's part of a trial vector
  double* dC = nullptr;                        // n x n orbitals (resident)
};

struct CisState {
  int n = 0, nconf = 0, nspin = 1, device = 0;
  bool triplet = false;
  double c1 = 1.0, c2 = 1.0;
  long long nVec = 0;
  CisSpin spin[2];
  KBatch b{};
  const double* WK = nullptr;
  void* resident = nullptr;  // layout, integrals, orbitals, energy differences
  double* dInts = nullptr;
  double* dDE = nullptr;     // e_a - e_i of every configuration (nconf)
  void* work = nullptr;      // buffers of one block of vectors
  int blockCap = 0;
  double *dX = nullptr, *dS = nullptr, *dJvec = nullptr;
  double *dT[2] = {}, *dD[2] = {}, *dF[2] = {}, *dU[2] = {};
  double* dDJ = nullptr;     // D^alpha + D^beta (unrestricted)
  GemmJob* dJobs = nullptr;
  int* dTiles = nullptr;
  double* pinned = nullptr;  // page-locked staging of one block of trial and sigma vectors (host path)
};

namespace {

// Vectors per block: at most 64, and the block's n x n stacks (D and F per spin, D^alpha + D^beta)
// within about 2 GiB.
int cisBlock(int n, int nspin, int nvec) {
  const double perVec = 8.0 * n * (double)n * (nspin == 2 ? 5 : 2);
  const int byMem = (int)std::max(1.0, std::floor(2147483648.0 / perVec));
  return std::max(1, std::min({nvec, 64, byMem}));
}

GemmJob cisJob(int M, int N, int K, const double* A, long long sAi, long long sAk, const double* B, long long sBj,
               long long sBk, double* C, long long sCi, long long sCj, double beta) {
  GemmJob g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.B = B;
  g.C = C;
  g.D = C;
  g.sAi = sAi;
  g.sAk = sAk;
  g.sBj = sBj;
  g.sBk = sBk;
  g.sCi = g.sDi = sCi;
  g.sCj = g.sDj = sCj;
  g.alpha = 1.0;
  g.beta = beta;
  return g;
}

}  // namespace

CisState* Engine::cisCreate(const std::string& methodName, const MoleculeInput& mol, const double* C, const double* eps,
                            int nocc, bool triplet, double c1, double c2) {
  return cisCreateSpins(methodName, mol, 1, &C, &eps, &nocc, triplet, c1, c2);
}

CisState* Engine::cisCreateUnrestricted(const std::string& methodName, const MoleculeInput& mol, const double* Ca,
                                        const double* Cb, const double* epsA, const double* epsB, int noccA, int noccB,
                                        double c1, double c2) {
  const double* C[2] = {Ca, Cb};
  const double* e[2] = {epsA, epsB};
  const int no[2] = {noccA, noccB};
  return cisCreateSpins(methodName, mol, 2, C, e, no, false, c1, c2);
}

CisState* Engine::cisCreateSpins(const std::string& methodName, const MoleculeInput& mol, int nspin, const double* const* C,
                                 const double* const* eps, const int* nocc, bool triplet, double c1, double c2) {
  HIPCHK(hipSetDevice(device_));
  ParameterSet& ps = method(methodName);
  Layout L = buildLayout(ps, tables_, {mol}, 2, /*closedShell*/ false);
  const int n = L.mols[0].n;
  auto st = std::make_unique<CisState>();
  st->n = n;
  st->nspin = nspin;
  st->device = device_;
  st->triplet = triplet;
  st->c1 = c1;
  st->c2 = c2;
  st->nVec = std::max<long long>(L.nVec, 1);
  for (int sp = 0; sp < nspin; ++sp) {
    if (nocc[sp] <= 0 || nocc[sp] >= n) throw std::runtime_error("CIS: need 0 < nocc < number of orbitals");
    CisSpin& S = st->spin[sp];
    S.nocc = nocc[sp];
    S.nvir = n - nocc[sp];
    S.nconf = S.nvir * S.nocc;
    S.off = st->nconf;
    st->nconf += S.nconf;
  }
  const size_t nn = (size_t)n * n;
  struct P {
    ElementDev* elems;
    DiatomicDev* diat;
    AtomDev* atoms;
    PairDev* pairs;
    MolDev* mols;
    int* classPairs;
    int* atomList;
    int* molList;
    double* ints;
    double* vuvb;
    double* C[2];
    double* dE;
  };
  auto carve = [&](Carver& cv) {
    P p{};
    p.elems = cv.take<ElementDev>(L.elems.size());
    p.diat = cv.take<DiatomicDev>(L.diat.size());
    p.atoms = cv.take<AtomDev>(L.atoms.size());
    p.pairs = cv.take<PairDev>(L.pairs.size());
    p.mols = cv.take<MolDev>(1);
    p.classPairs = cv.take<int>(L.classPairs.size());
    p.atomList = cv.take<int>(L.atoms.size());
    p.molList = cv.take<int>(1);
    p.ints = cv.take<double>(L.nInts);
    p.vuvb = cv.take<double>(L.nVec);
    for (int sp = 0; sp < nspin; ++sp) p.C[sp] = cv.take<double>(nn);
    p.dE = cv.take<double>(st->nconf);
    return p;
  };
  Carver count{nullptr, 0, 0};
  carve(count);
  HIPCHK(hipMalloc(&st->resident, count.off));
  Carver cv{reinterpret_cast<char*>(st->resident), 0, count.off};
  P p = carve(cv);
  hipStream_t s = stream_;
  HIPCHK(hipMemcpyAsync(p.elems, L.elems.data(), L.elems.size() * sizeof(ElementDev), hipMemcpyHostToDevice, s));
  if (!L.diat.empty()) HIPCHK(hipMemcpyAsync(p.diat, L.diat.data(), L.diat.size() * sizeof(DiatomicDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(p.atoms, L.atoms.data(), L.atoms.size() * sizeof(AtomDev), hipMemcpyHostToDevice, s));
  if (!L.pairs.empty()) HIPCHK(hipMemcpyAsync(p.pairs, L.pairs.data(), L.pairs.size() * sizeof(PairDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(p.mols, L.mols.data(), sizeof(MolDev), hipMemcpyHostToDevice, s));
  if (!L.classPairs.empty())
    HIPCHK(hipMemcpyAsync(p.classPairs, L.classPairs.data(), L.classPairs.size() * sizeof(int), hipMemcpyHostToDevice, s));
  std::vector<int> atomList;  // atoms without d shells first
  for (int pass = 0; pass < 2; ++pass)
    for (size_t a = 0; a < L.atoms.size(); ++a)
      if ((L.elems[L.atoms[a].elem].l == 2) == (pass == 1)) atomList.push_back((int)a);
  int nAtomListSp = 0;
  for (const AtomDev& a : L.atoms) nAtomListSp += L.elems[a.elem].l == 2 ? 0 : 1;
  const int molList = 0;
  HIPCHK(hipMemcpyAsync(p.atomList, atomList.data(), atomList.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(p.molList, &molList, sizeof(int), hipMemcpyHostToDevice, s));
  std::vector<double> dE(st->nconf);
  for (int sp = 0; sp < nspin; ++sp) {
    const CisSpin& S = st->spin[sp];
    for (int i = 0; i < S.nocc; ++i)
      for (int a = 0; a < S.nvir; ++a) dE[S.off + (size_t)i * S.nvir + a] = eps[sp][S.nocc + a] - eps[sp][i];
    HIPCHK(hipMemcpyAsync(p.C[sp], C[sp], nn * sizeof(double), hipMemcpyHostToDevice, s));
    st->spin[sp].dC = p.C[sp];
  }
  HIPCHK(hipMemcpyAsync(p.dE, dE.data(), dE.size() * sizeof(double), hipMemcpyHostToDevice, s));
  KBatch& b = st->b;
  b.atomList = p.atomList;
  b.nAtomList = (int)atomList.size();
  b.nAtomListSp = nAtomListSp;
  b.molList = p.molList;
  b.nMolList = 1;
  b.elems = p.elems;
  b.diatomic = p.diat;
  b.nElem = (int)L.elems.size();
  b.methodKind = (int)ps.kind();
  b.oneCW = oneCenterTable(ps);
  b.atoms = p.atoms;
  b.pairs = p.pairs;
  b.mols = p.mols;
  b.natoms = (int)L.atoms.size();
  b.npairs = (int)L.pairs.size();
  b.nmols = 1;
  b.classPairs = p.classPairs;
  for (int c = 0; c < 10; ++c) b.classStart[c] = L.classStart[c];
  st->WK = oneCenterExchangeTable(ps);
  st->dInts = p.ints;
  st->dDE = p.dE;
  launchTwoCenter(ktab_, b, p.ints, p.vuvb, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return st.release();
}

int Engine::cisConfigurations(const CisState* st) const { return st ? st->nconf : 0; }

void Engine::cisDestroy(CisState* st) {
  if (!st) return;
  HIPCHK(hipSetDevice(st->device));
  HIPCHK(hipStreamSynchronize(stream_));
  if (st->work) (void)hipFree(st->work);
  if (st->pinned) (void)hipHostFree(st->pinned);
  if (st->resident) (void)hipFree(st->resident);
  delete st;
}

void Engine::cisSigma(CisState* st, int nvec, const double* X, double* S, double* deviceMs, bool deviceBuffers) {
  if (!st) throw std::runtime_error("CIS: no evaluator");
  if (nvec < 0) throw std::runtime_error("CIS: negative number of vectors");
  HIPCHK(hipSetDevice(device_));
  if (st->device != device_) throw std::runtime_error("CIS: evaluator belongs to another device");
  if (nvec == 0) {
    if (deviceMs) *deviceMs = 0.0;
    return;
  }
  const int n = st->n, nconf = st->nconf, nspin = st->nspin;
  const long long nn = (long long)n * n;
  const int nb = cisBlock(n, nspin, nvec);
  const int nj = nspin * nb;  // jobs per GEMM stage
  hipStream_t s = stream_;
  if (nb > st->blockCap) {
    if (st->work) {
      HIPCHK(hipStreamSynchronize(s));
      HIPCHK(hipFree(st->work));
      st->work = nullptr;
    }
    if (st->pinned) {
      HIPCHK(hipHostFree(st->pinned));
      st->pinned = nullptr;
    }
    auto carve = [&](Carver& cv) {
      st->dX = cv.take<double>((size_t)nconf * nb);
      st->dS = cv.take<double>((size_t)nconf * nb);
      for (int sp = 0; sp < nspin; ++sp) {
        st->dT[sp] = cv.take<double>((size_t)n * st->spin[sp].nvir * nb);
        st->dD[sp] = cv.take<double>((size_t)nn * nb);
        st->dF[sp] = cv.take<double>((size_t)nn * nb);
        st->dU[sp] = cv.take<double>((size_t)n * st->spin[sp].nocc * nb);
      }
      st->dDJ = nspin == 2 ? cv.take<double>((size_t)nn * nb) : nullptr;
      st->dJvec = cv.take<double>((size_t)st->nVec * nb);
      st->dJobs = cv.take<GemmJob>(4 * (size_t)nj);
      st->dTiles = cv.take<int>(4 * ((size_t)nj + 1));
    };
    Carver count{nullptr, 0, 0};
    carve(count);
    HIPCHK(hipMalloc(&st->work, count.off));
    Carver cv{reinterpret_cast<char*>(st->work), 0, count.off};
    carve(cv);
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&st->pinned), 2 * (size_t)nconf * nb * sizeof(double),
                         hipHostMallocDefault));
    st->blockCap = nb;
  }
  // restricted singlet cJ = 2 c1, triplet cJ = 0; unrestricted cJ = c1 on D^alpha + D^beta
  const double cJ = nspin == 2 ? st->c1 : (st->triplet ? 0.0 : 2.0 * st->c1), cK = st->c2;
  const int cap = st->blockCap;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  float msTotal = 0.0f;
  std::vector<GemmJob> jobs(4 * (size_t)nj);
  std::vector<int> tiles(4 * ((size_t)nj + 1));
  for (int v0 = 0; v0 < nvec; v0 += nb) {
    const int nv = std::min(nb, nvec - v0);
    // trial and sigma vectors of the block: the caller's device buffers, or the staging buffers
    const double* dXb = deviceBuffers ? X + (size_t)v0 * nconf : st->dX;
    double* dSb = deviceBuffers ? S + (size_t)v0 * nconf : st->dS;
    double* pinX = st->pinned;
    double* pinS = st->pinned + (size_t)nconf * cap;
    if (!deviceBuffers) {
      std::memcpy(pinX, X + (size_t)v0 * nconf, (size_t)nconf * nv * sizeof(double));
      HIPCHK(hipMemcpyAsync(st->dX, pinX, (size_t)nconf * nv * sizeof(double), hipMemcpyHostToDevice, s));
    }
    // the four products of every vector and spin of the block: T = C_occ x^T, D = T C_vir^T,
    // U = F C_occ, S += C_vir^T U (jobs of one stage ordered by spin, then vector)
    int tileOf[4];
    const int njb = nspin * nv;
    for (int st4 = 0; st4 < 4; ++st4) {
      int* ts = tiles.data() + st4 * (nj + 1);
      ts[0] = 0;
      // one tile edge per launch: 64 when M and N of the stage are at least 128 for every spin
      bool big = true;
      for (int sp = 0; sp < nspin; ++sp) {
        const CisSpin& P = st->spin[sp];
        const int M = st4 == 3 ? P.nvir : n;
        const int N = st4 == 0 ? P.nvir : st4 == 1 ? n : P.nocc;
        big = big && M >= 128 && N >= 128;
      }
      tileOf[st4] = big ? 64 : 32;
      for (int sp = 0; sp < nspin; ++sp) {
        const CisSpin& P = st->spin[sp];
        const double* Co = P.dC;
        const double* Cv = P.dC + (long long)P.nocc * n;
        for (int k = 0; k < nv; ++k) {
          const int j = sp * nv + k;
          const double* Xk = dXb + (size_t)k * nconf + P.off;
          double* Sk = dSb + (size_t)k * nconf + P.off;
          double* Tk = st->dT[sp] + (size_t)k * n * P.nvir;
          double* Dk = st->dD[sp] + (size_t)k * nn;
          double* Fk = st->dF[sp] + (size_t)k * nn;
          double* Uk = st->dU[sp] + (size_t)k * n * P.nocc;
          GemmJob g;
          if (st4 == 0) g = cisJob(n, P.nvir, P.nocc, Co, 1, n, Xk, 1, P.nvir, Tk, 1, n, 0.0);
          else if (st4 == 1) g = cisJob(n, n, P.nvir, Tk, 1, n, Cv, 1, n, Dk, 1, n, 0.0);
          else if (st4 == 2) g = cisJob(n, P.nocc, n, Fk, 1, n, Co, n, 1, Uk, 1, n, 0.0);
          else g = cisJob(P.nvir, P.nocc, n, Cv, n, 1, Uk, n, 1, Sk, 1, P.nvir, 1.0);
          jobs[(size_t)st4 * nj + j] = g;
          const int t = tileOf[st4];
          ts[j + 1] = ts[j] + ((g.M + t - 1) / t) * ((g.N + t - 1) / t);
        }
      }
    }
    HIPCHK(hipMemcpyAsync(st->dJobs, jobs.data(), jobs.size() * sizeof(GemmJob), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(st->dTiles, tiles.data(), tiles.size() * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(e0, s));
    auto gemm = [&](int st4) {
      const int* ts = tiles.data() + st4 * (nj + 1);
      launchGemm(st->dJobs + (size_t)st4 * nj, st->dTiles + st4 * (nj + 1), njb, ts[njb], tileOf[st4], s);
    };
    gemm(0);
    gemm(1);
    if (nspin == 1) {
      launchCisFock(ktab_, st->b, st->dInts, st->WK, nv, st->dD[0], st->dD[0], nn, cJ, cK, st->dJvec, st->nVec, st->dF[0], s);
    } else {
      launchMatAdd(nn * nv, st->dD[0], st->dD[1], st->dDJ, s);
      for (int sp = 0; sp < 2; ++sp)
        launchCisFock(ktab_, st->b, st->dInts, st->WK, nv, st->dDJ, st->dD[sp], nn, cJ, cK, st->dJvec, st->nVec, st->dF[sp],
                      s);
    }
    gemm(2);
    launchCisDiag(nconf, nv, st->dDE, dXb, dSb, s);
    gemm(3);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, s));
    if (!deviceBuffers)
      HIPCHK(hipMemcpyAsync(pinS, st->dS, (size_t)nconf * nv * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (!deviceBuffers) std::memcpy(S + (size_t)v0 * nconf, pinS, (size_t)nconf * nv * sizeof(double));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    msTotal += ms;
  }
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  if (deviceMs) *deviceMs = msTotal;
}

}  // namespace sparrow_amd
