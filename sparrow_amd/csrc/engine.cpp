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

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <stdexcept>

namespace sparrow_amd {

#define HIPCHK(x)                                                                                               \
  do {                                                                                                          \
    hipError_t e_ = (x);                                                                                        \
    if (e_ != hipSuccess)                                                                                       \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " (" #x ") at " __FILE__ ":" + \
                               std::to_string(__LINE__));                                                       \
  } while (0)

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

double nowSeconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
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

Layout buildLayout(ParameterSet& ps, const ModelTables& t, const std::vector<MoleculeInput>& mols, int slots) {
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
    if (ne2 < 0 || ne2 % 2 != 0) throw std::runtime_error("restricted closed-shell SCF needs an even, non-negative electron count");
    md.nocc = (int)(ne2 / 2);
    if (md.nocc > md.n) throw std::runtime_error("more electrons than spin orbitals");
    L.nel.push_back(nel);
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
        L.nInts += (long long)dimOf(lA) * dimOf(lB);
        L.nVec += dimOf(lA) + dimOf(lB);
        byClass[p.cls].push_back((int)L.pairs.size());
        L.pairs.push_back(p);
      }
    md.npairs = (int)L.pairs.size() - md.pair0;
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

// Bump allocator over the arena.
struct Carver {
  char* base;
  size_t off = 0, cap;
  template <class T>
  T* take(size_t count) {
    size_t bytes = std::max<size_t>(count * sizeof(T), 8);
    off = (off + 255) & ~size_t(255);
    T* p = reinterpret_cast<T*>(base + off);
    off += bytes;
    return p;
  }
};

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
}

Engine::~Engine() {
  for (void* p : tableAllocs_) (void)hipFree(p);
  if (stream_) (void)hipStreamDestroy(stream_);
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
                                          const ScfSettings& st, bool wantOrbitals, bool wantEps, PhaseTimes* times) {
  HIPCHK(hipSetDevice(device_));
  counters_ = SolveCounters{};
  const bool sp2 = st.densityBuilder == 0;
  const bool wantSpectrum = wantOrbitals || wantEps;
  const double t0 = nowSeconds();
  ParameterSet& ps = method(methodName);
  const int slots = std::max(2, st.diisSize);
  Layout L = buildLayout(ps, tables_, inputs, slots);
  const int nm = (int)L.mols.size();
  const long long nMat = std::max<long long>(L.nMat, 1);
  const long long nVecN = std::max<long long>(L.nVecN, 1);
  const size_t npairs = L.pairs.size(), natoms = L.atoms.size();

  // ---- tile counts of the per-molecule matrix passes ----
  long long tiles64 = 0;
  for (const MolDev& m : L.mols) tiles64 += (long long)((m.n + 63) / 64) * ((m.n + 63) / 64);
  const int tileX = tiles64 >= 512 ? 64 : 32;  // GEMM tile (results do not depend on it)
  auto tilesOf = [&](const MolDev& m, int kind) -> int {
    if (kind == 0) {  // GEMM tiles
      const int tn = (m.n + tileX - 1) / tileX;
      return tn * tn;
    }
    const int tn = (m.n + 31) / 32;
    return kind == 1 ? tn * tn : tn * (tn + 1) / 2;  // 1: 32 x 32 tiles, 2: on and above the diagonal
  };
  long long totX = 0, tot32 = 0, totS = 0;
  for (const MolDev& m : L.mols) {
    totX += tilesOf(m, 0);
    tot32 += tilesOf(m, 1);
    totS += tilesOf(m, 2);
  }
  const size_t redLen = (size_t)std::max<long long>(1, tot32) * std::max(2, slots);
  // mixed precision: split-BF16 copies of the SP2 iterates, ldp x ldp per molecule and buffer
  const bool mixed = st.precision == 1;
  if (mixed && !sp2) throw std::runtime_error("mixed precision needs the SP2 density builder");
  std::vector<long long> lpOff(nm + 1, 0);
  for (int i = 0; i < nm; ++i) {
    const long long ldp = ((L.mols[i].n + 31) / 32) * 32;
    lpOff[i + 1] = lpOff[i] + (mixed ? ldp * ldp : 0);
  }

  // ---- device memory ----
  size_t need = 0;
  need += sizeOf<ElementDev>(L.elems.size()) + sizeOf<DiatomicDev>(L.diat.size()) + sizeOf<AtomDev>(natoms) +
          sizeOf<PairDev>(npairs) + 2 * sizeOf<MolDev>(nm) + 2 * sizeOf<int>(npairs) + sizeOf<int>(natoms) + sizeOf<int>(nm);
  need += sizeOf<double>(L.nInts) + 2 * sizeOf<double>(L.nVec);
  need += 9 * sizeOf<double>(nMat);  // S H F P Pold Tmp Fx A Z
  need += 5 * sizeOf<double>(nMat);  // eigensolver work
  need += sizeOf<unsigned char>(nMat);
  need += 2 * sizeOf<double>(std::max<long long>(L.nHist, 1));
  need += 10 * sizeOf<double>(nVecN);
  need += 2 * sizeOf<double>(npairs) + sizeOf<double>(natoms) + 5 * sizeOf<double>(nm) + 2 * sizeOf<double>((size_t)nm * slots) +
          sizeOf<double>((size_t)nm * slots * slots) + sizeOf<ScfMolDev>(nm);
  need += 3 * sizeOf<EigJob>(nm) + 3 * sizeOf<GemmJob>(nm) + 2 * sizeOf<Sp2StepJob>(nm) + 6 * sizeOf<int>(nm + 1);
  need += 2 * sizeOf<double>(2 * (size_t)std::max<long long>(totS, 1)) + 2 * sizeOf<Sp2Dev>(nm) + sizeOf<double>(redLen);
  need += 4 * sizeOf<unsigned short>(std::max<long long>(lpOff[nm], 1));
  arena_.reserve(need);
  Carver cv{arena_.base(), 0, arena_.capacity()};
  ElementDev* dElems = cv.take<ElementDev>(L.elems.size());
  DiatomicDev* dDiat = cv.take<DiatomicDev>(L.diat.size());
  AtomDev* dAtoms = cv.take<AtomDev>(natoms);
  PairDev* dPairs = cv.take<PairDev>(npairs);
  MolDev* dMols = cv.take<MolDev>(nm);
  MolDev* dMolsAct = cv.take<MolDev>(nm);  // molecules still iterating (compacted)
  int* dClassPairs = cv.take<int>(npairs);
  int* dClassPairsAct = cv.take<int>(npairs);
  int* dAtomList = cv.take<int>(natoms);
  int* dMolList = cv.take<int>(nm);
  double* dInts = cv.take<double>(L.nInts);
  double* dVuvb = cv.take<double>(L.nVec);
  double* dJvec = cv.take<double>(L.nVec);
  double* dS = cv.take<double>(nMat);
  double* dH = cv.take<double>(nMat);
  double* dF = cv.take<double>(nMat);
  double* dP = cv.take<double>(nMat);
  double* dPold = cv.take<double>(nMat);
  double* dTmp = cv.take<double>(nMat);
  double* dFx = cv.take<double>(nMat);
  double* dA = cv.take<double>(nMat);
  double* dZ = cv.take<double>(nMat);
  double* dWork = cv.take<double>(5 * nMat);
  unsigned char* dPiv = cv.take<unsigned char>(nMat);
  double* dHistF = cv.take<double>(std::max<long long>(L.nHist, 1));
  double* dHistE = cv.take<double>(std::max<long long>(L.nHist, 1));
  double* dEd = cv.take<double>(nVecN);
  double* dEe = cv.take<double>(nVecN);
  double* dEtau = cv.take<double>(nVecN);
  double* dEv = cv.take<double>(2 * nVecN);
  double* dEw = cv.take<double>(2 * nVecN);
  double* dEval = cv.take<double>(nVecN);
  double* dLo = cv.take<double>(nVecN);
  double* dHi = cv.take<double>(nVecN);
  double* dErepPair = cv.take<double>(npairs);
  double* dEPair = cv.take<double>(npairs);
  double* dEAtom = cv.take<double>(natoms);
  double* dEMol = cv.take<double>(nm);
  double* dErep = cv.take<double>(nm);
  double* dRmsd = cv.take<double>(nm);
  double* dNel = cv.take<double>(nm);
  double* dDots = cv.take<double>((size_t)nm * slots);
  double* dCoef = cv.take<double>((size_t)nm * slots);
  double* dBmat = cv.take<double>((size_t)nm * slots * slots);
  ScfMolDev* dScf = cv.take<ScfMolDev>(nm);
  EigJob* dEigOcc = cv.take<EigJob>(nm);
  EigJob* dEigAll = cv.take<EigJob>(nm);
  EigJob* dEigSub = cv.take<EigJob>(nm);
  GemmJob* dGemmFP = cv.take<GemmJob>(nm);
  GemmJob* dGemmSub = cv.take<GemmJob>(nm);
  Sp2StepJob* dStepX0 = cv.take<Sp2StepJob>(nm);
  Sp2StepJob* dStepX1 = cv.take<Sp2StepJob>(nm);
  int* dTilesFP = cv.take<int>(nm + 1);
  int* dTilesSub = cv.take<int>(nm + 1);
  int* dTiles32 = cv.take<int>(nm + 1);
  int* dTilesS = cv.take<int>(nm + 1);
  // SP2 controller state and per-tile partial traces, ping-pong between consecutive steps
  double* dPart[2] = {cv.take<double>(2 * (size_t)std::max<long long>(totS, 1)),
                      cv.take<double>(2 * (size_t)std::max<long long>(totS, 1))};
  Sp2Dev* dSp2[2] = {cv.take<Sp2Dev>(nm), cv.take<Sp2Dev>(nm)};
  double* dRed = cv.take<double>(redLen);
  unsigned short* dLp[4];  // X_hi, X_lo of the even iterates, then of the odd iterates
  for (auto& q : dLp) q = cv.take<unsigned short>(std::max<long long>(lpOff[nm], 1));
  hipStream_t s = stream_;
  auto timed = [&](const char* name, auto&& launch) {
    if (prof_.enabled()) prof_.begin(name, s);
    launch();
    if (prof_.enabled()) prof_.end(s);
  };
  auto up = [&](void* d, const void* h, size_t bytes) {
    if (bytes) HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
  };
  up(dElems, L.elems.data(), L.elems.size() * sizeof(ElementDev));
  up(dDiat, L.diat.data(), L.diat.size() * sizeof(DiatomicDev));
  up(dAtoms, L.atoms.data(), natoms * sizeof(AtomDev));
  up(dPairs, L.pairs.data(), npairs * sizeof(PairDev));
  up(dMols, L.mols.data(), nm * sizeof(MolDev));
  up(dClassPairs, L.classPairs.data(), L.classPairs.size() * sizeof(int));
  up(dNel, L.nel.data(), nm * sizeof(double));
  {
    std::vector<ScfMolDev> init(nm);
    for (auto& x : init) {
      x = ScfMolDev{};
      x.iter = 1;
      x.lowp = mixed ? 1 : 0;
    }
    up(dScf, init.data(), nm * sizeof(ScfMolDev));
  }

  // ---- per-molecule job templates (device pointers into this batch's workspace) ----
  std::vector<EigJob> eigOcc(nm), eigAll(nm);
  std::vector<GemmJob> gP(nm), gFP(nm);
  std::vector<Sp2StepJob> gX0(nm), gX1(nm);
  for (int i = 0; i < nm; ++i) {
    const MolDev& m = L.mols[i];
    EigJob J{};
    J.n = m.n;
    J.nvec = m.nocc;
    J.A = dA + m.matOff;
    J.d = dEd + m.vecOff;
    J.e = dEe + m.vecOff;
    J.tau = dEtau + m.vecOff;
    J.v = dEv + 2 * m.vecOff;
    J.w = dEw + 2 * m.vecOff;
    J.wval = dEval + m.vecOff;
    J.Z = dZ + m.matOff;
    J.work = dWork + 5 * m.matOff;
    J.ipiv = dPiv + m.matOff;
    eigOcc[i] = J;
    J.nvec = m.n;
    eigAll[i] = J;
    // density P = 2 Z_occ Z_occ^T (Z row-major n x nocc)
    GemmJob g{};
    g.M = g.N = m.n;
    g.K = m.nocc;
    g.A = g.B = dZ + m.matOff;
    g.C = dP + m.matOff;
    g.D = g.C;
    g.sAi = g.sBj = m.nocc;
    g.sAk = g.sBk = 1;
    g.sCi = g.sDi = 1;
    g.sCj = g.sDj = m.n;
    g.alpha = 2.0;
    gP[i] = g;
    // DIIS commutator product F P; B(j,k) = P(j,k) = P(k,j) (P symmetric, unit stride along j)
    GemmJob f{};
    f.M = f.N = f.K = m.n;
    f.A = dF + m.matOff;
    f.B = dP + m.matOff;
    f.C = dTmp + m.matOff;
    f.D = f.C;
    f.sAi = f.sBj = 1;
    f.sAk = f.sBk = m.n;
    f.sCi = f.sDi = 1;
    f.sCj = f.sDj = m.n;
    f.alpha = 1.0;
    gFP[i] = f;
    // SP2 steps between the ping-pong buffers A (even iterates) and Z (odd iterates)
    Sp2StepJob x{};
    x.n = m.n;
    x.mol = i;
    x.X = dA + m.matOff;
    x.Y = dZ + m.matOff;
    if (mixed) {
      x.ldp = ((m.n + 31) / 32) * 32;
      x.Xh = dLp[0] + lpOff[i];
      x.Xl = dLp[1] + lpOff[i];
      x.Yh = dLp[2] + lpOff[i];
      x.Yl = dLp[3] + lpOff[i];
    }
    gX0[i] = x;
    x.X = dZ + m.matOff;
    x.Y = dA + m.matOff;
    if (mixed) {
      x.Xh = dLp[2] + lpOff[i];
      x.Xl = dLp[3] + lpOff[i];
      x.Yh = dLp[0] + lpOff[i];
      x.Yl = dLp[1] + lpOff[i];
    }
    gX1[i] = x;
  }
  up(dEigOcc, eigOcc.data(), nm * sizeof(EigJob));
  up(dEigAll, eigAll.data(), nm * sizeof(EigJob));

  KBatch b{};
  b.elems = dElems;
  b.diatomic = dDiat;
  b.nElem = (int)L.elems.size();
  b.methodKind = (int)ps.kind();
  b.atoms = dAtoms;
  b.pairs = dPairs;
  b.mols = dMols;
  b.natoms = (int)natoms;
  b.npairs = (int)npairs;
  b.nmols = nm;
  b.classPairs = dClassPairs;
  for (int c = 0; c < 10; ++c) b.classStart[c] = L.classStart[c];

  // ---- density-independent part ----
  timed("twocenter", [&] { launchTwoCenter(ktab_, b, dInts, dVuvb, s); });
  timed("pair_overlap", [&] { launchPairOverlap(b, dS, dH, dErepPair, s); });
  timed("atom_hcore", [&] { launchAtomHcore(ktab_, b, dVuvb, dH, s); });
  if (b.npairs > 0) launchMolPairSum(b, dErepPair, dErep, s);
  else HIPCHK(hipMemsetAsync(dErep, 0, nm * sizeof(double), s));
  launchDensityGuess(b, dS, dNel, dP, s);
  for (int i = 0; i < nm; ++i)
    if (inputs[i].densityGuess)
      up(dP + L.mols[i].matOff, inputs[i].densityGuess, (size_t)L.mols[i].n * L.mols[i].n * sizeof(double));
  HIPCHK(hipMemcpyAsync(dPold, dP, nMat * sizeof(double), hipMemcpyDeviceToDevice, s));
  std::vector<double> erep(nm);
  HIPCHK(hipMemcpyAsync(erep.data(), dErep, nm * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const double tSetup = nowSeconds();

  // ---- work lists of the lock-step iterations ----
  // Fock list: molecules still iterating plus those that converged in the previous iteration
  // (their Fock build gives the energy at the converged density).  Iterating list: molecules
  // not yet converged; their P and F_extrap are frozen once they converge.
  std::vector<ScfMolDev> scf(nm);
  std::vector<char> inFock(nm, 1), inIter(nm, 1);
  int nIter = nm, nFockAct = nm;
  struct IterLists {
    std::vector<MolDev> mols;
    std::vector<int> tFP, t32, tS;
    int ntFP = 0, nt32 = 0, ntS = 0, maxN = 0, maxNocc = 0;
  } IL;
  auto buildFockLists = [&]() {
    std::vector<int> cp, atomsL, molsL;
    int cs[10];
    for (int c = 0; c < 9; ++c) {
      cs[c] = (int)cp.size();
      for (int k = L.classStart[c]; k < L.classStart[c + 1]; ++k) {
        const int pid = L.classPairs[k];
        if (inFock[L.pairs[pid].mol]) cp.push_back(pid);
      }
    }
    cs[9] = (int)cp.size();
    for (int i = 0; i < nm; ++i)
      if (inFock[i]) {
        molsL.push_back(i);
        for (int a = 0; a < L.mols[i].natoms; ++a) atomsL.push_back(L.mols[i].atom0 + a);
      }
    up(dClassPairsAct, cp.data(), cp.size() * sizeof(int));
    up(dAtomList, atomsL.data(), atomsL.size() * sizeof(int));
    up(dMolList, molsL.data(), molsL.size() * sizeof(int));
    b.classPairs = dClassPairsAct;
    for (int c = 0; c < 10; ++c) b.classStart[c] = cs[c];
    b.atomList = dAtomList;
    b.nAtomList = (int)atomsL.size();
    b.molList = dMolList;
    b.nMolList = (int)molsL.size();
    HIPCHK(hipStreamSynchronize(s));  // host vectors above must outlive the uploads
  };
  auto buildIterLists = [&]() {
    IL = IterLists{};
    std::vector<GemmJob> fp;
    std::vector<Sp2StepJob> x0, x1;
    IL.tFP.push_back(0);
    IL.t32.push_back(0);
    IL.tS.push_back(0);
    for (int i = 0; i < nm; ++i) {
      if (!inIter[i]) continue;
      const MolDev& m = L.mols[i];
      IL.mols.push_back(m);
      fp.push_back(gFP[i]);
      x0.push_back(gX0[i]);
      x1.push_back(gX1[i]);
      IL.tFP.push_back(IL.tFP.back() + tilesOf(m, 0));
      IL.t32.push_back(IL.t32.back() + tilesOf(m, 1));
      IL.tS.push_back(IL.tS.back() + tilesOf(m, 2));
      IL.maxN = std::max(IL.maxN, m.n);
      IL.maxNocc = std::max(IL.maxNocc, m.nocc);
    }
    const int k = (int)IL.mols.size();
    IL.ntFP = IL.tFP.back();
    IL.nt32 = IL.t32.back();
    IL.ntS = IL.tS.back();
    up(dMolsAct, IL.mols.data(), k * sizeof(MolDev));
    up(dGemmFP, fp.data(), k * sizeof(GemmJob));
    up(dStepX0, x0.data(), k * sizeof(Sp2StepJob));
    up(dStepX1, x1.data(), k * sizeof(Sp2StepJob));
    up(dTilesFP, IL.tFP.data(), (k + 1) * sizeof(int));
    up(dTiles32, IL.t32.data(), (k + 1) * sizeof(int));
    up(dTilesS, IL.tS.data(), (k + 1) * sizeof(int));
    HIPCHK(hipStreamSynchronize(s));
  };
  buildFockLists();
  buildIterLists();

  // ---- SCF iterations ----
  double tF = 0, tD = 0, tE = 0, tP = 0;
  int it = 0;
  int lastSp2Steps = 0;
  std::vector<Sp2Dev> sp2State(nm);
  for (it = 1; it <= st.maxIterations && nIter > 0; ++it) {
    const int k = (int)IL.mols.size();
    double ta = nowSeconds();
    timed("fock_pairs", [&] { launchFockPairs(ktab_, b, dInts, dP, dH, dF, dJvec, dEPair, s); });
    timed("fock_atoms", [&] { launchFockAtoms(ktab_, b, dP, dH, dJvec, dF, dEAtom, s); });
    launchEnergyReduce(b, dEAtom, dEPair, dEMol, s);
    double tb = nowSeconds();
    if (st.mixer == 1) {
      launchCopyMats(dMolsAct, k, dF, false, dHistF, true, dScf, slots, s);
      timed("gemm_fp", [&] { launchGemm(dGemmFP, dTilesFP, k, IL.ntFP, tileX, s); });
      timed("diis_dots", [&] { launchCommutatorDots(dMolsAct, dTiles32, k, IL.nt32, dTmp, dHistE, slots, dScf, dRed, dDots, s); });
      launchDiisSolve(dMolsAct, k, dScf, dDots, dBmat, dCoef, slots, s);
      launchDiisExtrapolate(dMolsAct, k, dHistF, slots, dCoef, dFx, s);
    } else {
      launchCopyMats(dMolsAct, k, dF, false, dFx, false, nullptr, slots, s);
    }
    double tc = nowSeconds();
    // density of the (extrapolated) Fock matrix Fx
    if (sp2) {
      timed("sp2_init", [&] {
        launchSp2Init(dMolsAct, k, IL.maxN, dFx, dStepX0, dTilesS, IL.ntS, dLo, dHi, dPart[0], dSp2[0], st.sp2MaxSteps, dScf,
                      s);
      });
      // launches in chunks; a molecule whose X is final leaves the steps (its blocks return at once)
      int steps = 0;
      int chunk = lastSp2Steps > 0 ? std::max(4, lastSp2Steps - 1) : 16;
      for (;;) {
        for (int c = 0; c < chunk && steps < st.sp2MaxSteps; ++c, ++steps) {
          const int e = steps % 2;
          timed("gemm_sp2", [&] {
            launchSp2Step(e == 0 ? dStepX0 : dStepX1, dTilesS, k, IL.ntS, dPart[e], dPart[1 - e], dSp2[e], dSp2[1 - e], s);
          });
        }
        // decision on the newest X, for the host's check only (the next step repeats it)
        const int e = steps % 2;
        launchSp2Decide(e == 0 ? dStepX0 : dStepX1, dTilesS, k, dPart[e], dSp2[e], dSp2[1 - e], s);
        HIPCHK(hipMemcpyAsync(sp2State.data(), dSp2[1 - e], nm * sizeof(Sp2Dev), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        bool all = true;
        for (const MolDev& m : IL.mols) all &= sp2State[m.id].done != 0;
        if (all || steps >= st.sp2MaxSteps) break;
        chunk = 4;
      }
      counters_.sp2Steps += steps;
      int used = 0;
      for (const MolDev& m : IL.mols) {
        const Sp2Dev& q = sp2State[m.id];
        used = std::max(used, q.doneStep);
        counters_.sp2Flops += 2LL * m.n * m.n * m.n * q.doneStep;
        counters_.sp2LowpFlops += 2LL * m.n * m.n * m.n * q.lowpSteps;
      }
      lastSp2Steps = used;
      launchPFromX(dMolsAct, k, dSp2[1 - steps % 2], dA, dZ, dP, s);
      // molecules whose projection did not converge (e.g. a vanishing HOMO-LUMO gap) are
      // diagonalised instead
      std::vector<int> bad;
      for (const MolDev& m : IL.mols)
        if (!sp2State[m.id].converged) bad.push_back(m.id);
      if (!bad.empty()) {
        counters_.sp2Fallbacks += (long long)bad.size();
        const int nb = (int)bad.size();
        std::vector<EigJob> ej(nb);
        std::vector<GemmJob> gj(nb);
        std::vector<int> ts(nb + 1, 0);
        int maxN = 0, maxO = 0;
        for (int q = 0; q < nb; ++q) {
          const int i = bad[q];
          const MolDev& m = L.mols[i];
          ej[q] = eigOcc[i];
          gj[q] = gP[i];
          ts[q + 1] = ts[q] + tilesOf(m, 0);
          maxN = std::max(maxN, m.n);
          maxO = std::max(maxO, m.nocc);
          HIPCHK(hipMemcpyAsync(dA + m.matOff, dFx + m.matOff, (size_t)m.n * m.n * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        up(dEigSub, ej.data(), nb * sizeof(EigJob));
        up(dGemmSub, gj.data(), nb * sizeof(GemmJob));
        up(dTilesSub, ts.data(), (nb + 1) * sizeof(int));
        launchSytrd(dEigSub, nb, s);
        launchStebz(dEigSub, nb, maxN, s);
        launchStein(dEigSub, nb, maxO, s);
        launchOrmtr(dEigSub, nb, maxO, s);
        launchGemm(dGemmSub, dTilesSub, nb, ts[nb], tileX, s);
        ++counters_.eigensolves;
        HIPCHK(hipStreamSynchronize(s));  // the host job vectors above must outlive the uploads
      }
    } else {
      std::vector<EigJob> ej;
      std::vector<GemmJob> gj;
      std::vector<int> ts(1, 0);
      for (const MolDev& m : IL.mols) {
        ej.push_back(eigOcc[m.id]);
        gj.push_back(gP[m.id]);
        ts.push_back(ts.back() + tilesOf(m, 0));
      }
      up(dEigSub, ej.data(), k * sizeof(EigJob));
      up(dGemmSub, gj.data(), k * sizeof(GemmJob));
      up(dTilesSub, ts.data(), (k + 1) * sizeof(int));
      launchCopyMats(dMolsAct, k, dFx, false, dA, false, nullptr, slots, s);
      timed("sytrd", [&] { launchSytrd(dEigSub, k, s); });
      timed("stebz", [&] { launchStebz(dEigSub, k, IL.maxN, s); });
      timed("stein", [&] { launchStein(dEigSub, k, IL.maxNocc, s); });
      timed("ormtr", [&] { launchOrmtr(dEigSub, k, IL.maxNocc, s); });
      timed("gemm_p", [&] { launchGemm(dGemmSub, dTilesSub, k, ts.back(), tileX, s); });
      ++counters_.eigensolves;
      HIPCHK(hipStreamSynchronize(s));
    }
    double td = nowSeconds();
    timed("rmsd", [&] { launchRmsdUpdate(dMolsAct, dTiles32, k, IL.nt32, dP, dPold, dRed, dRmsd, s); });
    launchScfCheck(dMolList, b.nMolList, dScf, dEMol, dRmsd, st.energyTol, st.rmsdTol, st.lowpTol, s);
    HIPCHK(hipMemcpyAsync(scf.data(), dScf, nm * sizeof(ScfMolDev), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double te = nowSeconds();
    tF += tb - ta;
    tD += tc - tb;
    tE += td - tc;
    tP += te - td;
    bool iterChanged = false, fockChanged = false;
    for (int i = 0; i < nm; ++i) {
      if (inIter[i] && scf[i].converged) {
        inIter[i] = 0;
        --nIter;
        iterChanged = true;
      }
      if (inFock[i] && scf[i].haveEnergy) {
        inFock[i] = 0;
        --nFockAct;
        fockChanged = true;
      }
    }
    if (nIter > 0 && fockChanged) buildFockLists();
    if (nIter > 0 && iterChanged) buildIterLists();
  }
  // final Fock build from the current densities for the molecules still without an energy
  for (int i = 0; i < nm; ++i) inFock[i] = !scf[i].haveEnergy;
  buildFockLists();
  launchFockPairs(ktab_, b, dInts, dP, dH, dF, dJvec, dEPair, s);
  launchFockAtoms(ktab_, b, dP, dH, dJvec, dF, dEAtom, s);
  launchEnergyReduce(b, dEAtom, dEPair, dEMol, s);
  std::vector<double> eMol(nm);
  HIPCHK(hipMemcpyAsync(eMol.data(), dEMol, nm * sizeof(double), hipMemcpyDeviceToHost, s));
  // results: P (frozen at convergence, or the last one), energies
  std::vector<MoleculeResult> res(nm);
  for (int i = 0; i < nm; ++i) {
    const MolDev& m = L.mols[i];
    res[i].nao = m.n;
    res[i].nocc = m.nocc;
    res[i].P.resize((size_t)m.n * m.n);
    HIPCHK(hipMemcpyAsync(res[i].P.data(), dP + m.matOff, res[i].P.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  for (int i = 0; i < nm; ++i) {
    const ScfMolDev& S = scf[i];
    res[i].electronicEnergy = S.haveEnergy ? S.Eel : eMol[i];
    res[i].repulsionEnergy = erep[i];
    res[i].energy = res[i].electronicEnergy + erep[i];
    res[i].converged = S.converged != 0;
    res[i].iterations = S.converged ? S.convIter : std::min(it - 1, st.maxIterations);
  }
  // orbital energies (and all eigenvectors) of each molecule's last diagonalised / projected
  // matrix, F_extrap, which stays in place once a molecule has converged
  if (wantSpectrum) {
    launchCopyMats(dMols, nm, dFx, false, dA, false, nullptr, slots, s);
    launchSytrd(dEigAll, nm, s);
    launchStebz(dEigAll, nm, L.maxN, s);
    if (wantOrbitals) {
      launchStein(dEigAll, nm, L.maxN, s);
      launchOrmtr(dEigAll, nm, L.maxN, s);
    }
    ++counters_.eigensolves;
    std::vector<double> Zr;
    for (int i = 0; i < nm; ++i) {
      const MolDev& m = L.mols[i];
      res[i].eps.resize(m.n);
      HIPCHK(hipMemcpyAsync(res[i].eps.data(), dEval + m.vecOff, m.n * sizeof(double), hipMemcpyDeviceToHost, s));
      if (wantOrbitals) {
        Zr.resize((size_t)m.n * m.n);
        res[i].C.resize((size_t)m.n * m.n);
        HIPCHK(hipMemcpyAsync(Zr.data(), dZ + m.matOff, Zr.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      }
      HIPCHK(hipStreamSynchronize(s));
      // Z is row-major n x n (row = AO, column = MO) == column-major C with C[k*n + r]
      if (wantOrbitals)
        for (int r = 0; r < m.n; ++r)
          for (int q = 0; q < m.n; ++q) res[i].C[(size_t)q * m.n + r] = Zr[(size_t)r * m.n + q];
    }
  }
  counters_.iterations = it - 1;
  HIPCHK(hipStreamSynchronize(s));
  if (prof_.enabled()) prof_.collect();
  if (times) {
    times->setup = tSetup - t0;
    times->fock = tF;
    times->diis = tD;
    times->eig = tE;
    times->density = tP;
    times->total = nowSeconds() - t0;
    times->iterations = it - 1;
  }
  return res;
}

void Engine::densityIndependent(const std::string& methodName, const MoleculeInput& mol, double* S, double* H, double* Erep) {
  HIPCHK(hipSetDevice(device_));
  ParameterSet& ps = method(methodName);
  Layout L = buildLayout(ps, tables_, {mol}, 2);
  const long long nMat = std::max<long long>(L.nMat, 1);
  size_t need = sizeOf<ElementDev>(L.elems.size()) + sizeOf<DiatomicDev>(L.diat.size()) + sizeOf<AtomDev>(L.atoms.size()) +
                sizeOf<PairDev>(L.pairs.size()) + sizeOf<MolDev>(1) + sizeOf<int>(L.classPairs.size()) +
                sizeOf<double>(L.nInts) + sizeOf<double>(L.nVec) + 2 * sizeOf<double>(nMat) + sizeOf<double>(L.pairs.size()) +
                sizeOf<double>(1);
  arena_.reserve(need);
  Carver cv{arena_.base(), 0, arena_.capacity()};
  ElementDev* dElems = cv.take<ElementDev>(L.elems.size());
  DiatomicDev* dDiat = cv.take<DiatomicDev>(L.diat.size());
  AtomDev* dAtoms = cv.take<AtomDev>(L.atoms.size());
  PairDev* dPairs = cv.take<PairDev>(L.pairs.size());
  MolDev* dMols = cv.take<MolDev>(1);
  int* dClassPairs = cv.take<int>(L.classPairs.size());
  double* dInts = cv.take<double>(L.nInts);
  double* dVuvb = cv.take<double>(L.nVec);
  double* dS = cv.take<double>(nMat);
  double* dH = cv.take<double>(nMat);
  double* dErepPair = cv.take<double>(L.pairs.size());
  double* dErep = cv.take<double>(1);
  hipStream_t s = stream_;
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
  b.atoms = dAtoms;
  b.pairs = dPairs;
  b.mols = dMols;
  b.natoms = (int)L.atoms.size();
  b.npairs = (int)L.pairs.size();
  b.nmols = 1;
  b.classPairs = dClassPairs;
  for (int c = 0; c < 10; ++c) b.classStart[c] = L.classStart[c];
  launchTwoCenter(ktab_, b, dInts, dVuvb, s);
  launchPairOverlap(b, dS, dH, dErepPair, s);
  launchAtomHcore(ktab_, b, dVuvb, dH, s);
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

void Engine::fock(const std::string& methodName, const MoleculeInput& mol, const double* P, double* F, double* Eel) {
  HIPCHK(hipSetDevice(device_));
  ParameterSet& ps = method(methodName);
  Layout L = buildLayout(ps, tables_, {mol}, 2);
  const long long nMat = std::max<long long>(L.nMat, 1);
  size_t need = sizeOf<ElementDev>(L.elems.size()) + sizeOf<DiatomicDev>(L.diat.size()) + sizeOf<AtomDev>(L.atoms.size()) +
                sizeOf<PairDev>(L.pairs.size()) + sizeOf<MolDev>(1) + sizeOf<int>(L.classPairs.size()) +
                sizeOf<double>(L.nInts) + 2 * sizeOf<double>(L.nVec) + 4 * sizeOf<double>(nMat) +
                2 * sizeOf<double>(L.pairs.size()) + sizeOf<double>(L.atoms.size()) + sizeOf<double>(1) +
                sizeOf<int>(L.atoms.size()) + sizeOf<int>(1);
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
  double* dJvec = cv.take<double>(L.nVec);
  double* dS = cv.take<double>(nMat);
  double* dH = cv.take<double>(nMat);
  double* dP = cv.take<double>(nMat);
  double* dF = cv.take<double>(nMat);
  double* dErepPair = cv.take<double>(L.pairs.size());
  double* dEPair = cv.take<double>(L.pairs.size());
  double* dEAtom = cv.take<double>(L.atoms.size());
  double* dEMol = cv.take<double>(1);
  hipStream_t s = stream_;
  const size_t nn = (size_t)L.mols[0].n * L.mols[0].n;
  HIPCHK(hipMemcpyAsync(dElems, L.elems.data(), L.elems.size() * sizeof(ElementDev), hipMemcpyHostToDevice, s));
  if (!L.diat.empty()) HIPCHK(hipMemcpyAsync(dDiat, L.diat.data(), L.diat.size() * sizeof(DiatomicDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dAtoms, L.atoms.data(), L.atoms.size() * sizeof(AtomDev), hipMemcpyHostToDevice, s));
  if (!L.pairs.empty()) HIPCHK(hipMemcpyAsync(dPairs, L.pairs.data(), L.pairs.size() * sizeof(PairDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dMols, L.mols.data(), sizeof(MolDev), hipMemcpyHostToDevice, s));
  if (!L.classPairs.empty())
    HIPCHK(hipMemcpyAsync(dClassPairs, L.classPairs.data(), L.classPairs.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dP, P, nn * sizeof(double), hipMemcpyHostToDevice, s));
  std::vector<int> atomList(L.atoms.size());
  for (size_t a = 0; a < atomList.size(); ++a) atomList[a] = (int)a;
  const int molList = 0;
  HIPCHK(hipMemcpyAsync(dAtomList, atomList.data(), atomList.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dMolList, &molList, sizeof(int), hipMemcpyHostToDevice, s));
  KBatch b{};
  b.atomList = dAtomList;
  b.nAtomList = (int)atomList.size();
  b.molList = dMolList;
  b.nMolList = 1;
  b.elems = dElems;
  b.diatomic = dDiat;
  b.nElem = (int)L.elems.size();
  b.methodKind = (int)ps.kind();
  b.atoms = dAtoms;
  b.pairs = dPairs;
  b.mols = dMols;
  b.natoms = (int)L.atoms.size();
  b.npairs = (int)L.pairs.size();
  b.nmols = 1;
  b.classPairs = dClassPairs;
  for (int c = 0; c < 10; ++c) b.classStart[c] = L.classStart[c];
  launchTwoCenter(ktab_, b, dInts, dVuvb, s);
  launchPairOverlap(b, dS, dH, dErepPair, s);
  launchAtomHcore(ktab_, b, dVuvb, dH, s);
  launchFockPairs(ktab_, b, dInts, dP, dH, dF, dJvec, dEPair, s);
  launchFockAtoms(ktab_, b, dP, dH, dJvec, dF, dEAtom, s);
  launchEnergyReduce(b, dEAtom, dEPair, dEMol, s);
  double e = 0;
  HIPCHK(hipMemcpyAsync(&e, dEMol, sizeof(double), hipMemcpyDeviceToHost, s));
  if (F) HIPCHK(hipMemcpyAsync(F, dF, nn * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (Eel) *Eel = e;
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
    L.pairs = {p};
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

void Engine::eigh(int n, int nvec, const double* A, double* w, double* V) {
  HIPCHK(hipSetDevice(device_));
  if (n <= 0) return;
  if (nvec < 0 || nvec > n) throw std::runtime_error("eigh: nvec out of range");
  const long long nn = (long long)n * n;
  arena_.reserve(2 * sizeOf<double>(nn) + 5 * sizeOf<double>(nn) + sizeOf<unsigned char>(nn) + 8 * sizeOf<double>(2 * n) +
                 sizeOf<EigJob>(1));
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
  hipStream_t s = stream_;
  HIPCHK(hipMemcpyAsync(J.A, A, nn * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dJ, &J, sizeof(EigJob), hipMemcpyHostToDevice, s));
  launchSytrd(dJ, 1, s);
  launchStebz(dJ, 1, n, s);
  launchStein(dJ, 1, nvec, s);
  launchOrmtr(dJ, 1, nvec, s);
  std::vector<double> Z((size_t)n * std::max(nvec, 1));
  HIPCHK(hipMemcpyAsync(w, J.wval, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (nvec > 0) HIPCHK(hipMemcpyAsync(Z.data(), J.Z, (size_t)n * nvec * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
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

}  // namespace sparrow_amd
