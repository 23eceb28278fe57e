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

// Small dense solve with partial pivoting; returns false if singular.
bool solveDense(int m, std::vector<double> A, std::vector<double> b, std::vector<double>& x) {
  for (int k = 0; k < m; ++k) {
    int p = k;
    for (int r = k + 1; r < m; ++r)
      if (std::abs(A[r * m + k]) > std::abs(A[p * m + k])) p = r;
    if (A[p * m + k] == 0.0) return false;
    if (p != k) {
      for (int c = 0; c < m; ++c) std::swap(A[k * m + c], A[p * m + c]);
      std::swap(b[k], b[p]);
    }
    for (int r = k + 1; r < m; ++r) {
      double f = A[r * m + k] / A[k * m + k];
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
  for (double v : x)
    if (!std::isfinite(v)) return false;
  return true;
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
  // ---- device memory ----
  size_t need = 0;
  need += sizeOf<ElementDev>(L.elems.size()) + sizeOf<DiatomicDev>(L.diat.size()) + sizeOf<AtomDev>(L.atoms.size()) +
          sizeOf<PairDev>(L.pairs.size()) + sizeOf<MolDev>(nm) + sizeOf<int>(L.classPairs.size());
  need += sizeOf<double>(L.nInts) + 2 * sizeOf<double>(L.nVec);
  need += 9 * sizeOf<double>(nMat);                       // S H F P Pold Tmp Fx A Z
  need += 5 * sizeOf<double>(nMat);                       // eigensolver work
  need += sizeOf<unsigned char>(nMat);                    // pivots
  need += 2 * sizeOf<double>(std::max<long long>(L.nHist, 1));
  need += 10 * sizeOf<double>(std::max<long long>(L.nVecN, 1));
  need += 2 * sizeOf<double>(L.pairs.size()) + sizeOf<double>(L.atoms.size()) + 5 * sizeOf<double>(nm) +
          2 * sizeOf<double>((size_t)nm * slots);
  need += 3 * sizeOf<EigJob>(nm) + 5 * sizeOf<GemmJob>(nm) + 4 * sizeOf<int>(nm + 1);
  // SP2: GEMM tile size by the amount of work, per-tile trace partials, controller state
  long long tiles64 = 0;
  for (const MolDev& m : L.mols) tiles64 += (long long)((m.n + 63) / 64) * ((m.n + 63) / 64);
  const int tileX = tiles64 >= 512 ? 64 : 32;
  std::vector<int> tilesX(nm + 1, 0);
  for (int i = 0; i < nm; ++i) {
    const int tn = (L.mols[i].n + tileX - 1) / tileX;
    tilesX[i + 1] = tilesX[i] + tn * tn;
  }
  // SP2 steps: 32 x 32 tiles on and above the diagonal
  std::vector<int> tilesS(nm + 1, 0);
  for (int i = 0; i < nm; ++i) {
    const int tn = (L.mols[i].n + 31) / 32;
    tilesS[i + 1] = tilesS[i] + tn * (tn + 1) / 2;
  }
  need += 2 * sizeOf<double>(2 * (size_t)std::max(tilesS[nm], 1)) + 2 * sizeOf<Sp2Dev>(nm) + sizeOf<int>(nm + 1);
  // 32 x 32 tiles of the element-wise reductions (commutator/DIIS dots, RMSD, SP2 start)
  std::vector<int> tiles32(nm + 1, 0);
  for (int i = 0; i < nm; ++i) {
    const int tn = (L.mols[i].n + 31) / 32;
    tiles32[i + 1] = tiles32[i] + tn * tn;
  }
  const int ntiles32 = tiles32[nm];
  const size_t redLen = (size_t)std::max(1, ntiles32) * std::max(2, slots);
  need += sizeOf<double>(redLen) + sizeOf<int>(nm + 1);
  arena_.reserve(need);
  Carver cv{arena_.base(), 0, arena_.capacity()};
  ElementDev* dElems = cv.take<ElementDev>(L.elems.size());
  DiatomicDev* dDiat = cv.take<DiatomicDev>(L.diat.size());
  AtomDev* dAtoms = cv.take<AtomDev>(L.atoms.size());
  PairDev* dPairs = cv.take<PairDev>(L.pairs.size());
  MolDev* dMols = cv.take<MolDev>(nm);
  int* dClassPairs = cv.take<int>(L.classPairs.size());
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
  double* dEd = cv.take<double>(L.nVecN);
  double* dEe = cv.take<double>(L.nVecN);
  double* dEtau = cv.take<double>(L.nVecN);
  double* dEv = cv.take<double>(2 * L.nVecN);
  double* dEw = cv.take<double>(2 * L.nVecN);
  double* dEval = cv.take<double>(L.nVecN);
  double* dErepPair = cv.take<double>(L.pairs.size());
  double* dEPair = cv.take<double>(L.pairs.size());
  double* dEAtom = cv.take<double>(L.atoms.size());
  double* dEMol = cv.take<double>(nm);
  double* dErep = cv.take<double>(nm);
  double* dRmsd = cv.take<double>(nm);
  double* dNel = cv.take<double>(nm);
  double* dDots = cv.take<double>((size_t)nm * slots);
  double* dCoef = cv.take<double>((size_t)nm * slots);
  EigJob* dEigOcc = cv.take<EigJob>(nm);
  EigJob* dEigAll = cv.take<EigJob>(nm);
  GemmJob* dGemmP = cv.take<GemmJob>(nm);
  GemmJob* dGemmFP = cv.take<GemmJob>(nm);
  int* dTilesP = cv.take<int>(nm + 1);
  int* dTilesFP = cv.take<int>(nm + 1);
  EigJob* dEigSub = cv.take<EigJob>(nm);
  GemmJob* dGemmSub = cv.take<GemmJob>(nm);
  Sp2StepJob* dStepX0 = cv.take<Sp2StepJob>(nm);
  Sp2StepJob* dStepX1 = cv.take<Sp2StepJob>(nm);
  int* dTilesS = cv.take<int>(nm + 1);
  int* dTilesSub = cv.take<int>(nm + 1);
  int* dTilesX = cv.take<int>(nm + 1);
  // SP2 controller state and per-tile partial traces, ping-pong between consecutive steps
  double* dPart[2] = {cv.take<double>(2 * (size_t)std::max(tilesS[nm], 1)), cv.take<double>(2 * (size_t)std::max(tilesS[nm], 1))};
  Sp2Dev* dSp2[2] = {cv.take<Sp2Dev>(nm), cv.take<Sp2Dev>(nm)};
  double* dRed = cv.take<double>(redLen);
  int* dTiles32 = cv.take<int>(nm + 1);
  double* dLo = cv.take<double>(L.nVecN);
  double* dHi = cv.take<double>(L.nVecN);
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
  up(dAtoms, L.atoms.data(), L.atoms.size() * sizeof(AtomDev));
  up(dPairs, L.pairs.data(), L.pairs.size() * sizeof(PairDev));
  up(dMols, L.mols.data(), L.mols.size() * sizeof(MolDev));
  up(dClassPairs, L.classPairs.data(), L.classPairs.size() * sizeof(int));
  up(dNel, L.nel.data(), nm * sizeof(double));
  // eigen / gemm job lists
  std::vector<EigJob> eigOcc(nm), eigAll(nm);
  std::vector<GemmJob> gP(nm), gFP(nm);
  std::vector<Sp2StepJob> gX0(nm), gX1(nm);
  std::vector<int> tilesP(nm + 1, 0), tilesFP(nm + 1, 0);
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
    GemmJob g{};
    g.M = m.n;
    g.N = m.n;
    g.K = m.nocc;
    g.A = dZ + m.matOff;
    g.B = dZ + m.matOff;
    g.C = dP + m.matOff;
    g.sAi = m.nocc;
    g.sAk = 1;
    g.sBj = m.nocc;
    g.sBk = 1;
    g.sCi = 1;
    g.sCj = m.n;
    g.D = g.C;
    g.sDi = g.sCi;
    g.sDj = g.sCj;
    g.alpha = 2.0;
    g.beta = 0.0;
    gP[i] = g;
    GemmJob f{};
    f.M = m.n;
    f.N = m.n;
    f.K = m.n;
    f.A = dF + m.matOff;
    f.B = dP + m.matOff;  // B(j,k) = P(j,k) = P(k,j): P is symmetric; unit stride along j
    f.C = dTmp + m.matOff;
    f.sAi = 1;
    f.sAk = m.n;
    f.sBj = 1;
    f.sBk = m.n;
    f.sCi = 1;
    f.sCj = m.n;
    f.D = f.C;
    f.sDi = f.sCi;
    f.sDj = f.sCj;
    f.alpha = 1.0;
    f.beta = 0.0;
    gFP[i] = f;
    // SP2 steps X' = alpha X X + beta X between the ping-pong buffers A and Z, with alpha and
    // beta chosen on the device by the controller
    Sp2StepJob x{};
    x.n = m.n;
    x.X = dA + m.matOff;
    x.Y = dZ + m.matOff;
    gX0[i] = x;
    x.X = dZ + m.matOff;
    x.Y = dA + m.matOff;
    gX1[i] = x;
    tilesP[i + 1] = tilesX[i + 1];
    tilesFP[i + 1] = tilesX[i + 1];
  }
  up(dEigOcc, eigOcc.data(), nm * sizeof(EigJob));
  up(dEigAll, eigAll.data(), nm * sizeof(EigJob));
  up(dGemmP, gP.data(), nm * sizeof(GemmJob));
  up(dGemmFP, gFP.data(), nm * sizeof(GemmJob));
  up(dTilesP, tilesP.data(), (nm + 1) * sizeof(int));
  up(dTilesFP, tilesFP.data(), (nm + 1) * sizeof(int));
  up(dStepX0, gX0.data(), nm * sizeof(Sp2StepJob));
  up(dStepX1, gX1.data(), nm * sizeof(Sp2StepJob));
  up(dTilesS, tilesS.data(), (nm + 1) * sizeof(int));
  up(dTilesX, tilesX.data(), (nm + 1) * sizeof(int));
  up(dTiles32, tiles32.data(), (nm + 1) * sizeof(int));

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

  // ---- SCF iterations ----
  struct MolState {
    std::vector<double> B;  // slots x slots error overlaps
    std::vector<int> stored;
    double Eold = 0;
    bool converged = false, haveEnergy = false;
    int convIter = 0;
    double Eel = 0;
  };
  std::vector<MolState> ms(nm);
  for (auto& m : ms) m.B.assign((size_t)slots * slots, 0.0);
  std::vector<MoleculeResult> res(nm);
  std::vector<std::vector<double>> finalF(nm);
  std::vector<double> eMol(nm), dots((size_t)nm * slots), coef((size_t)nm * slots), rmsd(nm);
  const int ntP = tilesP[nm], ntFP = tilesFP[nm];
  double tF = 0, tD = 0, tE = 0, tP = 0;
  int it = 0;
  auto snapshot = [&](int i) {
    const MolDev& m = L.mols[i];
    MoleculeResult& r = res[i];
    r.nao = m.n;
    r.nocc = m.nocc;
    r.P.resize((size_t)m.n * m.n);
    HIPCHK(hipMemcpyAsync(r.P.data(), dP + m.matOff, r.P.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    if (wantSpectrum) {
      finalF[i].resize((size_t)m.n * m.n);  // the matrix the final density was built from
      HIPCHK(hipMemcpyAsync(finalF[i].data(), dFx + m.matOff, finalF[i].size() * sizeof(double), hipMemcpyDeviceToHost, s));
    }
  };
  int nconv = 0;
  int lastSp2Steps = 0;
  std::vector<Sp2Dev> sp2State(nm);
  for (it = 1; it <= st.maxIterations && nconv < nm; ++it) {
    double ta = nowSeconds();
    timed("fock_pairs", [&] { launchFockPairs(ktab_, b, dInts, dP, dH, dF, dJvec, dEPair, s); });
    timed("fock_atoms", [&] { launchFockAtoms(ktab_, b, dP, dH, dJvec, dF, dEAtom, s); });
    launchEnergyReduce(b, dEAtom, dEPair, dEMol, s);
    double tb = nowSeconds();
    const int slot = (it - 1) % slots;
    if (st.mixer == 1) {
      launchCopyMats(dMols, nm, dF, false, dHistF, true, slot, s);
      timed("gemm_fp", [&] { launchGemm(dGemmFP, dTilesFP, nm, ntFP, tileX, s); });
      timed("diis_dots", [&] { launchCommutatorDots(dMols, dTiles32, nm, ntiles32, dTmp, dHistE, slots, slot, dRed, dDots, s); });
      HIPCHK(hipMemcpyAsync(dots.data(), dDots, dots.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHK(hipMemcpyAsync(eMol.data(), dEMol, nm * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      for (int i = 0; i < nm; ++i) {
        MolState& S = ms[i];
        auto pos = std::find(S.stored.begin(), S.stored.end(), slot);
        if (pos != S.stored.end()) S.stored.erase(pos);
        S.stored.push_back(slot);
        for (int q : S.stored) {
          S.B[(size_t)slot * slots + q] = dots[(size_t)i * slots + q];
          S.B[(size_t)q * slots + slot] = dots[(size_t)i * slots + q];
        }
        std::fill(coef.begin() + (size_t)i * slots, coef.begin() + (size_t)(i + 1) * slots, 0.0);
        const int mm = (int)S.stored.size();
        bool ok = false;
        if (mm >= 2) {
          const int dim = mm + 1;
          std::vector<double> A((size_t)dim * dim, 0.0), rhs(dim, 0.0), x;
          double scale = 0;
          for (int a = 0; a < mm; ++a) scale = std::max(scale, std::abs(S.B[(size_t)S.stored[a] * slots + S.stored[a]]));
          if (scale <= 0) scale = 1;
          for (int a = 0; a < mm; ++a)
            for (int c = 0; c < mm; ++c) A[(size_t)a * dim + c] = S.B[(size_t)S.stored[a] * slots + S.stored[c]] / scale;
          for (int a = 0; a < mm; ++a) A[(size_t)a * dim + mm] = A[(size_t)mm * dim + a] = -1.0;
          rhs[mm] = -1.0;
          ok = solveDense(dim, A, rhs, x);
          if (ok)
            for (int a = 0; a < mm; ++a) coef[(size_t)i * slots + S.stored[a]] = x[a];
        }
        if (!ok) coef[(size_t)i * slots + slot] = 1.0;
      }
      HIPCHK(hipMemcpyAsync(dCoef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, s));
      launchDiisExtrapolate(dMols, nm, dHistF, slots, dCoef, dFx, s);
    } else {
      launchCopyMats(dMols, nm, dF, false, dFx, false, 0, s);
      HIPCHK(hipMemcpyAsync(eMol.data(), dEMol, nm * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    double tc = nowSeconds();
    // density of the (extrapolated) Fock matrix Fx
    if (sp2) {
      // X0 from Fx into A, then GEMM steps between A and Z; every molecule takes the same number
      // of launches (finished ones copy), so the result of all of them is in one buffer
      timed("sp2_init", [&] {
        launchSp2Init(dMols, nm, L.nVecN, dFx, dStepX0, dTilesS, tilesS[nm], dLo, dHi, dPart[0], dSp2[0], st.sp2MaxSteps, s);
      });
      int steps = 0;
      int chunk = lastSp2Steps > 0 ? std::max(4, lastSp2Steps - 1) : 16;
      for (;;) {
        for (int c = 0; c < chunk && steps < st.sp2MaxSteps; ++c, ++steps) {
          timed("gemm_sp2", [&] {
            const int e = steps % 2;
            launchSp2Step(e == 0 ? dStepX0 : dStepX1, dTilesS, nm, tilesS[nm], dPart[e], dPart[1 - e], dSp2[e], dSp2[1 - e], s);
          });
        }
        // decision on the newest X, for the convergence check only (the next step repeats it)
        const int e = steps % 2;
        launchSp2Decide(dTilesS, nm, dPart[e], dSp2[e], dSp2[1 - e], s);
        HIPCHK(hipMemcpyAsync(sp2State.data(), dSp2[1 - e], nm * sizeof(Sp2Dev), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        bool all = true;
        for (int i = 0; i < nm; ++i) all &= sp2State[i].done != 0;
        if (all || steps >= st.sp2MaxSteps) break;
        chunk = 4;
      }
      counters_.sp2Steps += steps;
      int used = 0;
      for (int i = 0; i < nm; ++i) {
        if (!ms[i].converged) used = std::max(used, sp2State[i].doneStep);
        counters_.sp2Flops += 2LL * L.mols[i].n * L.mols[i].n * L.mols[i].n * sp2State[i].doneStep;
      }
      lastSp2Steps = used;
      launchScaleMats(dMols, nm, steps % 2 == 0 ? dA : dZ, 2.0, dP, s);
      // molecules whose projection did not converge (e.g. a vanishing HOMO-LUMO gap):
      // diagonalise their Fx instead
      std::vector<int> bad;
      for (int i = 0; i < nm; ++i)
        if (!sp2State[i].converged) bad.push_back(i);
      if (!bad.empty()) {
        counters_.sp2Fallbacks += (long long)bad.size();
        const int nb = (int)bad.size();
        std::vector<EigJob> ej(nb);
        std::vector<GemmJob> gj(nb);
        std::vector<int> ts(nb + 1, 0);
        int maxN = 0, maxO = 0;
        for (int k = 0; k < nb; ++k) {
          const int i = bad[k];
          const MolDev& m = L.mols[i];
          ej[k] = eigOcc[i];
          gj[k] = gP[i];
          ts[k + 1] = ts[k] + (tilesP[i + 1] - tilesP[i]);
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
      launchCopyMats(dMols, nm, dFx, false, dA, false, 0, s);
      timed("sytrd", [&] { launchSytrd(dEigOcc, nm, s); });
      timed("stebz", [&] { launchStebz(dEigOcc, nm, L.maxN, s); });
      timed("stein", [&] { launchStein(dEigOcc, nm, L.maxNocc, s); });
      timed("ormtr", [&] { launchOrmtr(dEigOcc, nm, L.maxNocc, s); });
      ++counters_.eigensolves;
    }
    HIPCHK(hipStreamSynchronize(s));
    double td = nowSeconds();
    if (!sp2) timed("gemm_p", [&] { launchGemm(dGemmP, dTilesP, nm, ntP, tileX, s); });
    timed("rmsd", [&] { launchRmsdUpdate(dMols, dTiles32, nm, ntiles32, dP, dPold, dRed, dRmsd, s); });
    HIPCHK(hipMemcpyAsync(rmsd.data(), dRmsd, nm * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double te = nowSeconds();
    tF += tb - ta;
    tD += tc - tb;
    tE += td - tc;
    tP += te - td;
    for (int i = 0; i < nm; ++i) {
      MolState& S = ms[i];
      if (S.converged) {
        if (!S.haveEnergy) {
          S.Eel = eMol[i];
          S.haveEnergy = true;
        }
        continue;
      }
      const double E = eMol[i];
      if (it > 1 && std::abs(E - S.Eold) < st.energyTol && rmsd[i] < st.rmsdTol) {
        S.converged = true;
        S.convIter = it;
        ++nconv;
        snapshot(i);
      }
      S.Eold = E;
    }
  }
  // final Fock build from the current densities for the molecules still without an energy
  launchFockPairs(ktab_, b, dInts, dP, dH, dF, dJvec, dEPair, s);
  launchFockAtoms(ktab_, b, dP, dH, dJvec, dF, dEAtom, s);
  launchEnergyReduce(b, dEAtom, dEPair, dEMol, s);
  HIPCHK(hipMemcpyAsync(eMol.data(), dEMol, nm * sizeof(double), hipMemcpyDeviceToHost, s));
  for (int i = 0; i < nm; ++i)
    if (!ms[i].converged) snapshot(i);
  HIPCHK(hipStreamSynchronize(s));
  for (int i = 0; i < nm; ++i) {
    MolState& S = ms[i];
    if (!S.haveEnergy) S.Eel = eMol[i];
    res[i].electronicEnergy = S.Eel;
    res[i].repulsionEnergy = erep[i];
    res[i].energy = S.Eel + erep[i];
    res[i].converged = S.converged;
    res[i].iterations = S.converged ? S.convIter : std::min(it - 1, st.maxIterations);
  }
  // orbital energies (and all eigenvectors) of each molecule's last diagonalised / projected matrix
  if (wantSpectrum) {
    for (int i = 0; i < nm; ++i) up(dA + L.mols[i].matOff, finalF[i].data(), finalF[i].size() * sizeof(double));
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
          for (int k = 0; k < m.n; ++k) res[i].C[(size_t)k * m.n + r] = Zr[(size_t)r * m.n + k];
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
                2 * sizeOf<double>(L.pairs.size()) + sizeOf<double>(L.atoms.size()) + sizeOf<double>(1);
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
