// Device-resident lock-step SCF over a changing set of molecules (see workspace.hpp).
#include "workspace.hpp"

#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "util.hpp"

namespace sparrow_amd {

// ---------------------------------------------------------------------------------------------
// RangeAlloc
// ---------------------------------------------------------------------------------------------
void RangeAlloc::reset(long long capacity) {
  free_.clear();
  cap_ = capacity;
  used_ = 0;
  if (capacity > 0) free_[0] = capacity;
}

bool RangeAlloc::alloc(long long len, long long* off) {
  if (len <= 0) {
    *off = 0;
    return true;
  }
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < len) continue;
    *off = it->first;
    const long long rest = it->second - len, start = it->first + len;
    free_.erase(it);
    if (rest > 0) free_[start] = rest;
    used_ += len;
    return true;
  }
  return false;
}

void RangeAlloc::release(long long off, long long len) {
  if (len <= 0) return;
  used_ -= len;
  auto next = free_.lower_bound(off);
  if (next != free_.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {  // merge with the range before
      off = prev->first;
      len += prev->second;
      free_.erase(prev);
    }
  }
  next = free_.lower_bound(off);
  if (next != free_.end() && off + len == next->first) {  // and with the one after
    len += next->second;
    free_.erase(next);
  }
  free_[off] = len;
}

namespace {

int gemmTiles(int n, int tile) {
  const int tn = (n + tile - 1) / tile;
  return tn * tn;
}
int tiles32(int n) {
  const int tn = (n + 31) / 32;
  return tn * tn;
}
int tilesSym(int n, int tile) {
  const int tn = (n + tile - 1) / tile;
  return tn * (tn + 1) / 2;
}

// Upper bound of sum over molecules of ceil(n/32)^2 given sum n^2, sum n and the count.
long long tileBound(const WorkspaceCaps& c) { return c.mat / 1024 + 2 * (c.vecN / 32 + 1) + 2LL * c.mols; }

struct Pools {
  ElementDev* elems;
  DiatomicDev* diat;
  AtomDev* atoms;
  PairDev* pairs;
  MolDev *mols, *molsAct;
  int *classPairs, *atomList, *molList, *slotPairs, *slotAtoms, *segs;
  double *ints, *vuvb, *jvec;
  double *S, *H, *F, *P, *Pold, *Tmp, *Fx, *A, *Z, *work;
  unsigned char* piv;
  double *histF, *histE, *histP;
  double *ed, *ee, *etau, *ev, *ew, *eval, *lo, *hi, *lz, *lzCoef;
  double *erepPair, *ePair, *eAtom;
  double *eMol, *erep, *rmsd, *nel, *dots, *coef, *bmat, *eHist, *pf, *mmat;
  ScfMolDev* scf;
  EigJob *eigOcc, *eigSub;
  GemmJob* gemmSub;
  Sp2StepJob *stepX0, *stepX1;
  int *tilesSub, *tiles32, *tilesS, *tiles64;
  double* part[2];
  Sp2Dev* sp2[2];
  double* red;
  unsigned short* lp[4];
};

Pools carve(Carver& cv, const WorkspaceCaps& c, int nElem, int slots) {
  Pools p{};
  const int M = std::max(c.mols, 1);
  const long long tb = tileBound(c);
  p.elems = cv.take<ElementDev>(nElem);
  p.diat = cv.take<DiatomicDev>((size_t)nElem * nElem);
  p.atoms = cv.take<AtomDev>(c.atoms);
  p.pairs = cv.take<PairDev>(c.pairs);
  p.mols = cv.take<MolDev>(M);
  p.molsAct = cv.take<MolDev>(M);
  p.classPairs = cv.take<int>(c.pairs);
  p.atomList = cv.take<int>(c.atoms);
  p.molList = cv.take<int>(M);
  p.slotPairs = cv.take<int>(c.pairs);
  p.slotAtoms = cv.take<int>(c.atoms);
  p.segs = cv.take<int>(3 * ((size_t)M * 11 + 1));
  p.ints = cv.take<double>(c.ints);
  p.vuvb = cv.take<double>(c.pvec);
  p.jvec = cv.take<double>(c.pvec);
  for (double** m : {&p.S, &p.H, &p.F, &p.P, &p.Pold, &p.Tmp, &p.Fx, &p.A, &p.Z}) *m = cv.take<double>(c.mat);
  p.work = cv.take<double>(5 * c.mat);
  p.piv = cv.take<unsigned char>(c.mat);
  p.histF = cv.take<double>(c.hist);
  p.histE = cv.take<double>(c.hist);
  p.histP = cv.take<double>(c.hist);
  for (double** v : {&p.ed, &p.ee, &p.etau, &p.eval, &p.lo, &p.hi}) *v = cv.take<double>(c.vecN);
  p.ev = cv.take<double>(2 * c.vecN);
  p.ew = cv.take<double>(2 * c.vecN);
  p.lz = cv.take<double>(5 * c.vecN);
  p.lzCoef = cv.take<double>((size_t)M * sp2LanczosCoefs());
  p.erepPair = cv.take<double>(c.pairs);
  p.ePair = cv.take<double>(c.pairs);
  p.eAtom = cv.take<double>(c.atoms);
  for (double** v : {&p.eMol, &p.erep, &p.rmsd, &p.nel}) *v = cv.take<double>(M);
  p.dots = cv.take<double>((size_t)M * slots);
  p.coef = cv.take<double>((size_t)M * slots);
  p.bmat = cv.take<double>((size_t)M * slots * slots);
  p.eHist = cv.take<double>((size_t)M * slots);
  p.pf = cv.take<double>((size_t)M * 2 * slots);
  p.mmat = cv.take<double>((size_t)M * slots * slots);
  p.scf = cv.take<ScfMolDev>(M);
  p.eigOcc = cv.take<EigJob>(M);
  p.eigSub = cv.take<EigJob>(M);
  p.gemmSub = cv.take<GemmJob>(M);
  p.stepX0 = cv.take<Sp2StepJob>(M);
  p.stepX1 = cv.take<Sp2StepJob>(M);
  for (int** t : {&p.tilesSub, &p.tiles32, &p.tilesS, &p.tiles64}) *t = cv.take<int>(M + 1);
  p.part[0] = cv.take<double>(2 * (size_t)tb);
  p.part[1] = cv.take<double>(2 * (size_t)tb);
  p.sp2[0] = cv.take<Sp2Dev>(M);
  p.sp2[1] = cv.take<Sp2Dev>(M);
  p.red = cv.take<double>((size_t)tb * 2 * std::max(2, slots));
  for (auto& q : p.lp) q = cv.take<unsigned short>(c.lowp);
  return p;
}

}  // namespace

size_t Workspace::bytesFor(const WorkspaceCaps& caps, int diisSlots) {
  Carver cv{nullptr, 0, 0};
  carve(cv, caps, 128, std::max(2, diisSlots));
  return cv.off + 256;
}

WorkspaceCaps Workspace::capsFor(const std::vector<MolDemand>& mols) {
  WorkspaceCaps c;
  c.mols = (int)mols.size();
  for (const auto& d : mols) {
    c.atoms += d.atoms;
    c.pairs += d.pairs;
    c.ints += d.ints;
    c.pvec += d.pvec;
    c.mat += d.mat;
    c.vecN += d.vecN;
    c.hist += d.hist;
    c.lowp += d.lowp;
  }
  return c;
}

Workspace::Workspace(Engine& engine, ParameterSet& ps, const ScfSettings& st, const WorkspaceCaps& caps, char* base,
                     size_t bytes)
    : eng_(engine), ps_(ps), st_(st), caps_(caps) {
  slotsDiis_ = std::max(2, st.diisSize);
  sp2_ = st.densityBuilder == 0;
  mixed_ = st.precision == 1;
  if (mixed_ && !sp2_) throw std::runtime_error("mixed precision needs the SP2 density builder");
  // element tables: every element of the parameter set
  std::vector<ElementDev> elems;
  for (int Z : ps.elements()) {
    elemIndex_[Z] = (int)elems.size();
    elems.push_back(ps.element(Z, eng_.hostTables()));
  }
  const int ne = (int)elems.size();
  if (ne > 128) throw std::runtime_error("parameter set with more than 128 elements");
  std::vector<DiatomicDev> diat((size_t)ne * ne, DiatomicDev{});
  if (ps.diatomic())
    for (int a = 0; a < ne; ++a)
      for (int b = 0; b < ne; ++b) diat[(size_t)a * ne + b] = ps.diatomicParams(elems[a].Z, elems[b].Z);
  Carver cv{base, 0, bytes};
  Pools p = carve(cv, caps, ne, slotsDiis_);
  dElems_ = p.elems;
  dDiat_ = p.diat;
  dAtoms_ = p.atoms;
  dPairs_ = p.pairs;
  dMols_ = p.mols;
  dMolsAct_ = p.molsAct;
  dClassPairs_ = p.classPairs;
  dAtomList_ = p.atomList;
  dMolList_ = p.molList;
  dSlotPairs_ = p.slotPairs;
  dSlotAtoms_ = p.slotAtoms;
  dSegs_ = p.segs;
  segHost_.assign(3 * ((size_t)std::max(caps.mols, 1) * 11 + 1), 0);
  dInts_ = p.ints;
  dVuvb_ = p.vuvb;
  dJvec_ = p.jvec;
  dS_ = p.S;
  dH_ = p.H;
  dF_ = p.F;
  dP_ = p.P;
  dPold_ = p.Pold;
  dTmp_ = p.Tmp;
  dFx_ = p.Fx;
  dA_ = p.A;
  dZ_ = p.Z;
  dWork_ = p.work;
  dPiv_ = p.piv;
  dHistF_ = p.histF;
  dHistE_ = p.histE;
  dHistP_ = p.histP;
  dEd_ = p.ed;
  dEe_ = p.ee;
  dEtau_ = p.etau;
  dEv_ = p.ev;
  dEw_ = p.ew;
  dEval_ = p.eval;
  dLo_ = p.lo;
  dHi_ = p.hi;
  dLz_ = p.lz;
  dLzCoef_ = p.lzCoef;
  if (const char* b = std::getenv("SPARROW_AMD_SP2_BOUNDS")) {
    const std::string v(b);
    if (v == "gershgorin") lanczosMode_ = 0;
    else if (v == "lanczos") lanczosMode_ = 1;
    else if (v != "auto") throw std::runtime_error("SPARROW_AMD_SP2_BOUNDS must be gershgorin, lanczos or auto");
  }
  dErepPair_ = p.erepPair;
  dEPair_ = p.ePair;
  dEAtom_ = p.eAtom;
  dEMol_ = p.eMol;
  dErep_ = p.erep;
  dRmsd_ = p.rmsd;
  dNel_ = p.nel;
  dDots_ = p.dots;
  dCoef_ = p.coef;
  dBmat_ = p.bmat;
  dEHist_ = p.eHist;
  dPf_ = p.pf;
  dMmat_ = p.mmat;
  dScf_ = p.scf;
  dEigOcc_ = p.eigOcc;
  dEigSub_ = p.eigSub;
  dGemmSub_ = p.gemmSub;
  dStepX0_ = p.stepX0;
  dStepX1_ = p.stepX1;
  dTilesSub_ = p.tilesSub;
  dTiles32_ = p.tiles32;
  dTilesS_ = p.tilesS;
  dTiles64_ = p.tiles64;
  for (int k = 0; k < 2; ++k) {
    dPart_[k] = p.part[k];
    dSp2_[k] = p.sp2[k];
  }
  dRed_ = p.red;
  for (int k = 0; k < 4; ++k) dLp_[k] = p.lp[k];

  hipStream_t s = eng_.stream();
  HIPCHK(hipMemcpyAsync(dElems_, elems.data(), ne * sizeof(ElementDev), hipMemcpyHostToDevice, s));
  if (!diat.empty()) HIPCHK(hipMemcpyAsync(dDiat_, diat.data(), diat.size() * sizeof(DiatomicDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));

  useGraphs_ = st.graphs;
  aAtoms_.reset(caps.atoms);
  aPairs_.reset(caps.pairs);
  aInts_.reset(caps.ints);
  aPvec_.reset(caps.pvec);
  aMat_.reset(caps.mat);
  aVecN_.reset(caps.vecN);
  aHist_.reset(caps.hist);
  aLowp_.reset(caps.lowp);
  slots_.assign(std::max(caps.mols, 0), Slot{});
  eigOcc_.assign(slots_.size(), EigJob{});
  eigAll_.assign(slots_.size(), EigJob{});
  gP_.assign(slots_.size(), GemmJob{});
  gX0_.assign(slots_.size(), Sp2StepJob{});
  gX1_.assign(slots_.size(), Sp2StepJob{});
  sp2State_.assign(std::max<size_t>(slots_.size(), 1), Sp2Dev{});
  scfHost_.assign(std::max<size_t>(slots_.size(), 1), ScfMolDev{});

  long long tiles64 = caps.mat / 4096 + 1;  // GEMM tile by the amount of work the pools can hold
  tileX_ = tiles64 >= 512 ? 64 : 32;

  b_.elems = dElems_;
  b_.diatomic = dDiat_;
  b_.nElem = ne;
  b_.methodKind = (int)ps.kind();
  b_.oneCW = eng_.oneCenterTable(ps);
  b_.atoms = dAtoms_;
  b_.pairs = dPairs_;
  b_.mols = dMols_;
  b_.natoms = (int)caps.atoms;
  b_.npairs = (int)caps.pairs;
  b_.nmols = caps.mols;
}

MolDemand Workspace::demandOf(ParameterSet& ps, const ModelTables& tables, const MoleculeInput& in, int diisSlots,
                               bool mixed) {
  if (in.Z.empty()) throw std::runtime_error("empty molecular structure");
  if (in.xyz.size() != 3 * in.Z.size()) throw std::runtime_error("position array does not match the number of atoms");
  MolDemand d;
  double core = 0;
  std::vector<int> l(in.Z.size());
  for (size_t k = 0; k < in.Z.size(); ++k) {
    if (!ps.hasElement(in.Z[k]))
      throw std::runtime_error("element Z=" + std::to_string(in.Z[k]) + " has no parameters in this method");
    const ElementDev& e = ps.element(in.Z[k], tables);
    d.n += e.nao;
    core += std::lround(e.core);
    l[k] = e.l;
  }
  if (ps.diatomic())
    for (size_t i = 0; i < in.Z.size(); ++i)
      for (size_t j = 0; j <= i; ++j)
        if (!ps.hasPair(in.Z[i], in.Z[j]))
          throw std::runtime_error("NDDO parameters do not exist for element pair Z=" + std::to_string(in.Z[i]) + "-" +
                                   std::to_string(in.Z[j]));
  const long long nel = std::llround(core - in.charge);
  if (nel < 0 || nel % 2 != 0) throw std::runtime_error("restricted closed-shell SCF needs an even, non-negative electron count");
  d.nocc = (int)(nel / 2);
  if (d.nocc > d.n) throw std::runtime_error("more electrons than spin orbitals");
  const long long na = (long long)in.Z.size();
  d.atoms = na;
  d.pairs = na * (na - 1) / 2;
  for (size_t i = 0; i < in.Z.size(); ++i)
    for (size_t j = i + 1; j < in.Z.size(); ++j) {
      d.ints += (long long)nChargeDist(l[i]) * nChargeDist(l[j]);
      d.pvec += nChargeDist(l[i]) + nChargeDist(l[j]);
    }
  const long long n = d.n;
  d.mat = n * n;
  d.vecN = n;
  d.hist = (long long)std::max(2, diisSlots) * n * n;
  const long long ldp = ((n + 31) / 32) * 32;
  d.lowp = mixed ? ldp * ldp : 0;
  return d;
}

int Workspace::admit(const MoleculeInput& in) { return admitImpl(in, nullptr); }
int Workspace::admitWarm(const MoleculeInput& in, const double* devGuess) { return admitImpl(in, devGuess); }

int Workspace::admitImpl(const MoleculeInput& in, const double* devGuess) {
  const MolDemand d = demand(in);
  int slot = -1;
  for (size_t i = 0; i < slots_.size(); ++i)
    if (!slots_[i].used) {
      slot = (int)i;
      break;
    }
  if (slot < 0) return -1;
  long long oA, oP, oI, oV, oM, oN, oH, oL;
  struct Got {
    RangeAlloc* a;
    long long len;
    long long* off;
  } want[] = {{&aAtoms_, d.atoms, &oA}, {&aPairs_, d.pairs, &oP}, {&aInts_, d.ints, &oI}, {&aPvec_, d.pvec, &oV},
              {&aMat_, d.mat, &oM},     {&aVecN_, d.vecN, &oN},   {&aHist_, d.hist, &oH}, {&aLowp_, d.lowp, &oL}};
  int got = 0;
  for (; got < 8; ++got)
    if (!want[got].a->alloc(want[got].len, want[got].off)) break;
  if (got < 8) {
    for (int k = 0; k < got; ++k) want[k].a->release(*want[k].off, want[k].len);
    return -1;
  }
  Slot& S = slots_[slot];
  S = Slot{};
  S.used = true;
  S.dm = d;
  S.intOff = oI;
  S.pvecOff = oV;
  S.lpOff = oL;
  MolDev& md = S.md;
  md.id = slot;
  md.n = d.n;
  md.nocc = d.nocc;
  md.atom0 = (int)oA;
  md.natoms = (int)d.atoms;
  md.pair0 = (int)oP;
  md.npairs = (int)d.pairs;
  md.matOff = oM;
  md.vecOff = oN;
  md.histOff = oH;
  // atoms and pairs (triangular order within the molecule, as the kernels expect)
  std::vector<AtomDev> atoms(d.atoms);
  std::vector<int> lOf(d.atoms);
  int ao = 0;
  double core = 0;
  for (long long k = 0; k < d.atoms; ++k) {
    AtomDev a{};
    a.elem = elemIndex_.at(in.Z[k]);
    a.mol = slot;
    a.ao = ao;
    const ElementDev& e = ps_.element(in.Z[k], eng_.hostTables());
    a.nao = e.nao;
    a.x = in.xyz[3 * k];
    a.y = in.xyz[3 * k + 1];
    a.z = in.xyz[3 * k + 2];
    ao += a.nao;
    core += std::lround(e.core);
    lOf[k] = e.l;
    atoms[k] = a;
  }
  std::vector<int> atomOrder;  // atoms without d shells first
  for (int pass = 0; pass < 2; ++pass)
    for (long long k = 0; k < d.atoms; ++k)
      if ((lOf[k] == 2) == (pass == 1)) atomOrder.push_back(md.atom0 + (int)k);
  for (long long k = 0; k < d.atoms; ++k) S.nAtomsSp += lOf[k] == 2 ? 0 : 1;
  S.nel = core - in.charge;
  std::vector<PairDev> pairs(d.pairs);
  long long intOff = oI, vecOff = oV, q = 0;
  for (int i = 0; i < md.natoms; ++i)
    for (int j = i + 1; j < md.natoms; ++j, ++q) {
      PairDev p{};
      p.a = md.atom0 + i;
      p.b = md.atom0 + j;
      p.mol = slot;
      p.cls = lOf[i] * 3 + lOf[j];
      p.intOff = intOff;
      p.vecOff = vecOff;
      p.aoA = atoms[i].ao;
      p.aoB = atoms[j].ao;
      p.n = md.n;
      p.matOff = md.matOff;
      intOff += (long long)nChargeDist(lOf[i]) * nChargeDist(lOf[j]);
      vecOff += nChargeDist(lOf[i]) + nChargeDist(lOf[j]);
      ++S.classCnt[p.cls];
      pairs[q] = p;
    }
  {
    std::vector<int> dk(d.atoms);
    for (long long k = 0; k < d.atoms; ++k) dk[k] = nChargeDist(lOf[k]);
    setCoulombPartialOffsets(atoms.data(), pairs.data(), md.natoms, dk.data(), oV);
  }
  std::vector<int> pairOrder(d.pairs);  // pair ids sorted by class (triangular order within a class)
  {
    std::array<long long, 9> at{};
    for (int c = 1; c < 9; ++c) at[c] = at[c - 1] + S.classCnt[c - 1];
    for (long long k = 0; k < d.pairs; ++k) pairOrder[at[pairs[k].cls]++] = md.pair0 + (int)k;
  }
  if (in.densityGuess) S.guess.assign(in.densityGuess, in.densityGuess + (size_t)d.n * d.n);
  S.devGuess = devGuess;
  hipStream_t s = eng_.stream();
  HIPCHK(hipMemcpyAsync(dAtoms_ + md.atom0, atoms.data(), atoms.size() * sizeof(AtomDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dSlotAtoms_ + md.atom0, atomOrder.data(), atomOrder.size() * sizeof(int), hipMemcpyHostToDevice, s));
  if (!pairs.empty()) {
    HIPCHK(hipMemcpyAsync(dPairs_ + md.pair0, pairs.data(), pairs.size() * sizeof(PairDev), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dSlotPairs_ + md.pair0, pairOrder.data(), pairOrder.size() * sizeof(int), hipMemcpyHostToDevice, s));
  }
  HIPCHK(hipMemcpyAsync(dMols_ + slot, &md, sizeof(MolDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dNel_ + slot, &S.nel, sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));  // host vectors above go out of scope

  // job templates of this slot
  const MolDev& m = md;
  EigJob J{};
  J.n = m.n;
  J.nvec = m.nocc;
  J.A = dA_ + m.matOff;
  J.d = dEd_ + m.vecOff;
  J.e = dEe_ + m.vecOff;
  J.tau = dEtau_ + m.vecOff;
  J.v = dEv_ + 2 * m.vecOff;
  J.w = dEw_ + 2 * m.vecOff;
  J.wval = dEval_ + m.vecOff;
  J.Z = dZ_ + m.matOff;
  J.work = dWork_ + 5 * m.matOff;
  J.ipiv = dPiv_ + m.matOff;
  eigOcc_[slot] = J;
  J.nvec = m.n;
  eigAll_[slot] = J;
  GemmJob g{};  // P = 2 Z_occ Z_occ^T (Z row-major n x nocc)
  g.M = g.N = m.n;
  g.K = m.nocc;
  g.A = g.B = dZ_ + m.matOff;
  g.C = dP_ + m.matOff;
  g.D = g.C;
  g.sAi = g.sBj = m.nocc;
  g.sAk = g.sBk = 1;
  g.sCi = g.sDi = 1;
  g.sCj = g.sDj = m.n;
  g.alpha = 2.0;
  gP_[slot] = g;
  Sp2StepJob x{};  // SP2 steps between A (even iterates) and Z (odd iterates)
  x.n = m.n;
  x.mol = slot;
  x.tile = eng_.sp2Tile(m.n, st_.batched);
  x.X = dA_ + m.matOff;
  x.Y = dZ_ + m.matOff;
  if (mixed_) {
    x.ldp = ((m.n + 31) / 32) * 32;
    x.Xh = dLp_[0] + S.lpOff;
    x.Xl = dLp_[1] + S.lpOff;
    x.Yh = dLp_[2] + S.lpOff;
    x.Yl = dLp_[3] + S.lpOff;
  }
  gX0_[slot] = x;
  x.X = dZ_ + m.matOff;
  x.Y = dA_ + m.matOff;
  if (mixed_) {
    x.Xh = dLp_[2] + S.lpOff;
    x.Xl = dLp_[3] + S.lpOff;
    x.Yh = dLp_[0] + S.lpOff;
    x.Yl = dLp_[1] + S.lpOff;
  }
  gX1_[slot] = x;
  pending_.push_back(slot);
  return slot;
}

// Set-up of the molecules admitted since the last iteration: two-centre integrals, S, H,
// repulsion, density guess (or the warm-start density), SCF state.
void Workspace::runSetup() {
  if (pending_.empty()) return;
  hipStream_t s = eng_.stream();
  KernelProfiler& prof = eng_.profiler();
  auto timed = [&](const char* name, auto&& launch) {
    if (prof.enabled()) prof.begin(name, s);
    launch();
    if (prof.enabled()) prof.end(s);
  };
  const double t0 = nowSeconds();
  HIPCHK(hipMemcpyAsync(dMolList_, pending_.data(), pending_.size() * sizeof(int), hipMemcpyHostToDevice, s));
  KBatch b = b_;
  gatherLists(pending_, b);
  b.molList = dMolList_;
  b.nMolList = (int)pending_.size();
  timed("twocenter", [&] { launchTwoCenter(eng_.tables(), b, dInts_, dVuvb_, s); });
  timed("pair_overlap", [&] { launchPairOverlap(b, dS_, dH_, dErepPair_, s); });
  timed("atom_hcore", [&] { launchAtomHcore(eng_.tables(), b, dVuvb_, dH_, s); });
  launchMirrorLower(b, dS_, dH_, s);
  launchMolPairSum(b, dErepPair_, dErep_, s);  // molecules without pairs: their sum is 0
  launchDensityGuess(b, dS_, dNel_, dP_, s);
  std::vector<ScfMolDev> init(pending_.size());
  for (size_t k = 0; k < pending_.size(); ++k) {
    Slot& S = slots_[pending_[k]];
    const size_t nn = (size_t)S.md.n * S.md.n;
    if (S.devGuess)
      HIPCHK(hipMemcpyAsync(dP_ + S.md.matOff, S.devGuess, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
    else if (!S.guess.empty())
      HIPCHK(hipMemcpyAsync(dP_ + S.md.matOff, S.guess.data(), nn * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dPold_ + S.md.matOff, dP_ + S.md.matOff, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
    init[k] = ScfMolDev{};
    init[k].iter = 1;
    init[k].lowpArmed = mixed_ ? 1 : 0;
    S.scf = init[k];
    HIPCHK(hipMemcpyAsync(dScf_ + pending_[k], &init[k], sizeof(ScfMolDev), hipMemcpyHostToDevice, s));
  }
  HIPCHK(hipStreamSynchronize(s));  // host lists above go out of scope
  for (int slot : pending_) {
    Slot& S = slots_[slot];
    S.setupDone = true;
    S.inFock = S.inIter = true;
    S.guess.clear();
    S.guess.shrink_to_fit();
    S.devGuess = nullptr;
    ++nActive_;
  }
  pending_.clear();
  fockDirty_ = iterDirty_ = true;
  times.setup += nowSeconds() - t0;
}

void Workspace::gatherLists(const std::vector<int>& slots, KBatch& b) {
  hipStream_t s = eng_.stream();
  int* seg = segHost_.data();
  int nseg = 0, dst = 0;
  auto add = [&](int from, int len) {
    if (len <= 0) return;
    seg[3 * nseg] = from;
    seg[3 * nseg + 1] = dst;
    seg[3 * nseg + 2] = len;
    ++nseg;
    dst += len;
  };
  for (int c = 0; c < 9; ++c) {  // pairs: class by class, the slots in list order within a class
    b.classStart[c] = dst;
    for (int slot : slots) {
      const Slot& S = slots_[slot];
      int first = S.md.pair0;
      for (int q = 0; q < c; ++q) first += S.classCnt[q];
      add(first, S.classCnt[c]);
    }
  }
  b.classStart[9] = dst;
  const int nPairSegs = nseg;
  dst = 0;
  for (int pass = 0; pass < 2; ++pass) {  // atoms without d shells first
    for (int slot : slots) {
      const Slot& S = slots_[slot];
      if (pass == 0) add(S.md.atom0, S.nAtomsSp);
      else add(S.md.atom0 + S.nAtomsSp, S.md.natoms - S.nAtomsSp);
    }
    if (pass == 0) b.nAtomListSp = dst;
  }
  b.nAtomList = dst;
  if ((size_t)(3 * nseg) > segHost_.size()) throw std::runtime_error("internal: work-list segments exceed their buffer");
  if (nseg > 0) HIPCHK(hipMemcpyAsync(dSegs_, seg, 3 * (size_t)nseg * sizeof(int), hipMemcpyHostToDevice, s));
  launchGatherSegments(dSlotPairs_, dClassPairs_, dSegs_, nPairSegs, s);
  launchGatherSegments(dSlotAtoms_, dAtomList_, dSegs_ + 3 * nPairSegs, nseg - nPairSegs, s);
  b.classPairs = dClassPairs_;
  b.atomList = dAtomList_;
}

void Workspace::buildFockLists() {
  hipStream_t s = eng_.stream();
  fockMols_.clear();
  for (size_t i = 0; i < slots_.size(); ++i)
    if (slots_[i].used && slots_[i].inFock) fockMols_.push_back((int)i);
  HIPCHK(hipMemcpyAsync(dMolList_, fockMols_.data(), fockMols_.size() * sizeof(int), hipMemcpyHostToDevice, s));
  gatherLists(fockMols_, b_);
  b_.molList = dMolList_;
  b_.nMolList = (int)fockMols_.size();
  HIPCHK(hipStreamSynchronize(s));
  fockDirty_ = false;
  ++listVersion_;
}

void Workspace::buildIterLists() {
  hipStream_t s = eng_.stream();
  IL_ = IterLists{};
  std::vector<Sp2StepJob> x0, x1;
  IL_.t32.push_back(0);
  IL_.tS.push_back(0);
  IL_.t64.push_back(0);
  std::vector<int> order;  // iterating slots, by SP2 tile edge (the step launches take them in ranges)
  for (int tile : {32, 64, 128})
    for (size_t i = 0; i < slots_.size(); ++i)
      if (slots_[i].used && slots_[i].inIter && gX0_[i].tile == tile) order.push_back((int)i);
  for (int i : order) {
    const Slot& S = slots_[i];
    const MolDev& m = S.md;
    IL_.mols.push_back(m);
    x0.push_back(gX0_[i]);
    x1.push_back(gX1_[i]);
    IL_.t32.push_back(IL_.t32.back() + tiles32(m.n));
    IL_.tS.push_back(IL_.tS.back() + tilesSym(m.n, gX0_[i].tile));
    (gX0_[i].tile == 32 ? IL_.ntS32 : gX0_[i].tile == 64 ? IL_.ntS64 : IL_.ntS128) += tilesSym(m.n, gX0_[i].tile);
    IL_.maxN = std::max(IL_.maxN, m.n);
    IL_.maxNocc = std::max(IL_.maxNocc, m.nocc);
    if (gX0_[i].tile == 32) {
      ++IL_.k32;
      IL_.nt32Small = IL_.t32.back();
    } else {
      const int t64 = (m.n + 63) / 64;
      IL_.t64.push_back(IL_.t64.back() + t64 * t64);
    }
  }
  IL_.nt64 = IL_.t64.back();
  const int k = (int)IL_.mols.size();
  IL_.nt32 = IL_.t32.back();
  IL_.ntS = IL_.tS.back();
  HIPCHK(hipMemcpyAsync(dMolsAct_, IL_.mols.data(), k * sizeof(MolDev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dStepX0_, x0.data(), k * sizeof(Sp2StepJob), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dStepX1_, x1.data(), k * sizeof(Sp2StepJob), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dTiles32_, IL_.t32.data(), (k + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dTilesS_, IL_.tS.data(), (k + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dTiles64_, IL_.t64.data(), IL_.t64.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  iterDirty_ = false;
  ++listVersion_;
}

// Diagonalise the extrapolated Fock matrices of the listed molecules and build P = 2 C_occ C_occ^T.
// Tridiagonalisation of the jobs in dJobs (device copies of jobs): the first nBatched together in
// one one-workgroup-per-matrix launch, the rest one after the other on the many-workgroup path.
static void tridiagonalise(EigJob* dJobs, const std::vector<EigJob>& jobs, int nBatched, hipStream_t s) {
  if (nBatched > 0) launchSytrd(dJobs, nBatched, s);
  for (size_t q = (size_t)nBatched; q < jobs.size(); ++q) launchSytrdWide(dJobs + q, jobs[q].n, s);
}

void Workspace::eigenDensity(const std::vector<int>& idsIn) {
  hipStream_t s = eng_.stream();
  KernelProfiler& prof = eng_.profiler();
  auto timed = [&](const char* name, auto&& launch) {
    if (prof.enabled()) prof.begin(name, s);
    launch();
    if (prof.enabled()) prof.end(s);
  };
  // molecules of kWideEigenMinN AOs or more are tridiagonalised on the many-workgroup path (a
  // choice by size alone, so a molecule's density does not depend on its batch), listed last
  std::vector<int> ids;
  for (int pass = 0; pass < 2; ++pass)
    for (int id : idsIn)
      if ((slots_[id].md.n >= kWideEigenMinN) == (pass == 1)) ids.push_back(id);
  const int nb = (int)ids.size();
  if (nb == 0) return;
  int nBatched = 0;
  for (int id : ids) nBatched += slots_[id].md.n < kWideEigenMinN ? 1 : 0;
  std::vector<EigJob> ej(nb);
  std::vector<GemmJob> gj(nb);
  std::vector<int> ts(nb + 1, 0);
  int maxN = 0, maxO = 0;
  for (int q = 0; q < nb; ++q) {
    const MolDev& m = slots_[ids[q]].md;
    ej[q] = eigOcc_[ids[q]];
    gj[q] = gP_[ids[q]];
    ts[q + 1] = ts[q] + gemmTiles(m.n, tileX_);
    maxN = std::max(maxN, m.n);
    maxO = std::max(maxO, m.nocc);
    HIPCHK(hipMemcpyAsync(dA_ + m.matOff, dFx_ + m.matOff, (size_t)m.n * m.n * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  HIPCHK(hipMemcpyAsync(dEigSub_, ej.data(), nb * sizeof(EigJob), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dGemmSub_, gj.data(), nb * sizeof(GemmJob), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dTilesSub_, ts.data(), (nb + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  timed("sytrd", [&] { tridiagonalise(dEigSub_, ej, nBatched, s); });
  timed("stebz", [&] { launchStebz(dEigSub_, nb, maxN, s); });
  timed("stein", [&] { launchStein(dEigSub_, nb, maxO, s); });
  timed("ormtr", [&] { launchOrmtr(dEigSub_, nb, maxN, maxO, s); });
  timed("gemm_p", [&] { launchGemm(dGemmSub_, dTilesSub_, nb, ts[nb], tileX_, s); });
  ++counters_.eigensolves;
  HIPCHK(hipStreamSynchronize(s));  // the host job vectors above must outlive the uploads
}

Workspace::~Workspace() {
  for (GraphSlot* g : {&segA_, &segB_})
    for (auto& kv : g->exec) (void)hipGraphExecDestroy(kv.second);
}

template <class Launches>
void Workspace::runSegment(GraphSlot& g, long long key, Launches&& launches) {
  hipStream_t s = eng_.stream();
  if (!useGraphs_ || eng_.profiler().enabled()) {
    launches();
    return;
  }
  auto it = g.exec.find(key);
  if (it == g.exec.end()) {
    if (g.seen.size() > 4096) g.seen.clear();  // keys of list versions long gone
    if (g.seen[key]++ == 0) {  // a launch sequence seen once is not worth a capture
      launches();
      return;
    }
    if (g.order.size() >= 16) {
      HIPCHK(hipGraphExecDestroy(g.exec.at(g.order.front())));
      g.exec.erase(g.order.front());
      g.order.erase(g.order.begin());
    }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      launches();
    } catch (...) {
      (void)hipStreamEndCapture(s, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    HIPCHK(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    it = g.exec.emplace(key, exec).first;
    g.order.push_back(key);
    ++counters_.graphCaptures;
  }
  HIPCHK(hipGraphLaunch(it->second, s));
  ++counters_.graphLaunches;
}

// Fock build and energy of the molecules in the Fock list, then the extrapolated Fock matrix Fx
// of the k iterating molecules.
void Workspace::launchFockAndMixer(int k) {
  hipStream_t s = eng_.stream();
  KernelProfiler& prof = eng_.profiler();
  auto timed = [&](const char* name, auto&& launch) {
    if (prof.enabled()) prof.begin(name, s);
    launch();
    if (prof.enabled()) prof.end(s);
  };
  const int slots = slotsDiis_;
  timed("fock_pairs", [&] { launchFockPairs(eng_.tables(), b_, dInts_, dP_, dH_, dF_, dJvec_, dEPair_, s); });
  timed("fock_atoms", [&] { launchFockAtoms(eng_.tables(), b_, dP_, dH_, dJvec_, dF_, dEAtom_, s); });
  launchEnergyReduce(b_, dEAtom_, dEPair_, dEMol_, s);
  if (k == 0) return;
  if (st_.mixer != 0) {  // 1 DIIS, 2 EDIIS, 3 EDIIS+DIIS
    launchCopyMats(dMolsAct_, k, IL_.maxN, dF_, false, dHistF_, true, dScf_, slots, s);
    if (st_.mixer != 2) {
      timed("gemm_fp", [&] {
        launchFPProduct(dMolsAct_, dTiles32_, k, IL_.nt32Small, dF_, dP_, dTmp_, s, IL_.k32, dTiles64_, IL_.nt64);
      });
      timed("diis_dots",
            [&] { launchCommutatorDots(dMolsAct_, dTiles32_, k, IL_.nt32, dTmp_, dHistE_, slots, dScf_, dRed_, dDots_, s); });
      launchDiisSolve(dMolsAct_, k, dScf_, dDots_, dBmat_, dCoef_, slots, s);
    }
    if (st_.mixer >= 2) {
      launchCopyMats(dMolsAct_, k, IL_.maxN, dP_, false, dHistP_, true, dScf_, slots, s);  // the P this F was built from
      timed("ediis", [&] {
        launchEdiis(dMolsAct_, dTiles32_, k, IL_.nt32, dHistP_, dHistF_, slots, dScf_, dEMol_, dEHist_, dRed_, dPf_, dDots_,
                    dMmat_, dCoef_, st_.mixer == 3, s);
      });
    }
    launchDiisExtrapolate(dMolsAct_, k, IL_.maxN, dHistF_, slots, dCoef_, dFx_, s);
  } else {
    launchCopyMats(dMolsAct_, k, IL_.maxN, dF_, false, dFx_, false, nullptr, slots, s);
  }
}

void Workspace::iterate() {
  runSetup();
  if (nActive_ == 0) return;
  if (fockDirty_) buildFockLists();
  if (iterDirty_) buildIterLists();
  hipStream_t s = eng_.stream();
  KernelProfiler& prof = eng_.profiler();
  auto timed = [&](const char* name, auto&& launch) {
    if (prof.enabled()) prof.begin(name, s);
    launch();
    if (prof.enabled()) prof.end(s);
  };
  const int k = (int)IL_.mols.size();
  // end of the iteration: density change, convergence test, SCF states to the host
  auto finish = [&] {
    if (k > 0) timed("rmsd", [&] { launchRmsdUpdate(dMolsAct_, dTiles32_, k, IL_.nt32, dP_, dPold_, dRed_, dRmsd_, s); });
    launchScfCheck(dMolList_, b_.nMolList, dScf_, dEMol_, dRmsd_, st_.energyTol, st_.rmsdTol, st_.lowpStart, st_.lowpTol,
                   st_.maxIterations, s);
    HIPCHK(hipMemcpyAsync(scfHost_.data(), dScf_, slots_.size() * sizeof(ScfMolDev), hipMemcpyDeviceToHost, s));
  };
  const double ta = nowSeconds();
  double tb = ta, tEig = 0;
  if (sp2_ && k > 0) {
    // segment A: Fock, mixer, SP2 start and a first chunk of SP2 steps sized by the last
    // iteration (steps after a molecule's X is final return at once), then the SP2 states
    const int maxSteps = st_.sp2MaxSteps;
    // (rounded up to a multiple of 4 for graph replay, so that the launch sequence repeats)
    const int want = lastSp2Steps_ > 0 ? lastSp2Steps_ + 1 : 32;
    const int chunk = std::min(maxSteps, useGraphs_ ? (want + 1 + 3) / 4 * 4 : want);
    runSegment(segA_, (listVersion_ << 12) | chunk, [&] {
      launchFockAndMixer(k);
      timed("sp2_init", [&] {
        // Lanczos bounds (9 launches of O(n^2) work per iteration) for molecules of at least
        // kLanczosMinN AOs, where about 4 SP2 steps saved outweigh them; the choice depends on the
        // molecule alone, so a molecule's SCF is the same in any batch
        const bool lz = lanczosMode_ != 0;
        launchSp2Init(dMolsAct_, k, IL_.maxN, dFx_, dStepX0_, dTilesS_, IL_.ntS, dLo_, dHi_, dPart_[0], dSp2_[0], maxSteps,
                      dScf_, s, lz ? dLz_ : nullptr, lz ? dLzCoef_ : nullptr, lanczosMode_ == 1 ? 0 : kLanczosMinN);
      });
      for (int c = 0; c < chunk; ++c) {
        const int e = c % 2;
        timed("gemm_sp2", [&] {
          launchSp2Step(e == 0 ? dStepX0_ : dStepX1_, dTilesS_, k, IL_.ntS32, IL_.ntS64, IL_.ntS128, dPart_[e], dPart_[1 - e],
                        dSp2_[e], dSp2_[1 - e], s);
        });
      }
      // decision on the newest X, for the host's check only (a further step repeats it)
      const int e = chunk % 2;
      launchSp2Decide(e == 0 ? dStepX0_ : dStepX1_, dTilesS_, k, dPart_[e], dSp2_[e], dSp2_[1 - e], s);
      HIPCHK(hipMemcpyAsync(sp2State_.data(), dSp2_[1 - e], slots_.size() * sizeof(Sp2Dev), hipMemcpyDeviceToHost, s));
    });
    HIPCHK(hipStreamSynchronize(s));
    tb = nowSeconds();
    int steps = chunk;
    for (;;) {  // further chunks of 4 while a molecule's projection runs
      bool all = true;
      for (const MolDev& m : IL_.mols) all &= sp2State_[m.id].done != 0;
      if (all || steps >= maxSteps) break;
      for (int c = 0; c < 4 && steps < maxSteps; ++c, ++steps) {
        const int e = steps % 2;
        timed("gemm_sp2", [&] {
          launchSp2Step(e == 0 ? dStepX0_ : dStepX1_, dTilesS_, k, IL_.ntS32, IL_.ntS64, IL_.ntS128, dPart_[e], dPart_[1 - e],
                        dSp2_[e], dSp2_[1 - e], s);
        });
      }
      const int e = steps % 2;
      launchSp2Decide(e == 0 ? dStepX0_ : dStepX1_, dTilesS_, k, dPart_[e], dSp2_[e], dSp2_[1 - e], s);
      HIPCHK(hipMemcpyAsync(sp2State_.data(), dSp2_[1 - e], slots_.size() * sizeof(Sp2Dev), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
    }
    counters_.sp2Steps += steps;
    int used = 0;
    std::vector<int> bad;  // projections that did not converge (e.g. a vanishing gap): diagonalised
    for (const MolDev& m : IL_.mols) {
      const Sp2Dev& q = sp2State_[m.id];
      used = std::max(used, q.doneStep);
      // symmetric product: one triangle with the diagonal, n^2 (n + 1) flops per step
      const long long symFlops = (long long)m.n * m.n * (m.n + 1);
      counters_.sp2Flops += symFlops * q.doneStep;
      counters_.sp2LowpFlops += symFlops * q.lowpSteps;
      if (!q.converged) bad.push_back(m.id);
    }
    lastSp2Steps_ = used;
    Sp2Dev* finalState = dSp2_[1 - steps % 2];
    if (bad.empty()) {
      // segment B: P = 2 X, RMSD, convergence test
      runSegment(segB_, (listVersion_ << 2) | (steps % 2), [&] {
        launchPFromX(dMolsAct_, k, finalState, dA_, dZ_, dP_, s);
        finish();
      });
    } else {
      launchPFromX(dMolsAct_, k, finalState, dA_, dZ_, dP_, s);
      counters_.sp2Fallbacks += (long long)bad.size();
      const double t0 = nowSeconds();
      eigenDensity(bad);
      tEig = nowSeconds() - t0;
      finish();
    }
  } else {  // diagonalisation builder, or final energies only
    launchFockAndMixer(k);
    if (k > 0) {
      std::vector<int> ids;
      for (const MolDev& m : IL_.mols) ids.push_back(m.id);
      const double t0 = nowSeconds();
      eigenDensity(ids);
      tEig = nowSeconds() - t0;
    }
    finish();
  }
  HIPCHK(hipStreamSynchronize(s));
  const double te = nowSeconds();
  times.fock += tb - ta;
  times.eig += tEig;
  times.density += te - tb - tEig;
  ++times.iterations;
  ++counters_.iterations;
  for (int id : fockMols_) {
    Slot& S = slots_[id];
    S.scf = scfHost_[id];
    if (S.inIter && (S.scf.converged || S.scf.stopped)) {
      S.inIter = false;
      iterDirty_ = true;
    }
    if (S.inFock && S.scf.haveEnergy) {
      S.inFock = false;
      S.finished = true;
      --nActive_;
      fockDirty_ = true;
    }
  }
}

std::vector<int> Workspace::finishedSlots() const {
  std::vector<int> v;
  for (size_t i = 0; i < slots_.size(); ++i)
    if (slots_[i].used && slots_[i].finished) v.push_back((int)i);
  return v;
}

const double* Workspace::densityOf(int slot) const { return dP_ + slots_[slot].md.matOff; }

std::vector<MoleculeResult> Workspace::collect(const std::vector<int>& ids, bool wantEps, bool wantOrbitals,
                                               bool wantDensity) {
  hipStream_t s = eng_.stream();
  std::vector<MoleculeResult> res(ids.size());
  std::vector<double> erep(slots_.size());  // repulsion energies of every slot, one copy
  if (!slots_.empty())
    HIPCHK(hipMemcpyAsync(erep.data(), dErep_, slots_.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  for (size_t q = 0; q < ids.size(); ++q) {
    const Slot& S = slots_.at(ids[q]);
    if (!S.used || !S.finished) throw std::runtime_error("collect: molecule not finished");
    const MolDev& m = S.md;
    MoleculeResult& r = res[q];
    r.nao = m.n;
    r.nocc = m.nocc;
    if (wantDensity) {
      r.P.resize((size_t)m.n * m.n);
      HIPCHK(hipMemcpyAsync(r.P.data(), dP_ + m.matOff, r.P.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    }
  }
  // orbital energies (and all eigenvectors) of each molecule's last extrapolated Fock matrix,
  // frozen in place since the molecule stopped iterating
  if ((wantEps || wantOrbitals) && !ids.empty()) {
    // a few molecules (a single solve, small batches): matrices of kWideSytrdMinN AOs or more on the
    // many-workgroup path, listed last; many molecules: all in one batched launch (these orbitals
    // do not feed back into any SCF)
    const bool few = ids.size() <= 4;
    std::vector<int> order;
    for (int pass = 0; pass < 2; ++pass)
      for (size_t q = 0; q < ids.size(); ++q)
        if ((few && slots_[ids[q]].md.n >= kWideSytrdMinN) == (pass == 1)) order.push_back((int)q);
    std::vector<EigJob> ej;
    int maxN = 0, nBatched = 0;
    for (int q : order) {
      const MolDev& m = slots_[ids[q]].md;
      ej.push_back(eigAll_[ids[q]]);
      maxN = std::max(maxN, m.n);
      nBatched += (few && m.n >= kWideSytrdMinN) ? 0 : 1;
      HIPCHK(hipMemcpyAsync(dA_ + m.matOff, dFx_ + m.matOff, (size_t)m.n * m.n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    const int nb = (int)ej.size();
    HIPCHK(hipMemcpyAsync(dEigSub_, ej.data(), nb * sizeof(EigJob), hipMemcpyHostToDevice, s));
    tridiagonalise(dEigSub_, ej, nBatched, s);
    launchStebz(dEigSub_, nb, maxN, s);
    DeviceBuffer gemmJobs;  // the reflector-block GEMMs of the matrices on the many-workgroup path
    if (wantOrbitals) {
      launchStein(dEigSub_, nb, maxN, s);
      launchOrmtr(dEigSub_, nBatched, maxN, maxN, s);
      int most = 0;
      for (int q = nBatched; q < nb; ++q) most = std::max(most, ormtrGemmJobs(ej[q].n));
      if (most > 0) {
        gemmJobs.allocate(sizeof(GemmJob) * (size_t)most + 256);
        GemmJob* dG = reinterpret_cast<GemmJob*>(gemmJobs.ptr);
        int* dZero = reinterpret_cast<int*>(gemmJobs.ptr + sizeof(GemmJob) * (size_t)most);
        HIPCHK(hipMemsetAsync(dZero, 0, sizeof(int), s));
        std::vector<GemmJob> hG(most);
        for (int q = nBatched; q < nb; ++q) {
          launchOrmtrGemm(ej[q], dEigSub_ + q, hG.data(), dG, dZero, s);
          HIPCHK(hipStreamSynchronize(s));  // hG and dG are reused by the next matrix
        }
      }
    }
    ++counters_.eigensolves;
    std::vector<double> Zr;
    for (size_t q = 0; q < ids.size(); ++q) {
      const MolDev& m = slots_[ids[q]].md;
      res[q].eps.resize(m.n);
      HIPCHK(hipMemcpyAsync(res[q].eps.data(), dEval_ + m.vecOff, m.n * sizeof(double), hipMemcpyDeviceToHost, s));
      if (wantOrbitals) {
        Zr.resize((size_t)m.n * m.n);
        res[q].C.resize((size_t)m.n * m.n);
        HIPCHK(hipMemcpyAsync(Zr.data(), dZ_ + m.matOff, Zr.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      }
      HIPCHK(hipStreamSynchronize(s));
      // Z is row-major n x n (row = AO, column = MO) == column-major C with C[k*n + r]
      if (wantOrbitals)
        for (int r = 0; r < m.n; ++r)
          for (int c = 0; c < m.n; ++c) res[q].C[(size_t)c * m.n + r] = Zr[(size_t)r * m.n + c];
    }
  }
  HIPCHK(hipStreamSynchronize(s));
  for (size_t q = 0; q < ids.size(); ++q) {
    Slot& S = slots_[ids[q]];
    MoleculeResult& r = res[q];
    r.electronicEnergy = S.scf.Eel;
    r.repulsionEnergy = erep[ids[q]];
    r.energy = r.electronicEnergy + r.repulsionEnergy;
    r.converged = S.scf.converged != 0;
    r.iterations = S.scf.converged ? S.scf.convIter : st_.maxIterations;
    // release the slot's ranges
    const MolDemand& d = S.dm;
    aAtoms_.release(S.md.atom0, d.atoms);
    aPairs_.release(S.md.pair0, d.pairs);
    aInts_.release(S.intOff, d.ints);
    aPvec_.release(S.pvecOff, d.pvec);
    aMat_.release(S.md.matOff, d.mat);
    aVecN_.release(S.md.vecOff, d.vecN);
    aHist_.release(S.md.histOff, d.hist);
    aLowp_.release(S.lpOff, d.lowp);
    S = Slot{};
  }
  return res;
}

}  // namespace sparrow_amd
