// Fock build F = H + G(P) on gfx950 (restricted, closed shell).
//
// Reference: NU/TwoElectronMatrix.cpp:63-166 (G), NU/FockMatrix.cpp:108-122 (F = H + G),
// NU/NDDOElectronicEnergyCalculator.cpp:48-67 (E_el).  The reference fills the lower triangle;
// here every matrix is stored full and symmetric (n x n, column-major per molecule).
//
//   atom block A:   G_AA(mu,nu) = sum_{ls in A} P_ls [(mu nu|l s) - 1/2 (mu l|nu s)]        (one-centre)
//                               + sum_{B != A} sum_{ls in B} P_ls (mu nu|l s)                 (Coulomb)
//   pair block AB:  G_BA(l,mu)  = -1/2 sum_{nu in A, s in B} P_{s nu} (mu nu|l s)              (exchange)
//
// The pair kernel (one wave per atom pair) stages the pair's d_A x d_B integral block in LDS and
// produces the exchange block (owned exclusively by that pair, so no atomics) plus the two
// Coulomb partial vectors J_A(P) = sum_Q I(P,Q) w_B(Q) and J_B(Q) = sum_P I(P,Q) w_A(P)
// (w = P over charge distributions with multiplicity 2 off the diagonal).  The atom kernel adds
// the one-centre part and the Coulomb partials of all partners in a fixed order, so the result
// is deterministic.  Both kernels also emit their share of E_el = 1/2 sum_ij P_ij (H_ij + F_ij).
#include <algorithm>
#include <cstdlib>
#include <string>

#include "kernels.hpp"

namespace sparrow_amd {

namespace {

__host__ __device__ constexpr int dimL(int l) { return l == 0 ? 1 : l == 1 ? 10 : 40; }
__host__ __device__ constexpr int naoL(int l) { return l == 0 ? 1 : l == 1 ? 4 : 9; }

// s/p charge distributions: (i, j), i <= j in {s, x, y, z}, numbered j (j + 1) / 2 + i, the
// order of the model's distribution table for its first ten entries; as arithmetic, so the
// s/p pair classes address their blocks without table lookups
__host__ __device__ constexpr int spPair(int i, int j) { return i > j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }
__host__ __device__ constexpr int spHi(int p) { return p >= 6 ? 3 : p >= 3 ? 2 : p >= 1 ? 1 : 0; }
__host__ __device__ constexpr int spLo(int p) { return p - spHi(p) * (spHi(p) + 1) / 2; }

__device__ inline double waveSum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}


// Per-wave LDS of one atom pair: the d_A x d_B integral block, then P_AA, P_BB and P_BA.
constexpr int kPairLds = 1600 + 5 * 81;  // + H of the exchange block and its energy terms

// Loads the pair's integral block and density blocks into the team's LDS region (all integral
// loads issued before the stores: one round of memory latency).  A team is TS lanes of a wave
// (TS = 64: the whole wave); tl is the lane's index in its team.
template <int LA, int LB, int TS = 64>
__device__ inline void fockPairStage(const KBatch& b, const double* __restrict__ ints, const double* __restrict__ P,
                                     const double* __restrict__ H, int pid, double* lds, int tl) {
  constexpr int DA = dimL(LA), DB = dimL(LB), NA = naoL(LA), NB = naoL(LB);
  double* sI = lds;
  double* sPA = sI + DA * DB;
  double* sPB = sPA + NA * NA;
  double* sPBA = sPB + NB * NB;
  const PairDev pr = b.pairs[pid];
  const long long n = pr.n;
  const double* Pm = P + pr.matOff;
  const int aoA = pr.aoA, aoB = pr.aoB;
  constexpr int NI = (DA * DB + TS - 1) / TS;
  double v[NI];
#pragma unroll
  for (int r = 0; r < NI; ++r) {
    const int o = tl + TS * r;
    v[r] = o < DA * DB ? ints[pr.intOff + o] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < NI; ++r) {
    const int o = tl + TS * r;
    if (o < DA * DB) sI[o] = v[r];
  }
  for (int k = tl; k < NA * NA; k += TS) sPA[k] = Pm[(long long)(aoA + k / NA) * n + aoA + k % NA];
  for (int k = tl; k < NB * NB; k += TS) sPB[k] = Pm[(long long)(aoB + k / NB) * n + aoB + k % NB];
  double* sH = sPBA + NB * NA;  // H of the exchange block, loaded with the densities
  const double* Hm = H + pr.matOff;
  for (int k = tl; k < NB * NA; k += TS) {
    int lam = k / NA, mu = k - lam * NA;
    const long long pos = (long long)(aoA + mu) * n + aoB + lam;
    sPBA[k] = Pm[pos];
    sH[k] = Hm[pos];
  }
}

// From the staged blocks: the two Coulomb partial vectors, the exchange block of F (both
// triangles) and the pair's share of E_el.  The energy terms of the block's elements go through
// LDS (sE, NA NB doubles after the staged blocks) and one lane adds them in element order, so the
// share does not depend on the team size.
template <int LA, int LB, int TS = 64>
__device__ inline void fockPairFinish(const KBatch& b, const double* __restrict__ H, double* __restrict__ F,
                                      double* __restrict__ jvec, double* __restrict__ ePair, int pid, double* lds,
                                      const uint8_t* sPairIdx, const uint8_t* sPairOrb, int tl) {
  constexpr int DA = dimL(LA), DB = dimL(LB), NA = naoL(LA), NB = naoL(LB);
  const double* sI = lds;
  const double* sPA = sI + DA * DB;
  const double* sPB = sPA + NA * NA;
  const double* sPBA = sPB + NB * NB;
  const double* sH = sPBA + NB * NA;
  double* sE = lds + DA * DB + NA * NA + NB * NB + 2 * NA * NB;
  const PairDev pr = b.pairs[pid];
  const long long n = pr.n;
  if constexpr (LA <= 1 && LB <= 1) {
    // s/p classes: the same sums in the same order, indices from spPair/spLo/spHi
    for (int Pp = tl; Pp < DA; Pp += TS) {
      double s = 0.0;
#pragma unroll
      for (int Q = 0; Q < DB; ++Q) {
        const int l = spLo(Q), sg = spHi(Q);
        const double wq = sPB[l * NB + sg] * (l == sg ? 1.0 : 2.0);
        s += sI[Pp * DB + Q] * wq;
      }
      jvec[pr.jA + Pp] = s;
    }
    for (int Q = tl; Q < DB; Q += TS) {
      double s = 0.0;
#pragma unroll
      for (int Pp = 0; Pp < DA; ++Pp) {
        const int mu = spLo(Pp), nu = spHi(Pp);
        const double wp = sPA[mu * NA + nu] * (mu == nu ? 1.0 : 2.0);
        s += sI[Pp * DB + Q] * wp;
      }
      jvec[pr.jB + Q] = s;
    }
    double* Fm = F + pr.matOff;
    for (int o = tl; o < NB * NA; o += TS) {
      const int lam = o / NA, mu = o - lam * NA;
      double k = 0.0;
#pragma unroll
      for (int nu = 0; nu < NA; ++nu) {
        const int Pp = spPair(mu, nu);
#pragma unroll
        for (int sg = 0; sg < NB; ++sg) k += sPBA[sg * NA + nu] * sI[Pp * DB + spPair(lam, sg)];
      }
      const long long r = pr.aoB + lam, c = pr.aoA + mu;
      const double h = sH[o];
      const double f = h - 0.5 * k;
      Fm[c * n + r] = f;
      Fm[r * n + c] = f;
      sE[o] = sPBA[o] * (h + f);
    }
  } else {
  // Coulomb partial vectors (unrolled by 4: the table and block reads of a group of terms are
  // issued ahead of their sum, which keeps its order)
  for (int Pp = tl; Pp < DA; Pp += TS) {
    double s = 0.0;
#pragma unroll 4
    for (int Q = 0; Q < DB; ++Q) {
      int l = sPairOrb[2 * Q], sg = sPairOrb[2 * Q + 1];
      double wq = sPB[l * NB + sg] * (l == sg ? 1.0 : 2.0);
      s += sI[Pp * DB + Q] * wq;
    }
    jvec[pr.jA + Pp] = s;
  }
  for (int Q = tl; Q < DB; Q += TS) {
    double s = 0.0;
#pragma unroll 4
    for (int Pp = 0; Pp < DA; ++Pp) {
      int mu = sPairOrb[2 * Pp], nu = sPairOrb[2 * Pp + 1];
      double wp = sPA[mu * NA + nu] * (mu == nu ? 1.0 : 2.0);
      s += sI[Pp * DB + Q] * wp;
    }
    jvec[pr.jB + Q] = s;
  }
  // exchange block and its energy terms: the element's NB distribution indices on B read once,
  // then the NA NB terms in (nu, sigma) order
  double* Fm = F + pr.matOff;
  for (int o = tl; o < NB * NA; o += TS) {
    const int lam = o / NA, mu = o - lam * NA;
    int qIdx[NB];
#pragma unroll
    for (int sg = 0; sg < NB; ++sg) qIdx[sg] = sPairIdx[lam * 9 + sg];
    double k = 0.0;
    for (int nu = 0; nu < NA; ++nu) {
      const int Pp = sPairIdx[mu * 9 + nu];
      if (Pp == kNoPair) continue;
#pragma unroll
      for (int sg = 0; sg < NB; ++sg) {
        if (qIdx[sg] == kNoPair) continue;
        k += sPBA[sg * NA + nu] * sI[Pp * DB + qIdx[sg]];
      }
    }
    const long long r = pr.aoB + lam, c = pr.aoA + mu;
    const double h = sH[o];
    const double f = h - 0.5 * k;
    Fm[c * n + r] = f;
    Fm[r * n + c] = f;
    sE[o] = sPBA[o] * (h + f);
  }
  }
  // the team's lanes wrote sE: within a wave, LDS writes are visible to the wave's later reads
  // once the wave waits for them (wave-level barrier of the team's lanes)
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this lane's LDS writes done
  __builtin_amdgcn_wave_barrier();
  if (tl == 0) {
    double e = 0.0;
    for (int o = 0; o < NB * NA; ++o) e += sE[o];
    ePair[pid] = e;
  }
}

// One launch per (l_A, l_B) class.  A pair takes a team of TS lanes sized to its work: 1 for H-H
// (one integral, one exchange element), 4 for H-sp (10 integrals, 4 exchange elements), 8 for d-H
// (40 and 9), 16 for sp-sp (100 and 16), the whole wave for a d atom with an sp or d atom; W waves
// per workgroup.  The results do not depend on the team size.
template <int LA, int LB>
struct FockTeam {
  static constexpr int TS = (LA == 0 && LB == 0)   ? 1
                            : (LA + LB == 1)       ? 4
                            : (LA == 2 || LB == 2) ? (LA + LB == 2 ? 8 : 64)
                                                   : 16;
  static constexpr int W = (dimL(LA) * dimL(LB) >= 1600) ? 2 : 4;
  static constexpr int perBlock = W * (64 / TS);  // pairs per workgroup
};

// Items of the persistent launch: a wave's share of one class, 64 / TS pairs of that class (one per
// team of TS lanes).  Item i of class c covers its pairs [i * 64 / TS, (i + 1) * 64 / TS).
template <int LA, int LB>
__device__ inline int fockItems(int pairs) {
  constexpr int per = 64 / FockTeam<LA, LB>::TS;
  return (pairs + per - 1) / per;
}

// One item: each team stages its pair into its part of the wave's LDS region, then (after a
// wave-level barrier: the region belongs to this wave alone) finishes it.
template <int LA, int LB>
__device__ inline void fockItem(const KBatch& b, const double* __restrict__ ints, const double* __restrict__ P,
                                const double* __restrict__ H, double* __restrict__ F, double* __restrict__ jvec,
                                double* __restrict__ ePair, int item, double* lds, const uint8_t* sPairIdx,
                                const uint8_t* sPairOrb, int lane) {
  constexpr int TS = FockTeam<LA, LB>::TS, DA = dimL(LA), DB = dimL(LB), NA = naoL(LA), NB = naoL(LB);
  constexpr int L = DA * DB + NA * NA + NB * NB + 3 * NA * NB;
  constexpr int cls = LA * 3 + LB;
  static_assert((64 / TS) * L <= kPairLds, "an item's staged blocks must fit the wave's LDS region");
  const int team = lane / TS, tl = lane % TS;
  const int idx = b.classStart[cls] + item * (64 / TS) + team;
  const bool active = idx < b.classStart[cls + 1];
  double* mine = lds + team * L;
  const int pid = active ? b.classPairs[idx] : 0;
  if (active) fockPairStage<LA, LB, TS>(b, ints, P, H, pid, mine, tl);
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this lane's LDS writes done
  __builtin_amdgcn_wave_barrier();
  if (active) fockPairFinish<LA, LB, TS>(b, H, F, jvec, ePair, pid, mine, sPairIdx, sPairOrb, tl);
  __builtin_amdgcn_s_waitcnt(0xc07f);  // the finish's LDS reads done before the region is reused
  __builtin_amdgcn_wave_barrier();
}

// The pair part of every listed molecule's Fock build in one persistent launch: the pairs of each
// (l_A, l_B) class are cut into items of 64 / TS pairs (the class's team size, as in the per-class
// launches), and wave g of the grid takes items g, g + G, g + 2G, ... of the concatenated item
// list (G = waves in the grid).  Waves run independently (each owns its LDS region); the results
// are those of the per-class launches.
__global__ __launch_bounds__(256) void k_fock_pairs_all(KTables t, KBatch b, const double* __restrict__ ints,
                                                        const double* __restrict__ P, const double* __restrict__ H,
                                                        double* __restrict__ F, double* __restrict__ jvec,
                                                        double* __restrict__ ePair) {
  __shared__ double sAll[4][kPairLds];
  __shared__ uint8_t sPairIdx[81], sPairOrb[80];  // charge-distribution tables, read in the inner loops
  for (int q = threadIdx.x; q < 81; q += blockDim.x) sPairIdx[q] = t.pairIndex[q];
  for (int q = threadIdx.x; q < 80; q += blockDim.x) sPairOrb[q] = t.pairOrb[q];
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* lds = sAll[w];
  int itemStart[10];
  itemStart[0] = 0;
  {
    const int* cs = b.classStart;
    itemStart[1] = itemStart[0] + fockItems<0, 0>(cs[1] - cs[0]);
    itemStart[2] = itemStart[1] + fockItems<0, 1>(cs[2] - cs[1]);
    itemStart[3] = itemStart[2] + fockItems<0, 2>(cs[3] - cs[2]);
    itemStart[4] = itemStart[3] + fockItems<1, 0>(cs[4] - cs[3]);
    itemStart[5] = itemStart[4] + fockItems<1, 1>(cs[5] - cs[4]);
    itemStart[6] = itemStart[5] + fockItems<1, 2>(cs[6] - cs[5]);
    itemStart[7] = itemStart[6] + fockItems<2, 0>(cs[7] - cs[6]);
    itemStart[8] = itemStart[7] + fockItems<2, 1>(cs[8] - cs[7]);
    itemStart[9] = itemStart[8] + fockItems<2, 2>(cs[9] - cs[8]);
  }
  const int G = gridDim.x * 4;
  for (int i = blockIdx.x * 4 + w; i < itemStart[9]; i += G) {
    int cls = 0;
    while (i >= itemStart[cls + 1]) ++cls;
    const int item = i - itemStart[cls];
    switch (cls) {
      case 0: fockItem<0, 0>(b, ints, P, H, F, jvec, ePair, item, lds, sPairIdx, sPairOrb, lane); break;
      case 1: fockItem<0, 1>(b, ints, P, H, F, jvec, ePair, item, lds, sPairIdx, sPairOrb, lane); break;
      case 2: fockItem<0, 2>(b, ints, P, H, F, jvec, ePair, item, lds, sPairIdx, sPairOrb, lane); break;
      case 3: fockItem<1, 0>(b, ints, P, H, F, jvec, ePair, item, lds, sPairIdx, sPairOrb, lane); break;
      case 4: fockItem<1, 1>(b, ints, P, H, F, jvec, ePair, item, lds, sPairIdx, sPairOrb, lane); break;
      case 5: fockItem<1, 2>(b, ints, P, H, F, jvec, ePair, item, lds, sPairIdx, sPairOrb, lane); break;
      case 6: fockItem<2, 0>(b, ints, P, H, F, jvec, ePair, item, lds, sPairIdx, sPairOrb, lane); break;
      case 7: fockItem<2, 1>(b, ints, P, H, F, jvec, ePair, item, lds, sPairIdx, sPairOrb, lane); break;
      default: fockItem<2, 2>(b, ints, P, H, F, jvec, ePair, item, lds, sPairIdx, sPairOrb, lane); break;
    }
  }
}

template <int LA, int LB>
__global__ __launch_bounds__(256) void k_fock_pairs(KTables t, KBatch b, const double* __restrict__ ints,
                                                    const double* __restrict__ P, const double* __restrict__ H,
                                                    double* __restrict__ F, double* __restrict__ jvec,
                                                    double* __restrict__ ePair) {
  constexpr int DA = dimL(LA), DB = dimL(LB), NA = naoL(LA), NB = naoL(LB);
  constexpr int TS = FockTeam<LA, LB>::TS, TPB = FockTeam<LA, LB>::perBlock;
  constexpr int cls = LA * 3 + LB;
  constexpr int L = DA * DB + NA * NA + NB * NB + 3 * NA * NB;
  __shared__ double sAll[TPB][L];
  __shared__ uint8_t sPairIdx[81], sPairOrb[80];
  for (int q = threadIdx.x; q < 81; q += blockDim.x) sPairIdx[q] = t.pairIndex[q];
  for (int q = threadIdx.x; q < 80; q += blockDim.x) sPairOrb[q] = t.pairOrb[q];
  const int team = threadIdx.x / TS, tl = threadIdx.x % TS;
  const int idx = b.classStart[cls] + blockIdx.x * TPB + team;
  const bool active = idx < b.classStart[cls + 1];
  const int pid = active ? b.classPairs[idx] : 0;
  if (active) fockPairStage<LA, LB, TS>(b, ints, P, H, pid, sAll[team], tl);
  __syncthreads();
  if (active) fockPairFinish<LA, LB, TS>(b, H, F, jvec, ePair, pid, sAll[team], sPairIdx, sPairOrb, tl);
}

template <int LA, int LB>
void launchFockClass(const KTables& t, const KBatch& b, const double* ints, const double* P, const double* H, double* F,
                     double* jvec, double* ePair, hipStream_t s) {
  constexpr int cls = LA * 3 + LB;
  int np = b.classStart[cls + 1] - b.classStart[cls];
  if (np <= 0) return;
  constexpr int TPB = FockTeam<LA, LB>::perBlock, W = FockTeam<LA, LB>::W;
  hipLaunchKernelGGL((k_fock_pairs<LA, LB>), dim3((np + TPB - 1) / TPB), dim3(64 * W), 0, s, t, b, ints, P, H, F, jvec,
                     ePair);
}

// One 4-wave workgroup per atom: one-centre part + Coulomb partials of all partners + H.  The
// one-centre coefficients W(ij,kl) = (ij|kl) - 1/2 (ik|jl) come from the parameter set's dense
// table (one nao^2 x nao^2 block per element).  MAXNA = 4 serves the atoms without d shells,
// MAXNA = 9 those with.
constexpr int kAtomThreads = 256;
template <int MAXNA>
__device__ inline void fockAtom(const KTables& t, const KBatch& b, int a, const double* __restrict__ P,
                                const double* __restrict__ H, const double* __restrict__ jvec, double* __restrict__ F,
                                double* __restrict__ eAtom) {
  constexpr int MAXD = MAXNA == 9 ? 40 : 10, MAXNA2 = MAXNA * MAXNA, NW = kAtomThreads / 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const AtomDev A = b.atoms[a];
  const ElementDev& E = b.elems[A.elem];
  const MolDev m = b.mols[A.mol];
  const long long n = m.n;
  const int na = E.nao, dA = dimL(E.l), na2 = na * na;
  const int nat = m.natoms;
  if (na > MAXNA) {  // the host partitions the list by shell type; a wrong list shows as NaN
    if (tid == 0) eAtom[a] = __builtin_nan("");
    return;
  }
  __shared__ double sJ[MAXD], sP[MAXNA2], sAcc[NW][MAXD], sE[NW];
  __shared__ uint8_t sPairIdx[81];
  for (int k = tid; k < na2; k += kAtomThreads) sP[k] = P[m.matOff + (long long)(A.ao + k / na) * n + A.ao + k % na];
  for (int k = tid; k < 81; k += kAtomThreads) sPairIdx[k] = t.pairIndex[k];
  // Coulomb partials J(Pp): the atom's nat - 1 partner vectors are contiguous (d_A doubles each,
  // partners in index order); each thread sums its partners in order, each wave its threads
  // (shuffle tree), then the waves in order
  double acc[MAXD];
#pragma unroll
  for (int Pp = 0; Pp < MAXD; ++Pp) acc[Pp] = 0.0;
  const double* __restrict__ jA = jvec + A.jOff;
  for (int q = tid; q < nat - 1; q += kAtomThreads) {
    const double* v = jA + (long long)q * dA;
#pragma unroll
    for (int Pp = 0; Pp < MAXD; ++Pp)
      if (Pp < dA) acc[Pp] += v[Pp];
  }
#pragma unroll
  for (int Pp = 0; Pp < MAXD; ++Pp)
    if (Pp < dA) {
      const double v = waveSum(acc[Pp]);
      if (lane == 0) sAcc[w][Pp] = v;
    }
  __syncthreads();
  if (tid < dA) {
    double v = 0.0;
    for (int q = 0; q < NW; ++q) v += sAcc[q][tid];
    sJ[tid] = v;
  }
  __syncthreads();
  // one-centre part, one thread per element ij of the block: G_ij = sum_kl P_kl W(ij,kl), the
  // table stored by column (W(ij,kl) at kl nao^2 + ij: consecutive threads read adjacent values)
  const double* __restrict__ W = b.oneCW + (long long)E.w1c * kOneCenterBlock;
  double e = 0.0;
  for (int o = tid; o < na2; o += kAtomThreads) {
    const int i = o / na, j = o - i * na;
    double g = 0.0;
#pragma unroll MAXNA
    for (int kl = 0; kl < na2; ++kl) g += sP[kl] * W[kl * na2 + o];
    const int Pp = sPairIdx[i * 9 + j];
    if (Pp != kNoPair) g += sJ[Pp];
    const long long pos = (long long)(A.ao + j) * n + A.ao + i;
    const double h = H[m.matOff + pos];
    const double f = h + g;
    F[m.matOff + pos] = f;
    e += 0.5 * sP[j * na + i] * (h + f);
  }
  e = waveSum(e);
  if (lane == 0) sE[w] = e;
  __syncthreads();
  if (tid == 0) {
    double v = 0.0;
    for (int q = 0; q < NW; ++q) v += sE[q];
    eAtom[a] = v;
  }
}

// atomList[first, first + grid): atoms of one shell type
template <int MAXNA>
__global__ __launch_bounds__(kAtomThreads) void k_fock_atoms(KTables t, KBatch b, int first, const double* __restrict__ P,
                                                             const double* __restrict__ H, const double* __restrict__ jvec,
                                                             double* __restrict__ F, double* __restrict__ eAtom) {
  fockAtom<MAXNA>(t, b, b.atomList[first + blockIdx.x], P, H, jvec, F, eAtom);
}

// both shell types in one launch (the few atoms with d shells first, so their longer blocks run
// beside the others instead of after them): blocks [0, nd) take atomList[nsp, nsp + nd), the
// rest atomList[0, nsp)
__global__ __launch_bounds__(kAtomThreads) void k_fock_atoms_mixed(KTables t, KBatch b, int nsp, int nd,
                                                                   const double* __restrict__ P, const double* __restrict__ H,
                                                                   const double* __restrict__ jvec, double* __restrict__ F,
                                                                   double* __restrict__ eAtom) {
  if ((int)blockIdx.x < nd)
    fockAtom<9>(t, b, b.atomList[nsp + blockIdx.x], P, H, jvec, F, eAtom);
  else
    fockAtom<4>(t, b, b.atomList[blockIdx.x - nd], P, H, jvec, F, eAtom);
}

// E_el of each molecule of molList: its atom and pair shares summed in a fixed order (each thread
// strided, four loads in flight, then the waves and the block).
__global__ __launch_bounds__(1024) void k_energy_reduce(KBatch b, const double* __restrict__ eAtom,
                                                        const double* __restrict__ ePair, double* __restrict__ eMol) {
  const int id = b.molList[blockIdx.x];
  const MolDev m = b.mols[id];
  constexpr int T = 1024;
  double s = 0.0;
  for (int k = threadIdx.x; k < m.natoms; k += T) s += eAtom[m.atom0 + k];
  const double* ep = ePair + m.pair0;
  int k = threadIdx.x;
  for (; k + 3 * T < m.npairs; k += 4 * T) {
    const double a0 = ep[k], a1 = ep[k + T], a2 = ep[k + 2 * T], a3 = ep[k + 3 * T];
    s += a0;
    s += a1;
    s += a2;
    s += a3;
  }
  for (; k < m.npairs; k += T) s += ep[k];
  __shared__ double red[T / 64];
  s = waveSum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < T / 64; ++w) t += red[w];
    eMol[id] = t;
  }
}

// ---------------------------------------------------------------------------------------------
// Spin-polarised (UHF) Fock matrices (TwoElectronMatrix spinPolarized_ branches,
// NU/TwoElectronMatrix.cpp:82-108,142-160): G^s = J(P) - K(P^s) with P = Pa + Pb.  In NDDO the
// two-centre Coulomb terms fall on atom-diagonal blocks and the two-centre exchange on
// off-diagonal blocks, so from restricted builds G_R(D) = J(D) - K(D)/2:
//   off-diagonal blocks  G^s = 2 G_R(P^s)
//   diagonal blocks      G^s = G_R(P) -+ X/2,  X_ij = sum_kl (Pa - Pb)_kl (ik|jl)  (one-centre)
// ---------------------------------------------------------------------------------------------
// out = a + b (n x n)
__global__ __launch_bounds__(256) void k_mat_add(long long nn, const double* __restrict__ a, const double* __restrict__ b,
                                                 double* __restrict__ out) {
  for (long long k = blockIdx.x * 256LL + threadIdx.x; k < nn; k += (long long)gridDim.x * 256) out[k] = a[k] + b[k];
}

// F^s from F_T = H + G_R(P) (diagonal blocks) and F_s = H + G_R(P^s) (off-diagonal: 2 F_s - H)
__global__ __launch_bounds__(256) void k_uhf_assemble(int n, const int* __restrict__ aoAtom, const double* __restrict__ H,
                                                      const double* __restrict__ FT, const double* __restrict__ Fa1,
                                                      const double* __restrict__ Fb1, double* __restrict__ Fa,
                                                      double* __restrict__ Fb) {
  const long long nn = (long long)n * n;
  for (long long k = blockIdx.x * 256LL + threadIdx.x; k < nn; k += (long long)gridDim.x * 256) {
    const int r = (int)(k % n), c = (int)(k / n);
    const bool same = aoAtom[r] == aoAtom[c];
    Fa[k] = same ? FT[k] : 2.0 * Fa1[k] - H[k];
    Fb[k] = same ? FT[k] : 2.0 * Fb1[k] - H[k];
  }
}

// the one-centre same-spin exchange of the spin density on every atom's diagonal block, one
// 256-thread workgroup per atom (K table stored by column: K(ij,kl) at kl nao^2 + ij)
__global__ __launch_bounds__(256) void k_uhf_onecenter(KBatch b, const double* __restrict__ WK, int n,
                                                       const double* __restrict__ Pa, const double* __restrict__ Pb,
                                                       double* __restrict__ Fa, double* __restrict__ Fb) {
  const int a = blockIdx.x, tid = threadIdx.x;
  const AtomDev A = b.atoms[a];
  const ElementDev& E = b.elems[A.elem];
  const int na = E.nao, na2 = na * na;
  __shared__ double sD[81];
  for (int k = tid; k < na2; k += 256) {
    const long long pos = (long long)(A.ao + k % na) * n + A.ao + k / na;  // (k / na, k % na)
    sD[k] = Pa[pos] - Pb[pos];
  }
  __syncthreads();
  const double* __restrict__ W = WK + (long long)E.w1c * kOneCenterBlock;
  for (int o = tid; o < na2; o += 256) {
    double x = 0.0;
    for (int kl = 0; kl < na2; ++kl) x += sD[kl] * W[kl * na2 + o];
    const long long pos = (long long)(A.ao + o % na) * n + A.ao + o / na;  // (i, j) = (o / na, o % na)
    Fa[pos] -= 0.5 * x;
    Fb[pos] += 0.5 * x;
  }
}

// E_el = 1/2 sum_ij [P_ij H_ij + Pa_ij Fa_ij + Pb_ij Fb_ij] (NDDOElectronicEnergyCalculator::
// unrestrictedEnergy, NU/NDDOElectronicEnergyCalculator.cpp:69-89), one workgroup, fixed order
__global__ __launch_bounds__(1024) void k_uhf_energy(long long nn, const double* __restrict__ P, const double* __restrict__ H,
                                                     const double* __restrict__ Pa, const double* __restrict__ Fa,
                                                     const double* __restrict__ Pb, const double* __restrict__ Fb,
                                                     double* __restrict__ e) {
  double v = 0.0;
  for (long long k = threadIdx.x; k < nn; k += 1024) v += P[k] * H[k] + Pa[k] * Fa[k] + Pb[k] * Fb[k];
  __shared__ double red[16];
  v = waveSum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += red[w];
    *e = 0.5 * t;
  }
}

}  // namespace

void launchFockPairs(const KTables& t, const KBatch& b, const double* ints, const double* P, const double* H, double* F,
                     double* jvec, double* ePair, hipStream_t s) {
  // SPARROW_AMD_FOCK=classes / persistent selects one launch per pair class / the persistent
  // launch for every list (kept for comparison)
  static const int forced = [] {
    const char* e = getenv("SPARROW_AMD_FOCK");
    return !e ? 0 : std::string(e) == "classes" ? 1 : std::string(e) == "persistent" ? 2 : 0;
  }();
  const bool classes = forced == 1;
  const int total = b.classStart[9];
  if (total <= 0) return;
  // Up to 160k pairs (one molecule of up to ~560 atoms, or a batch of 32 molecules of 100 atoms)
  // one persistent launch beats nine launches: its 64 KB of LDS per workgroup allows two
  // resident 4-wave workgroups per CU, enough for lists this size.  Larger lists keep more pairs
  // in flight with the class launches' smaller LDS footprints.  Measured per Fock build: one
  // molecule of 450 atoms 65 us persistent against 77 us per class, 600 atoms 114 us against
  // 97 us; 16 x 100 atoms 38 us against 74 us, 32 x 100 atoms 56 us against 76 us, and the
  // 256-molecule 10-300-atom mix (3M pairs) 1.25 ms against 0.55 ms.  Both give the same
  // results.
  if (forced == 2 || (!classes && total <= 160000)) {
    const int blocks = std::min((total + 3) / 4, 512);
    hipLaunchKernelGGL(k_fock_pairs_all, dim3(blocks), dim3(256), 0, s, t, b, ints, P, H, F, jvec, ePair);
    return;
  }
  launchFockClass<0, 0>(t, b, ints, P, H, F, jvec, ePair, s);
  launchFockClass<0, 1>(t, b, ints, P, H, F, jvec, ePair, s);
  launchFockClass<0, 2>(t, b, ints, P, H, F, jvec, ePair, s);
  launchFockClass<1, 0>(t, b, ints, P, H, F, jvec, ePair, s);
  launchFockClass<1, 1>(t, b, ints, P, H, F, jvec, ePair, s);
  launchFockClass<1, 2>(t, b, ints, P, H, F, jvec, ePair, s);
  launchFockClass<2, 0>(t, b, ints, P, H, F, jvec, ePair, s);
  launchFockClass<2, 1>(t, b, ints, P, H, F, jvec, ePair, s);
  launchFockClass<2, 2>(t, b, ints, P, H, F, jvec, ePair, s);
}

void launchFockAtoms(const KTables& t, const KBatch& b, const double* P, const double* H, const double* jvec, double* F,
                     double* eAtom, hipStream_t s) {
  const int nsp = b.nAtomListSp, nd = b.nAtomList - b.nAtomListSp;
  if (nd > 0)
    hipLaunchKernelGGL(k_fock_atoms_mixed, dim3(nd + nsp), dim3(kAtomThreads), 0, s, t, b, nsp, nd, P, H, jvec, F, eAtom);
  else if (nsp > 0)
    hipLaunchKernelGGL(k_fock_atoms<4>, dim3(nsp), dim3(kAtomThreads), 0, s, t, b, 0, P, H, jvec, F, eAtom);
}

void launchUnrestrictedFock(const KBatch& b, const double* WK, int n, const int* aoAtom, const double* H, const double* FT,
                            const double* Fa1, const double* Fb1, const double* P, const double* Pa, const double* Pb,
                            double* Fa, double* Fb, double* e, hipStream_t s) {
  const long long nn = (long long)n * n;
  const int blocks = (int)std::min<long long>((nn + 255) / 256, 4096);
  hipLaunchKernelGGL(k_uhf_assemble, dim3(blocks), dim3(256), 0, s, n, aoAtom, H, FT, Fa1, Fb1, Fa, Fb);
  if (b.natoms > 0) hipLaunchKernelGGL(k_uhf_onecenter, dim3(b.natoms), dim3(256), 0, s, b, WK, n, Pa, Pb, Fa, Fb);
  hipLaunchKernelGGL(k_uhf_energy, dim3(1), dim3(1024), 0, s, nn, P, H, Pa, Fa, Pb, Fb, e);
}

void launchMatAdd(long long nn, const double* a, const double* b, double* out, hipStream_t s) {
  if (nn > 0) hipLaunchKernelGGL(k_mat_add, dim3((unsigned)std::min<long long>((nn + 255) / 256, 4096)), dim3(256), 0, s, nn, a, b, out);
}

void launchEnergyReduce(const KBatch& b, const double* eAtom, const double* ePair, double* eMol, hipStream_t s) {
  if (b.nMolList == 0) return;
  hipLaunchKernelGGL(k_energy_reduce, dim3(b.nMolList), dim3(1024), 0, s, b, eAtom, ePair, eMol);
}

}  // namespace sparrow_amd
