// CIS AO Fock matrices of a block of pseudo-densities on gfx950.
//
// Reference: N/TimeDependent/LinearResponse/CISMatrixAOFockBuilder.cpp:48-170 (the Coulomb and
// exchange blocks of every atom and atom pair) and :171-330 (their contraction with the
// pseudo-density), N/TimeDependent/LinearResponse/CISSigmaVectorEvaluator.cpp:96-125 (sigma).
//
// For each vector v of the block (matrices n x n, column-major, v * matStride apart):
//   F = cJ J(DJ) - cK K(DK^T),   J(Y)_xy = sum_cd (xy|cd) Y_cd,   K(Y)_xy = sum_cd (xc|yd) Y_cd
// in the NDDO approximation.  The pseudo-density D = C_occ X^T C_vir^T is not symmetric, so
// the pair kernel writes both exchange blocks F_AB and F_BA, each from its own density block.
// Restricted singlet: cJ = 2 c1, DJ = DK = D; triplet: cJ = 0; unrestricted (spin s): cJ = c1,
// DJ = D^alpha + D^beta, DK = D^s; cK = c2.
//
// The two-centre integral blocks are the ones the SCF computes (resident in HBM, d_A x d_B per
// pair); each wave stages one pair's block once for a group of vectors, and the groups of one
// pair are adjacent work items, so the waves that share a block run side by side (L2 hits).
#include "kernels.hpp"

namespace sparrow_amd {

namespace {

__host__ __device__ constexpr int dimL(int l) { return l == 0 ? 1 : l == 1 ? 10 : 40; }
__host__ __device__ constexpr int naoL(int l) { return l == 0 ? 1 : l == 1 ? 4 : 9; }

__device__ inline double waveSum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wave per (pair, group of G vectors), G = 64 / (n_A n_B) (at least 1), so that the lanes of
// the small classes (s-s, sp-s, sp-sp) have work: stage the pair's integral block once and, per
// vector of the group, the four density blocks it needs; then the Coulomb partial vectors
// (J_A = sum_Q I(P,Q) w_B(Q), J_B = sum_P I(P,Q) w_A(P), w = DJ over charge distributions, both
// orders of an off-diagonal distribution) and the two exchange blocks
//   F(A i, B k) = -cK sum_{j in A, l in B} (ij|kl) DK(B l, A j)
//   F(B k, A i) = -cK sum_{j in A, l in B} (ij|kl) DK(A j, B l),
// lanes over (vector of the group, output element).
template <int LA, int LB>
struct CisPairShape {
  static constexpr int DA = dimL(LA), DB = dimL(LB), NA = naoL(LA), NB = naoL(LB);
  static constexpr int NO = NA * NB;                         // exchange outputs per vector
  static constexpr int G = NO >= 64 ? 1 : 64 / NO;           // vectors per wave
  static constexpr int W = (DA * DB >= 1600) ? 2 : 4;        // waves per workgroup
  static constexpr int LD = NA * NA + NB * NB + 2 * NA * NB;  // staged density doubles per vector
  static constexpr int L = DA * DB + G * LD;
};

template <int LA, int LB>
__global__ __launch_bounds__(256) void k_cis_pairs(KTables t, KBatch b, const double* __restrict__ ints, int nv,
                                                   const double* __restrict__ DJ, const double* __restrict__ DK,
                                                   long long matStride, double cJ, double cK, double* __restrict__ jvec,
                                                   long long vecStride, double* __restrict__ F) {
  using S = CisPairShape<LA, LB>;
  constexpr int DA = S::DA, DB = S::DB, NA = S::NA, NB = S::NB, NO = S::NO, G = S::G, W = S::W, LD = S::LD;
  constexpr int cls = LA * 3 + LB;
  __shared__ double sAll[W][S::L];
  __shared__ uint8_t sPairIdx[81], sPairOrb[80];
  for (int q = threadIdx.x; q < 81; q += blockDim.x) sPairIdx[q] = t.pairIndex[q];
  for (int q = threadIdx.x; q < 80; q += blockDim.x) sPairOrb[q] = t.pairOrb[q];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long np = b.classStart[cls + 1] - b.classStart[cls];
  const long long groups = (nv + G - 1) / G;
  const long long item = (long long)blockIdx.x * W + w;
  const bool active = item < np * groups;
  const int pid = active ? b.classPairs[b.classStart[cls] + item / groups] : 0;
  const int v0 = active ? (int)(item % groups) * G : 0;
  const int gv = active ? min(G, nv - v0) : 0;  // vectors of this wave
  const PairDev pr = b.pairs[pid];
  const AtomDev A = b.atoms[pr.a], B = b.atoms[pr.b];
  const MolDev m = b.mols[pr.mol];
  const long long n = m.n;
  double* sI = sAll[w];
  // per vector g at sD + g LD: DJ(A i, A j) at i NA + j | DJ(B k, B l) at k NB + l |
  // DK(B l, A j) at l NA + j | DK(A j, B l) at j NB + l
  double* sD = sI + DA * DB;
  constexpr int oJB = NA * NA, oKBA = oJB + NB * NB, oKAB = oKBA + NB * NA;
  if (active) {
    for (int o = lane; o < DA * DB; o += 64) sI[o] = ints[pr.intOff + o];
    for (int idx = lane; idx < gv * LD; idx += 64) {
      const int g = idx / LD, k = idx - g * LD;
      const long long off = (long long)(v0 + g) * matStride + m.matOff;
      double val;
      if (k < oJB) {
        val = cJ != 0.0 ? DJ[off + (A.ao + k % NA) * n + A.ao + k / NA] : 0.0;
      } else if (k < oKBA) {
        const int kk = k - oJB;
        val = cJ != 0.0 ? DJ[off + (B.ao + kk % NB) * n + B.ao + kk / NB] : 0.0;
      } else if (k < oKAB) {
        const int kk = k - oKBA;
        val = DK[off + (A.ao + kk % NA) * n + B.ao + kk / NA];
      } else {
        const int kk = k - oKAB;
        val = DK[off + (B.ao + kk % NB) * n + A.ao + kk / NB];
      }
      sD[idx] = val;
    }
  }
  __syncthreads();
  if (!active) return;
  if (cJ != 0.0) {
    for (int idx = lane; idx < gv * DA; idx += 64) {
      const int g = idx / DA, Pp = idx - g * DA;
      const double* sJB = sD + g * LD + oJB;
      double s = 0.0;
      for (int Q = 0; Q < DB; ++Q) {
        const int l = sPairOrb[2 * Q], sg = sPairOrb[2 * Q + 1];
        const double wq = l == sg ? sJB[l * NB + l] : sJB[l * NB + sg] + sJB[sg * NB + l];
        s += sI[Pp * DB + Q] * wq;
      }
      jvec[(long long)(v0 + g) * vecStride + pr.vecOff + Pp] = s;
    }
    for (int idx = lane; idx < gv * DB; idx += 64) {
      const int g = idx / DB, Q = idx - g * DB;
      const double* sJA = sD + g * LD;
      double s = 0.0;
      for (int Pp = 0; Pp < DA; ++Pp) {
        const int mu = sPairOrb[2 * Pp], nu = sPairOrb[2 * Pp + 1];
        const double wp = mu == nu ? sJA[mu * NA + mu] : sJA[mu * NA + nu] + sJA[nu * NA + mu];
        s += sI[Pp * DB + Q] * wp;
      }
      jvec[(long long)(v0 + g) * vecStride + pr.vecOff + DA + Q] = s;
    }
  }
  for (int idx = lane; idx < gv * NO; idx += 64) {
    const int g = idx / NO, o = idx - g * NO;
    const int i = o / NB, k = o - i * NB;
    const double* sKBA = sD + g * LD + oKBA;
    const double* sKAB = sD + g * LD + oKAB;
    double x1 = 0.0, x2 = 0.0;
    for (int j = 0; j < NA; ++j) {
      const int Pp = sPairIdx[i * 9 + j];
      if (Pp == kNoPair) continue;
      for (int l = 0; l < NB; ++l) {
        const int Q = sPairIdx[k * 9 + l];
        if (Q == kNoPair) continue;
        const double I = sI[Pp * DB + Q];
        x1 += I * sKBA[l * NA + j];
        x2 += I * sKAB[j * NB + l];
      }
    }
    double* Fv = F + (long long)(v0 + g) * matStride + m.matOff;
    Fv[(long long)(B.ao + k) * n + A.ao + i] = -cK * x1;
    Fv[(long long)(A.ao + i) * n + B.ao + k] = -cK * x2;
  }
}

template <int LA, int LB>
void launchCisClass(const KTables& t, const KBatch& b, const double* ints, int nv, const double* DJ, const double* DK,
                    long long matStride, double cJ, double cK, double* jvec, long long vecStride, double* F,
                    hipStream_t s) {
  constexpr int cls = LA * 3 + LB;
  constexpr int G = CisPairShape<LA, LB>::G, W = CisPairShape<LA, LB>::W;
  const long long items = (long long)(b.classStart[cls + 1] - b.classStart[cls]) * ((nv + G - 1) / G);
  if (items <= 0) return;
  hipLaunchKernelGGL((k_cis_pairs<LA, LB>), dim3((unsigned)((items + W - 1) / W)), dim3(64 * W), 0, s, t, b, ints, nv, DJ, DK,
                     matStride, cJ, cK, jvec, vecStride, F);
}

// One 4-wave workgroup per (atom, vector): the atom's diagonal block
//   F(A i, A j) = sum_{kl in A} [cJ (ij|kl) DJ(k,l) - cK (ik|jl) DK(l,k)] + cJ sum_{B != A} J_B->A(ij)
// with (ij|kl) = W(ij,kl) + 1/2 K(ij,kl) from the one-centre tables (W = (ij|kl) - 1/2 (ik|jl),
// K = (ik|jl), both stored by column: entry (ij,kl) at kl na^2 + ij), the partners' Coulomb
// partials summed in a fixed order.
constexpr int kAtomThreads = 256;
template <int MAXNA>
__global__ __launch_bounds__(kAtomThreads) void k_cis_atoms(KTables t, KBatch b, int first, const double* __restrict__ WK,
                                                            const double* __restrict__ DJ, const double* __restrict__ DK,
                                                            long long matStride, double cJ, double cK,
                                                            const double* __restrict__ jvec, long long vecStride,
                                                            double* __restrict__ F) {
  constexpr int MAXD = MAXNA == 9 ? 40 : 10, MAXNA2 = MAXNA * MAXNA, NW = kAtomThreads / 64;
  const int a = b.atomList[first + blockIdx.x], v = blockIdx.y, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const AtomDev A = b.atoms[a];
  const ElementDev& E = b.elems[A.elem];
  const MolDev m = b.mols[A.mol];
  const long long n = m.n;
  const int na = E.nao, dA = dimL(E.l), na2 = na * na;
  const int la = a - m.atom0, nat = m.natoms;
  double* Fv = F + v * matStride + m.matOff;
  if (na > MAXNA) {  // the host partitions the list by shell type; a wrong list shows as NaN
    if (tid == 0) Fv[(long long)A.ao * n + A.ao] = __builtin_nan("");
    return;
  }
  const double* Dj = DJ + v * matStride + m.matOff;
  const double* Dk = DK + v * matStride + m.matOff;
  __shared__ double sJ[MAXD], sDJ[MAXNA2], sDKT[MAXNA2], sAcc[NW][MAXD];
  __shared__ uint8_t sPairIdx[81];
  for (int k = tid; k < na2; k += kAtomThreads) {
    const int r = k / na, c = k - r * na;
    sDJ[k] = cJ != 0.0 ? Dj[(A.ao + c) * n + A.ao + r] : 0.0;  // DJ(k, l) at k na + l
    sDKT[k] = Dk[(A.ao + r) * n + A.ao + c];                      // DK(l, k) at k na + l
  }
  for (int k = tid; k < 81; k += kAtomThreads) sPairIdx[k] = t.pairIndex[k];
  if (cJ != 0.0) {
    double acc[MAXD];
#pragma unroll
    for (int Pp = 0; Pp < MAXD; ++Pp) acc[Pp] = 0.0;
    const double* jv = jvec + v * vecStride;
    for (int o = tid; o < nat; o += kAtomThreads) {
      if (o == la) continue;
      const int i = la < o ? la : o, j = la < o ? o : la;
      const long long pid = m.pair0 + (long long)i * nat - (long long)i * (i + 1) / 2 + (j - i - 1);
      const PairDev pr = b.pairs[pid];
      const int nO = b.atoms[m.atom0 + o].nao;  // the partner is the pair's first atom when o < la
      const long long off = la < o ? pr.vecOff : pr.vecOff + (nO == 1 ? 1 : nO == 4 ? 10 : 40);
#pragma unroll
      for (int Pp = 0; Pp < MAXD; ++Pp)
        if (Pp < dA) acc[Pp] += jv[off + Pp];
    }
#pragma unroll
    for (int Pp = 0; Pp < MAXD; ++Pp)
      if (Pp < dA) {
        const double s = waveSum(acc[Pp]);
        if (lane == 0) sAcc[w][Pp] = s;
      }
  }
  __syncthreads();
  if (tid < dA) {
    double s = 0.0;
    if (cJ != 0.0)
      for (int q = 0; q < NW; ++q) s += sAcc[q][tid];
    sJ[tid] = s;
  }
  __syncthreads();
  const double* __restrict__ Wc = b.oneCW + (long long)E.w1c * kOneCenterBlock;
  const double* __restrict__ Wx = WK + (long long)E.w1c * kOneCenterBlock;
  for (int o = tid; o < na2; o += kAtomThreads) {
    const int i = o / na, j = o - i * na;
    double g = 0.0;
#pragma unroll MAXNA
    for (int kl = 0; kl < na2; ++kl) {
      const double x = Wx[kl * na2 + o];
      g += cJ * (Wc[kl * na2 + o] + 0.5 * x) * sDJ[kl] - cK * x * sDKT[kl];
    }
    const int Pp = sPairIdx[i * 9 + j];
    if (Pp != kNoPair) g += cJ * sJ[Pp];
    Fv[(long long)(A.ao + j) * n + A.ao + i] = g;
  }
}

// S = dE .* X (nconf x nv, column-major), the orbital-energy part of the sigma vectors
__global__ __launch_bounds__(256) void k_cis_diag(long long nconf, long long total, const double* __restrict__ dE,
                                                  const double* __restrict__ X, double* __restrict__ S) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < total) S[i] = dE[i % nconf] * X[i];
}

}  // namespace

void launchCisFock(const KTables& t, const KBatch& b, const double* ints, const double* WK, int nv, const double* DJ,
                   const double* DK, long long matStride, double cJ, double cK, double* jvec, long long vecStride,
                   double* F, hipStream_t s) {
  if (nv <= 0) return;
  launchCisClass<0, 0>(t, b, ints, nv, DJ, DK, matStride, cJ, cK, jvec, vecStride, F, s);
  launchCisClass<0, 1>(t, b, ints, nv, DJ, DK, matStride, cJ, cK, jvec, vecStride, F, s);
  launchCisClass<0, 2>(t, b, ints, nv, DJ, DK, matStride, cJ, cK, jvec, vecStride, F, s);
  launchCisClass<1, 0>(t, b, ints, nv, DJ, DK, matStride, cJ, cK, jvec, vecStride, F, s);
  launchCisClass<1, 1>(t, b, ints, nv, DJ, DK, matStride, cJ, cK, jvec, vecStride, F, s);
  launchCisClass<1, 2>(t, b, ints, nv, DJ, DK, matStride, cJ, cK, jvec, vecStride, F, s);
  launchCisClass<2, 0>(t, b, ints, nv, DJ, DK, matStride, cJ, cK, jvec, vecStride, F, s);
  launchCisClass<2, 1>(t, b, ints, nv, DJ, DK, matStride, cJ, cK, jvec, vecStride, F, s);
  launchCisClass<2, 2>(t, b, ints, nv, DJ, DK, matStride, cJ, cK, jvec, vecStride, F, s);
  const int nSp = b.nAtomListSp, nD = b.nAtomList - b.nAtomListSp;
  if (nSp > 0)
    hipLaunchKernelGGL(k_cis_atoms<4>, dim3(nSp, nv), dim3(kAtomThreads), 0, s, t, b, 0, WK, DJ, DK, matStride, cJ, cK, jvec,
                       vecStride, F);
  if (nD > 0)
    hipLaunchKernelGGL(k_cis_atoms<9>, dim3(nD, nv), dim3(kAtomThreads), 0, s, t, b, nSp, WK, DJ, DK, matStride, cJ, cK,
                       jvec, vecStride, F);
}

void launchCisDiag(long long nconf, int nv, const double* dE, const double* X, double* S, hipStream_t s) {
  const long long total = nconf * nv;
  if (total <= 0) return;
  hipLaunchKernelGGL(k_cis_diag, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, nconf, total, dE, X, S);
}

}  // namespace sparrow_amd
