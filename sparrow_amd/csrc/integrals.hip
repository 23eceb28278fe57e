// Density-independent NDDO quantities on gfx950: two-centre two-electron integral blocks,
// VuvB core-attraction vectors, overlap, one-electron matrix H and core-core repulsion.
//
// Two-centre blocks (reference: IE/TwoCenterIntegralContainer.cpp:52-75 ->
// IE/Global2c2eMatrix.cpp:22-129 -> IE/Local2c2eMatrix.cpp -> IE/Local2c2eIntegralCalculator.cpp):
//   one 64-lane wave per atom pair, templated on (lA, lB) so that each class launch reserves
//   exactly its LDS.  Per pair:
//     1. point-charge sums of every multipole-multipole instance of the class (1/sqrt each),
//     2. local integrals L(p,q) = sum_k f_k * instance_k, with (xy|xy) = 1/2[(xx|xx)-(xx|yy)],
//     3. X = L * T^T and 4. G = T * X, where T(P,p) = r(mu,a) r(nu,b) + r(mu,b) r(nu,a)
//        is the charge-distribution rotation built from the 9x9 orbital rotation r
//        (IE/OrbitalRotation.cpp:77-154).  T only couples pairs of equal (l1,l2), so every output
//        sums over a block of at most 14 members.
//   The result is the d_A x d_B block (row-major) of the global integrals, identical to the
//   reference's term-list evaluation (IE/Global2c2eTerms.cpp:20-90) up to summation order.
#include "kernels.hpp"

namespace sparrow_amd {

namespace {

constexpr int kInstCap = 160;  // >= max instances of any class (151 for d-d)

__host__ __device__ constexpr int dimOfL(int l) { return l == 0 ? 1 : l == 1 ? 10 : 40; }

// Orbital rotation coefficients r[global*9 + local] for the axis Rab (IE/OrbitalRotation.cpp).
__device__ void fillRotation(double X, double Y, double Z, double* r) {
  for (int k = 0; k < 81; ++k) r[k] = 0.0;
  double R = sqrt(X * X + Y * Y + Z * Z);
  double pp = sqrt(X * X + Y * Y);
  double ca = 1.0, sa = 0.0;
  if (pp != 0.0) {
    ca = X / pp;
    sa = Y / pp;
  }
  double cb = Z / R, sb = pp / R;
  double s2b = sb * sb, c2b = cb * cb, saca = sa * ca, sbcb = sb * cb, dc = 2 * ca * ca - 1;
  const double r3 = 1.7320508075688772;
  enum { S = 0, Xo, Yo, Zo, X2Y2, XZ, Z2, YZ, XY };
  auto at = [&](int g, int l) -> double& { return r[g * 9 + l]; };
  at(S, S) = 1.0;
  at(Xo, Xo) = ca * cb;
  at(Yo, Xo) = sa * cb;
  at(Zo, Xo) = -sb;
  at(Xo, Yo) = -sa;
  at(Yo, Yo) = ca;
  at(Xo, Zo) = ca * sb;
  at(Yo, Zo) = sa * sb;
  at(Zo, Zo) = cb;
  at(X2Y2, X2Y2) = dc * c2b + 0.5 * dc * s2b;
  at(X2Y2, XZ) = dc * sbcb;
  at(X2Y2, Z2) = 0.5 * r3 * dc * s2b;
  at(X2Y2, YZ) = -2 * saca * sb;
  at(X2Y2, XY) = -2 * saca * cb;
  at(XZ, X2Y2) = -ca * sbcb;
  at(XZ, XZ) = ca * (2 * c2b - 1);
  at(XZ, Z2) = r3 * ca * sbcb;
  at(XZ, YZ) = -sa * cb;
  at(XZ, XY) = sa * sb;
  at(Z2, X2Y2) = 0.5 * r3 * s2b;
  at(Z2, XZ) = -r3 * sbcb;
  at(Z2, Z2) = c2b - 0.5 * s2b;
  at(YZ, X2Y2) = -sa * sbcb;
  at(YZ, XZ) = sa * (2 * c2b - 1);
  at(YZ, Z2) = r3 * sa * sbcb;
  at(YZ, YZ) = ca * cb;
  at(YZ, XY) = -ca * sb;
  at(XY, X2Y2) = 2 * saca * c2b + saca * s2b;
  at(XY, XZ) = 2 * saca * sbcb;
  at(XY, Z2) = r3 * saca * s2b;
  at(XY, YZ) = dc * sb;
  at(XY, XY) = dc * cb;
}

// T(P, q) for global pair P = (mu, nu) and local pair q = (a, b)
__device__ inline double pairRot(const double* r, const uint8_t* pairOrb, int P, int q) {
  int mu = pairOrb[2 * P], nu = pairOrb[2 * P + 1], a = pairOrb[2 * q], b = pairOrb[2 * q + 1];
  double v = r[mu * 9 + a] * r[nu * 9 + b];
  if (a != b) v += r[mu * 9 + b] * r[nu * 9 + a];
  return v;
}

// Point-charge sum of one multipole-multipole instance.
__device__ double evalInstance(const KTables& t, int id, double R, const double* DA, const double* rhoA, const double* DB,
                               const double* rhoB) {
  const MMInstanceDev in = t.inst[id];
  double d1 = in.p1 < kSs0 ? DA[in.p1] : 0.0;
  double d2 = in.p2 < kSs0 ? DB[in.p2] : 0.0;
  double rs = rhoA[in.p1] + rhoB[in.p2];
  double rs2 = rs * rs;
  double sum = 0.0;
  for (int c = 0; c < in.nCharge; ++c) {
    const ChargePairDev q = t.charges[in.charge0 + c];
    double dx = d2 * q.x2 - d1 * q.x1;
    double dy = d2 * q.y2 - d1 * q.y1;
    double dz = R + d2 * q.z2 - d1 * q.z1;
    sum += q.q / sqrt(dx * dx + dy * dy + dz * dz + rs2);
  }
  return sum;
}

template <int LA, int LB>
struct TwoCenterCfg {
  static constexpr int DA = dimOfL(LA), DB = dimOfL(LB);
  static constexpr int W = (DA * DB >= 1600) ? 1 : (DA * DB >= 400 ? 2 : 4);  // waves per block (LDS <= ~52 KB)
};

// Per-wave LDS of the two-centre evaluation of one pair (W waves per block).
template <int LA, int LB>
struct TwoCenterLds {
  static constexpr int W = TwoCenterCfg<LA, LB>::W, DA = TwoCenterCfg<LA, LB>::DA, DB = TwoCenterCfg<LA, LB>::DB;
  double rot[W][81], rotRev[W][81];
  double L[W][DA * DB], X[W][DA * DB];
  double inst[W][kInstCap];
  double VA[W][DA], VB[W][DB];        // local-frame VuvB vectors (scratch)
  double vuvbA[W][DA], vuvbB[W][DB];  // VuvB vectors in the global frame (result)
  double par[W][2][16];               // D[5], rho[8], pcore
};

// Two-centre block of one pair per active wave (all threads of the block call it): atom B at
// (X, Y, Z) from atom A.  On return L[wi] holds the global block G(P, Q) (P on A, Q on B,
// d_A x d_B) and vuvbA[wi] / vuvbB[wi] the VuvB vectors -Z_B (P | s_B s_B), -Z_A (s_A s_A | Q).
template <int LA, int LB>
__device__ void twoCenterPair(const KTables& t, const ElementDev* eA, const ElementDev* eB, double X, double Y, double Z,
                              bool active, int wi, int lane, TwoCenterLds<LA, LB>& S) {
  using Cfg = TwoCenterCfg<LA, LB>;
  constexpr int DA = Cfg::DA, DB = Cfg::DB;
  constexpr int cls = LA * 3 + LB;
  double R = 1;
  if (active) {
    R = sqrt(X * X + Y * Y + Z * Z);
    if (lane == 0) fillRotation(X, Y, Z, S.rot[wi]);
    if (lane == 1) fillRotation(-X, -Y, -Z, S.rotRev[wi]);
    if (lane < 5) {
      S.par[wi][0][lane] = eA->D[lane];
      S.par[wi][1][lane] = eB->D[lane];
    }
    if (lane < 8) {
      S.par[wi][0][5 + lane] = eA->rho[lane];
      S.par[wi][1][5 + lane] = eB->rho[lane];
    }
    for (int k = lane; k < DA * DB; k += 64) S.L[wi][k] = 0.0;
  }
  __syncthreads();
  // 1. multipole-multipole instances of class (LA, LB)
  const LocalClassDev C = t.classes[cls];
  if (active)
    for (int i = lane; i < C.nInst; i += 64)
      S.inst[wi][i] = evalInstance(t, C.inst0 + i, R, &S.par[wi][0][0], &S.par[wi][0][5], &S.par[wi][1][0], &S.par[wi][1][5]);
  __syncthreads();
  // 2. local integrals
  if (active)
    for (int e = lane; e < C.nEntry; e += 64) {
      const EntryDev E = t.entries[C.entry0 + e];
      double v = 0.0;
      for (int k = 0; k < E.nTerm; ++k) {
        const EntryTermDev T = t.eterms[E.term0 + k];
        v += T.f * S.inst[wi][T.inst];
      }
      S.L[wi][E.i * DB + E.j] = v;
    }
  __syncthreads();
  if (LA > 0 && LB > 0 && active && lane == 0) S.L[wi][4 * DB + 4] = 0.5 * (S.L[wi][2 * DB + 2] - S.L[wi][2 * DB + 5]);
  __syncthreads();
  // 3. X(p, Q) = sum_q L(p, q) T(Q, q)
  if (active)
    for (int o = lane; o < DA * DB; o += 64) {
      int p = o / DB, Q = o - p * DB;
      int blk = t.pairBlock[Q], m0 = t.blockStart[blk], ml = t.blockLen[blk];
      double v = 0.0;
      for (int k = 0; k < ml; ++k) {
        int q = t.blockMembers[m0 + k];
        v += S.L[wi][p * DB + q] * pairRot(S.rot[wi], t.pairOrb, Q, q);
      }
      S.X[wi][o] = v;
    }
  __syncthreads();
  // 4. G(P, Q) = sum_p T(P, p) X(p, Q)
  if (active)
    for (int o = lane; o < DA * DB; o += 64) {
      int P = o / DB, Q = o - P * DB;
      int blk = t.pairBlock[P], m0 = t.blockStart[blk], ml = t.blockLen[blk];
      double v = 0.0;
      for (int k = 0; k < ml; ++k) {
        int p = t.blockMembers[m0 + k];
        v += pairRot(S.rot[wi], t.pairOrb, P, p) * S.X[wi][p * DB + Q];
      }
      S.L[wi][o] = v;
    }
  __syncthreads();
  // 5. VuvB (IE/VuvB.cpp:17-31): V_A(P) = -Z_B (P | s_B s_B) with rho_B = pcore_B, D_B = 0.
  //    Uses the (LA,0) local class; for B the reversed axis and the (LB,0) class.
  double ZA = 0, ZB = 0;
  int pcA = 0, pcB = 0;
  double pcoreA = 0, pcoreB = 0;
  if (active) {
    ZA = eA->core;
    ZB = eB->core;
    pcA = eA->pcoreSpecified;
    pcB = eB->pcoreSpecified;
    pcoreA = eA->pcore;
    pcoreB = eB->pcore;
  }
  const LocalClassDev CA = t.classes[LA * 3];
  const LocalClassDev CB = t.classes[LB * 3];
  double coreD[5] = {0, 0, 0, 0, 0};
  double coreRhoA[8] = {0, 0, 0, 0, 0, pcoreA, 0, 0};
  double coreRhoB[8] = {0, 0, 0, 0, 0, pcoreB, 0, 0};
  // instances for A (first CA.nInst slots) and for B (following CB.nInst slots)
  if (active) {
    for (int i = lane; i < CA.nInst; i += 64)
      S.inst[wi][i] = evalInstance(t, CA.inst0 + i, R, &S.par[wi][0][0], &S.par[wi][0][5], coreD, coreRhoB);
    for (int i = lane; i < CB.nInst; i += 64)
      S.inst[wi][CA.nInst + i] = evalInstance(t, CB.inst0 + i, R, &S.par[wi][1][0], &S.par[wi][1][5], coreD, coreRhoA);
    for (int k = lane; k < DA; k += 64) S.VA[wi][k] = 0.0;
    for (int k = lane; k < DB; k += 64) S.VB[wi][k] = 0.0;
  }
  __syncthreads();
  if (active) {
    for (int e = lane; e < CA.nEntry; e += 64) {
      const EntryDev E = t.entries[CA.entry0 + e];
      double v = 0.0;
      for (int k = 0; k < E.nTerm; ++k) {
        const EntryTermDev T = t.eterms[E.term0 + k];
        v += T.f * S.inst[wi][T.inst];
      }
      S.VA[wi][E.i] = v;
    }
    for (int e = lane; e < CB.nEntry; e += 64) {
      const EntryDev E = t.entries[CB.entry0 + e];
      double v = 0.0;
      for (int k = 0; k < E.nTerm; ++k) {
        const EntryTermDev T = t.eterms[E.term0 + k];
        v += T.f * S.inst[wi][CA.nInst + T.inst];
      }
      S.VB[wi][E.i] = v;
    }
  }
  __syncthreads();
  // the VuvB vectors in the global frame
  if (active) {
    for (int P = lane; P < DA; P += 64) {
      double v;
      if (pcB) {
        int blk = t.pairBlock[P], m0 = t.blockStart[blk], ml = t.blockLen[blk];
        v = 0.0;
        for (int k = 0; k < ml; ++k) {
          int p = t.blockMembers[m0 + k];
          v += pairRot(S.rot[wi], t.pairOrb, P, p) * S.VA[wi][p];
        }
      } else {
        v = S.L[wi][P * DB];  // (P | s_B s_B) from the two-centre block
      }
      S.vuvbA[wi][P] = -ZB * v;
    }
    for (int Q = lane; Q < DB; Q += 64) {
      double v;
      if (pcA) {
        int blk = t.pairBlock[Q], m0 = t.blockStart[blk], ml = t.blockLen[blk];
        v = 0.0;
        for (int k = 0; k < ml; ++k) {
          int q = t.blockMembers[m0 + k];
          v += pairRot(S.rotRev[wi], t.pairOrb, Q, q) * S.VB[wi][q];
        }
      } else {
        v = S.L[wi][Q];  // (s_A s_A | Q)
      }
      S.vuvbB[wi][Q] = -ZA * v;
    }
  }
  __syncthreads();
}

template <int LA, int LB>
__global__ __launch_bounds__(256) void k_twocenter(KTables t, KBatch b, double* __restrict__ ints, double* __restrict__ vuvb) {
  using Cfg = TwoCenterCfg<LA, LB>;
  constexpr int DA = Cfg::DA, DB = Cfg::DB, W = Cfg::W;
  constexpr int cls = LA * 3 + LB;
  __shared__ TwoCenterLds<LA, LB> S;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int idx = b.classStart[cls] + blockIdx.x * W + w;
  const bool active = (w < W) && idx < b.classStart[cls + 1];
  const int wi = w < W ? w : 0;
  PairDev pr{};
  ElementDev const* eA = nullptr;
  ElementDev const* eB = nullptr;
  double X = 0, Y = 0, Z = 0;
  if (active) {
    pr = b.pairs[b.classPairs[idx]];
    const AtomDev A = b.atoms[pr.a], B = b.atoms[pr.b];
    eA = &b.elems[A.elem];
    eB = &b.elems[B.elem];
    X = B.x - A.x;
    Y = B.y - A.y;
    Z = B.z - A.z;
  }
  twoCenterPair<LA, LB>(t, eA, eB, X, Y, Z, active, wi, lane, S);
  if (active) {
    for (int o = lane; o < DA * DB; o += 64) ints[pr.intOff + o] = S.L[wi][o];
    for (int P = lane; P < DA; P += 64) vuvb[pr.vecOff + P] = S.vuvbA[wi][P];
    for (int Q = lane; Q < DB; Q += 64) vuvb[pr.vecOff + DA + Q] = S.vuvbB[wi][Q];
  }
}

template <int LA, int LB>
void launchClass(const KTables& t, const KBatch& b, double* ints, double* vuvb, hipStream_t s) {
  constexpr int cls = LA * 3 + LB;
  int n = b.classStart[cls + 1] - b.classStart[cls];
  if (n <= 0) return;
  constexpr int W = TwoCenterCfg<LA, LB>::W;
  int blocks = (n + W - 1) / W;
  hipLaunchKernelGGL((k_twocenter<LA, LB>), dim3(blocks), dim3(64 * W), 0, s, t, b, ints, vuvb);
}

// ---------------------------------------------------------------------------------------------
// Overlap (IE/GTOOverlapMatrixBlock.cpp:49-165): STO-6G shells, Obara-Saika per Cartesian
// direction, six Cartesian d folded into five spherical d (xy-type normalisation).  One thread
// per (pair, shell A, shell B).  The block is computed, like the reference, for the higher atom
// (rows) against the lower atom (columns) with R = R_lower - R_higher, and stored with its
// transpose.  H off-diagonal = 1/2 (beta_mu + beta_nu) S_mu,nu (NU/OneElectronMatrix.cpp:125-134).
// ---------------------------------------------------------------------------------------------
__device__ void shellPairOverlap(const ElementDev& EA, int sa, const ElementDev& EB, int sb, double x, double y, double z,
                                 double out[5][5]) {
  const int la = EA.shellL[sa], lb = EB.shellL[sb];
  const int nA = la == 0 ? 1 : la == 1 ? 3 : 6, nB = lb == 0 ? 1 : lb == 1 ? 3 : 6;
  const int startA = la == 0 ? 0 : la == 1 ? 1 : 4, startB = lb == 0 ? 0 : lb == 1 ? 1 : 4;
  // Cartesian exponents: s | x y z | xz yz xy zz yy xx
  const int mx[10] = {0, 1, 0, 0, 1, 0, 1, 0, 0, 2};
  const int my[10] = {0, 0, 1, 0, 0, 1, 1, 0, 2, 0};
  const int mz[10] = {0, 0, 0, 1, 1, 1, 0, 2, 0, 0};
  double res[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) res[i][j] = 0.0;
  const double pi = 3.141592653589793;
  const double rr[3] = {x, y, z};
  const double r2 = x * x + y * y + z * z;
  for (int ia = 0; ia < 6; ++ia)
    for (int ib = 0; ib < 6; ++ib) {
      double ea = EA.shE[sa][ia], eb = EB.shE[sb][ib], es = ea + eb;
      double fp = eb / es;
      double KD = exp(r2 * (-fp * ea)) * (EA.shC[sa][ia] * EB.shC[sb][ib]);
      double Sd[3][3][3];
      for (int d = 0; d < 3; ++d) {
        double p = rr[d] * fp, dist2 = p - rr[d];
        Sd[d][0][0] = sqrt(pi / es);
        for (int k = 0; k <= la; ++k)
          for (int l = 0; l <= lb; ++l) {
            if (k == 0 && l == 0) continue;
            double v = 0.0;
            if (l == 0) {
              if (k > 1) v = Sd[d][k - 2][0] * ((k - 1.0) / (2 * es));
              v += Sd[d][k - 1][0] * p;
            } else {
              if (l > 1) v = Sd[d][k][l - 2] * ((l - 1.0) / (2 * es));
              if (k > 0) v += Sd[d][k - 1][l - 1] * (k / (2 * es));
              v += Sd[d][k][l - 1] * dist2;
            }
            Sd[d][k][l] = v;
          }
      }
      for (int k = 0; k < nA; ++k)
        for (int l = 0; l < nB; ++l) {
          int ca = k + startA, cbb = l + startB;
          res[k][l] += KD * (Sd[0][mx[ca]][mx[cbb]] * Sd[1][my[ca]][my[cbb]] * Sd[2][mz[ca]][mz[cbb]]);
        }
    }
  const double r3 = 1.7320508075688772;
  if (nA == 6)
    for (int l = 0; l < nB; ++l) {
      res[3][l] = (res[3][l] - 0.5 * res[4][l] - 0.5 * res[5][l]) / r3;
      res[4][l] = 0.5 * (res[5][l] - res[4][l]);
    }
  if (nB == 6)
    for (int k = 0; k < nA; ++k) {
      res[k][3] = (res[k][3] - 0.5 * res[k][4] - 0.5 * res[k][5]) / r3;
      res[k][4] = 0.5 * (res[k][5] - res[k][4]);
    }
  const int dPos[5] = {1, 3, 4, 2, 0};  // xz yz xy z2 x2y2 -> AO slot x2y2 xz z2 yz xy
  for (int k = 0; k < 2 * la + 1; ++k)
    for (int l = 0; l < 2 * lb + 1; ++l) out[la == 2 ? dPos[k] : k][lb == 2 ? dPos[l] : l] = res[k][l];
}

// shellPairOverlap with the two shells' angular momenta fixed at compile time (every loop unrolled,
// the recurrences in registers)
template <int LA, int LB>
__device__ inline void shellPairOverlapT(const ElementDev& EA, int sa, const ElementDev& EB, int sb, double x, double y,
                                         double z, double out[5][5]) {
  constexpr int nA = LA == 0 ? 1 : LA == 1 ? 3 : 6, nB = LB == 0 ? 1 : LB == 1 ? 3 : 6;
  constexpr int startA = LA == 0 ? 0 : LA == 1 ? 1 : 4, startB = LB == 0 ? 0 : LB == 1 ? 1 : 4;
  constexpr int mx[10] = {0, 1, 0, 0, 1, 0, 1, 0, 0, 2};
  constexpr int my[10] = {0, 0, 1, 0, 0, 1, 1, 0, 2, 0};
  constexpr int mz[10] = {0, 0, 0, 1, 1, 1, 0, 2, 0, 0};
  double res[nA][nB];
#pragma unroll
  for (int i = 0; i < nA; ++i)
#pragma unroll
    for (int j = 0; j < nB; ++j) res[i][j] = 0.0;
  const double pi = 3.141592653589793;
  const double rr[3] = {x, y, z};
  const double r2 = x * x + y * y + z * z;
#pragma unroll 1
  for (int ia = 0; ia < 6; ++ia)
#pragma unroll 1
    for (int ib = 0; ib < 6; ++ib) {
      const double ea = EA.shE[sa][ia], eb = EB.shE[sb][ib], es = ea + eb;
      const double fp = eb / es;
      const double KD = exp(r2 * (-fp * ea)) * (EA.shC[sa][ia] * EB.shC[sb][ib]);
      double Sd[3][LA + 1][LB + 1];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double p = rr[d] * fp, dist2 = p - rr[d];
        Sd[d][0][0] = sqrt(pi / es);
#pragma unroll
        for (int k = 0; k <= LA; ++k)
#pragma unroll
          for (int l = 0; l <= LB; ++l) {
            if (k == 0 && l == 0) continue;
            double v = 0.0;
            if (l == 0) {
              if (k > 1) v = Sd[d][k - 2][0] * ((k - 1.0) / (2 * es));
              v += Sd[d][k - 1][0] * p;
            } else {
              if (l > 1) v = Sd[d][k][l - 2] * ((l - 1.0) / (2 * es));
              if (k > 0) v += Sd[d][k - 1][l - 1] * (k / (2 * es));
              v += Sd[d][k][l - 1] * dist2;
            }
            Sd[d][k][l] = v;
          }
      }
#pragma unroll
      for (int k = 0; k < nA; ++k)
#pragma unroll
        for (int l = 0; l < nB; ++l) {
          constexpr int dummy = 0;
          (void)dummy;
          const int ca = k + startA, cbb = l + startB;
          res[k][l] += KD * (Sd[0][mx[ca]][mx[cbb]] * Sd[1][my[ca]][my[cbb]] * Sd[2][mz[ca]][mz[cbb]]);
        }
    }
  const double r3 = 1.7320508075688772;
  if constexpr (nA == 6) {
#pragma unroll
    for (int l = 0; l < nB; ++l) {
      res[3][l] = (res[3][l] - 0.5 * res[4][l] - 0.5 * res[5][l]) / r3;
      res[4][l] = 0.5 * (res[5][l] - res[4][l]);
    }
  }
  if constexpr (nB == 6) {
#pragma unroll
    for (int k = 0; k < nA; ++k) {
      res[k][3] = (res[k][3] - 0.5 * res[k][4] - 0.5 * res[k][5]) / r3;
      res[k][4] = 0.5 * (res[k][5] - res[k][4]);
    }
  }
  constexpr int dPos[5] = {1, 3, 4, 2, 0};  // xz yz xy z2 x2y2 -> AO slot x2y2 xz z2 yz xy
#pragma unroll
  for (int k = 0; k < 2 * LA + 1; ++k)
#pragma unroll
    for (int l = 0; l < 2 * LB + 1; ++l) out[LA == 2 ? dPos[k] : k][LB == 2 ? dPos[l] : l] = res[k][l];
}

// shellPairOverlap split over the lanes of a group: this lane's share (primitive pairs
// ia * 6 + ib = part mod nparts) of the Cartesian block (rows A functions, columns B, in the
// Cartesian order s | x y z | xz yz xy zz yy xx), then the group's sum and the spherical block
// (cartesianToSpherical) on every lane.  All lanes of the wave call it.
__device__ void shellPairOverlapPart(const ElementDev& EA, int sa, const ElementDev& EB, int sb, double x, double y, double z,
                                     int part, int nparts, bool valid, double res[6][6]) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) res[i][j] = 0.0;
  if (!valid) return;
  const int la = EA.shellL[sa], lb = EB.shellL[sb];
  const int nA = la == 0 ? 1 : la == 1 ? 3 : 6, nB = lb == 0 ? 1 : lb == 1 ? 3 : 6;
  const int startA = la == 0 ? 0 : la == 1 ? 1 : 4, startB = lb == 0 ? 0 : lb == 1 ? 1 : 4;
  const int mx[10] = {0, 1, 0, 0, 1, 0, 1, 0, 0, 2};
  const int my[10] = {0, 0, 1, 0, 0, 1, 1, 0, 2, 0};
  const int mz[10] = {0, 0, 0, 1, 1, 1, 0, 2, 0, 0};
  const double pi = 3.141592653589793;
  const double rr[3] = {x, y, z};
  const double r2 = x * x + y * y + z * z;
  for (int q = part; q < 36; q += nparts) {
    const int ia = q / 6, ib = q - ia * 6;
    double ea = EA.shE[sa][ia], eb = EB.shE[sb][ib], es = ea + eb;
    double fp = eb / es;
    double KD = exp(r2 * (-fp * ea)) * (EA.shC[sa][ia] * EB.shC[sb][ib]);
    double Sd[3][3][3];
    for (int d = 0; d < 3; ++d) {
      double p = rr[d] * fp, dist2 = p - rr[d];
      Sd[d][0][0] = sqrt(pi / es);
      for (int k = 0; k <= la; ++k)
        for (int l = 0; l <= lb; ++l) {
          if (k == 0 && l == 0) continue;
          double v = 0.0;
          if (l == 0) {
            if (k > 1) v = Sd[d][k - 2][0] * ((k - 1.0) / (2 * es));
            v += Sd[d][k - 1][0] * p;
          } else {
            if (l > 1) v = Sd[d][k][l - 2] * ((l - 1.0) / (2 * es));
            if (k > 0) v += Sd[d][k - 1][l - 1] * (k / (2 * es));
            v += Sd[d][k][l - 1] * dist2;
          }
          Sd[d][k][l] = v;
        }
    }
    for (int k = 0; k < nA; ++k)
      for (int l = 0; l < nB; ++l) {
        int ca = k + startA, cbb = l + startB;
        res[k][l] += KD * (Sd[0][mx[ca]][mx[cbb]] * Sd[1][my[ca]][my[cbb]] * Sd[2][mz[ca]][mz[cbb]]);
      }
  }
}

// the five spherical d (slot order x2-y2 xz z2 yz xy) from the six Cartesian ones, as in
// shellPairOverlap
__device__ void cartesianToSpherical(int la, int lb, double res[6][6], double out[5][5]) {
  const int nA = la == 0 ? 1 : la == 1 ? 3 : 6, nB = lb == 0 ? 1 : lb == 1 ? 3 : 6;
  const double r3 = 1.7320508075688772;
  if (nA == 6)
    for (int l = 0; l < nB; ++l) {
      res[3][l] = (res[3][l] - 0.5 * res[4][l] - 0.5 * res[5][l]) / r3;
      res[4][l] = 0.5 * (res[5][l] - res[4][l]);
    }
  if (nB == 6)
    for (int k = 0; k < nA; ++k) {
      res[k][3] = (res[k][3] - 0.5 * res[k][4] - 0.5 * res[k][5]) / r3;
      res[k][4] = 0.5 * (res[k][5] - res[k][4]);
    }
  const int dPos[5] = {1, 3, 4, 2, 0};
  for (int k = 0; k < 2 * la + 1; ++k)
    for (int l = 0; l < 2 * lb + 1; ++l) out[la == 2 ? dPos[k] : k][lb == 2 ? dPos[l] : l] = res[k][l];
}

// PM6 (N/Pm6/PM6PairwiseRepulsion.h:76-175) and MNDO/AM1-type (N/Mndo/MNDOPairwiseRepulsion.h:62-125,
// N/Am1/AM1PairwiseRepulsion.h:64-121) core-core repulsion of one pair, Hartree.
__device__ double pairRepulsion(const KBatch& b, int ea, int eb, double R) {
  const ElementDev& A = b.elems[ea];
  const ElementDev& B = b.elems[eb];
  const double ps = A.pcore + B.pcore;
  const double ssss = 1.0 / sqrt(R * R + ps * ps);
  if (b.methodKind == 4) {
    const DiatomicDev d = b.diatomic[ea * b.nElem + eb];
    const double expCoef = 0.00001244878365801758178693335;  // 3e-4 A^-5 in bohr^-5
    const double R6 = R * R * R * R * R * R;
    double standard = 1.0 + 2 * d.x * exp(-d.alpha * (R + expCoef * R6));
    double par;
    if (d.kind == 1) par = 1.0 + 2 * d.x * exp(-d.alpha * (R * R));
    else if (d.kind == 2) par = standard + 9.28 * exp(-3.1644797213016 * R);
    else if (d.kind == 3) {
      double u = R - 5.480205761238679760340424;
      par = standard - 0.0007 * exp(-u * u);
    } else par = standard;
    double v = (A.zcbrt + B.zcbrt) / R;
    double v2 = v * v, v6 = v2 * v2 * v2;
    double additional = 0.0000207389327971913490822259 * v6 * v6 / 27.21138602;  // 1e-8 A^12 in bohr^12, eV->Eh
    double ga = 0.0, gb = 0.0;
    if (A.nGauss > 0) ga = A.gauss[0][0] * exp(-A.gauss[0][1] * (R - A.gauss[0][2]) * (R - A.gauss[0][2]));
    if (B.nGauss > 0) gb = B.gauss[0][0] * exp(-B.gauss[0][1] * (R - B.gauss[0][2]) * (R - B.gauss[0][2]));
    return A.core * B.core * ssss * par + additional + (ga + gb) / R * (A.core * B.core);
  }
  const double aw = 0.52917721067;  // Angstrom per bohr
  double par;
  if (A.Z == 1 && (B.Z == 7 || B.Z == 8)) par = 1.0 + exp(-A.alpha * R) + R * aw * exp(-B.alpha * R);
  else if (B.Z == 1 && (A.Z == 7 || A.Z == 8)) par = 1.0 + R * aw * exp(-A.alpha * R) + exp(-B.alpha * R);
  else par = 1.0 + exp(-A.alpha * R) + exp(-B.alpha * R);
  double e = A.core * B.core * ssss * par;
  if (b.methodKind != 0) {  // AM1, PM3, RM1: all Gaussians
    double g = 0.0;
    for (int k = 0; k < A.nGauss; ++k) g += A.gauss[k][0] * exp(-A.gauss[k][1] * (R - A.gauss[k][2]) * (R - A.gauss[k][2]));
    for (int k = 0; k < B.nGauss; ++k) g += B.gauss[k][0] * exp(-B.gauss[k][1] * (R - B.gauss[k][2]) * (R - B.gauss[k][2]));
    e += g / R * (A.core * B.core);
  }
  return e;
}

// Overlap S, off-diagonal H = (beta_mu + beta_nu)/2 S and the core repulsion of the pairs of one
// (l_A, l_B) class: thread t takes shell pair t / count (shell pair sp: shell sp / (LO + 1) of the
// pair's second atom, sp % (LO + 1) of its first; shell s has angular momentum s) of pair t % count,
// so a wave's threads evaluate the same shell-pair type.  The blocks go to the lower triangle
// (the second atom has the higher AOs); k_mirror_lower fills the upper one.
template <int HI, int LO>
__device__ inline void pairOverlapBlock(const KBatch& b, const ElementDev& EH, int sh, const ElementDev& EL, int sl,
                                        const AtomDev& hi, const AtomDev& lo, double x, double y, double z, double* S,
                                        double* H) {
  constexpr int offH = HI == 0 ? 0 : HI == 1 ? 1 : 4, offL = LO == 0 ? 0 : LO == 1 ? 1 : 4;
  if (offH >= hi.nao || offL >= lo.nao) return;
  double o[5][5];
  shellPairOverlapT<HI, LO>(EH, sh, EL, sl, x, y, z, o);
  const MolDev m = b.mols[hi.mol];
  const long long n = m.n;
  double* Sm = S + m.matOff;
  double* Hm = H + m.matOff;
#pragma unroll
  for (int k = 0; k < 2 * HI + 1; ++k)
#pragma unroll
    for (int l = 0; l < 2 * LO + 1; ++l) {
      const int ah = offH + k, al = offL + l;
      const long long r = hi.ao + ah, c = lo.ao + al;
      const double sv = o[k][l];
      const double bh = EH.beta[HI], bl = EL.beta[LO];
      Sm[c * n + r] = sv;
      Hm[c * n + r] = 0.5 * (bh + bl) * sv;
    }
}

template <int LA, int LB>
__global__ __launch_bounds__(256) void k_pair_overlap(KBatch b, double* __restrict__ S, double* __restrict__ H,
                                                      double* __restrict__ erep) {
  constexpr int cls = LA * 3 + LB, NSL = LA + 1, NSP = (LB + 1) * (LA + 1);  // LA: first atom, LB: second
  const int first = b.classStart[cls], count = b.classStart[cls + 1] - first;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int sp = (int)(gid / count), q = (int)(gid - (long long)sp * count);
  if (sp >= NSP) return;
  const int p = b.classPairs[first + q];
  const PairDev pr = b.pairs[p];
  const AtomDev lo = b.atoms[pr.a], hi = b.atoms[pr.b];
  const ElementDev& EL = b.elems[lo.elem];
  const ElementDev& EH = b.elems[hi.elem];
  const double x = lo.x - hi.x, y = lo.y - hi.y, z = lo.z - hi.z;
  if (sp == 0) erep[p] = pairRepulsion(b, hi.elem, lo.elem, sqrt(x * x + y * y + z * z));
  const int sh = sp / NSL, sl = sp - sh * NSL;  // shell on the second (higher) atom, on the first
  if (sh >= EH.nShell || sl >= EL.nShell) return;
  switch (sh * 3 + sl) {
    case 0: pairOverlapBlock<0, 0>(b, EH, sh, EL, sl, hi, lo, x, y, z, S, H); break;
    case 1: if constexpr (LA >= 1) pairOverlapBlock<0, 1>(b, EH, sh, EL, sl, hi, lo, x, y, z, S, H); break;
    case 2: if constexpr (LA >= 2) pairOverlapBlock<0, 2>(b, EH, sh, EL, sl, hi, lo, x, y, z, S, H); break;
    case 3: if constexpr (LB >= 1) pairOverlapBlock<1, 0>(b, EH, sh, EL, sl, hi, lo, x, y, z, S, H); break;
    case 4: if constexpr (LB >= 1 && LA >= 1) pairOverlapBlock<1, 1>(b, EH, sh, EL, sl, hi, lo, x, y, z, S, H); break;
    case 5: if constexpr (LB >= 1 && LA >= 2) pairOverlapBlock<1, 2>(b, EH, sh, EL, sl, hi, lo, x, y, z, S, H); break;
    case 6: if constexpr (LB >= 2) pairOverlapBlock<2, 0>(b, EH, sh, EL, sl, hi, lo, x, y, z, S, H); break;
    case 7: if constexpr (LB >= 2 && LA >= 1) pairOverlapBlock<2, 1>(b, EH, sh, EL, sl, hi, lo, x, y, z, S, H); break;
    default: if constexpr (LB >= 2 && LA >= 2) pairOverlapBlock<2, 2>(b, EH, sh, EL, sl, hi, lo, x, y, z, S, H); break;
  }
}

// Upper triangle of S and H (H may be null) from their lower triangle, 32 x 32 tiles on and above
// the diagonal of each listed molecule (grid-stride over a molecule's tiles): the mirrored tile is
// read down its columns into LDS and written down the columns of the upper tile.
__global__ __launch_bounds__(256) void k_mirror_lower(KBatch b, double* __restrict__ S, double* __restrict__ H) {
  const MolDev m = b.mols[b.molList[blockIdx.y]];
  const int n = m.n, tn = (n + 31) >> 5, ntiles = tn * (tn + 1) / 2;
  __shared__ double sT[2][32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows of 32 per pass
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int ti = 0, local = t;
    while (local >= tn - ti) {
      local -= tn - ti;
      ++ti;
    }
    const int tj = ti + local, r0 = ti * 32, c0 = tj * 32;
    __syncthreads();  // the previous tile's reads of sT are done
    for (int k = ty; k < 32; k += 8) {  // source column r0 + k, rows c0 + tx (lower triangle)
      const int sr = c0 + tx, sc = r0 + k;
      const bool in = sr < n && sc < n;
      sT[0][k][tx] = in ? S[m.matOff + (long long)sc * n + sr] : 0.0;
      if (H) sT[1][k][tx] = in ? H[m.matOff + (long long)sc * n + sr] : 0.0;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {  // destination column c0 + k, rows r0 + tx (upper triangle)
      const int r = r0 + tx, c = c0 + k;
      if (r < c && c < n) {
        S[m.matOff + (long long)c * n + r] = sT[0][tx][k];
        if (H) H[m.matOff + (long long)c * n + r] = sT[1][tx][k];
      }
    }
  }
}

// Diagonal blocks of H: U on the diagonal plus the VuvB vectors of every partner
// (NU/OneElectronMatrix.cpp:56-111).  One wave per atom; lane = charge distribution of the atom.
__global__ __launch_bounds__(64) void k_atom_hcore(KTables t, KBatch b, const double* __restrict__ vuvb, double* __restrict__ H) {
  const int a = b.atomList[blockIdx.x];
  const int lane = threadIdx.x;
  const AtomDev A = b.atoms[a];
  const ElementDev& E = b.elems[A.elem];
  const MolDev m = b.mols[A.mol];
  const int dA = dimOfL(E.l);
  const int la = a - m.atom0;
  const int nat = m.natoms;
  double acc = 0.0;
  if (lane < dA) {
    for (int o = 0; o < nat; ++o) {
      if (o == la) continue;
      int i = la < o ? la : o, j = la < o ? o : la;
      long long pid = m.pair0 + (long long)i * nat - (long long)i * (i + 1) / 2 + (j - i - 1);
      const PairDev pr = b.pairs[pid];
      long long off = la < o ? pr.vecOff + lane : pr.vecOff + dimOfL(b.elems[b.atoms[pr.a].elem].l) + lane;
      acc += vuvb[off];
    }
  }
  __shared__ double sV[40];
  if (lane < dA) sV[lane] = acc;
  __syncthreads();
  const long long n = m.n;
  double* Hm = H + m.matOff;
  const int na = E.nao;
  for (int k = lane; k < na * na; k += 64) {
    int i = k / na, j = k - i * na;
    int P = t.pairIndex[i * 9 + j];
    double v = P == kNoPair ? 0.0 : sV[P];
    if (i == j) v += i < 1 ? E.U[0] : i < 4 ? E.U[1] : E.U[2];
    Hm[(long long)(A.ao + j) * n + A.ao + i] = v;
  }
}

// ---------------------------------------------------------------------------------------------
// Nuclear gradient at a converged density (the SCF energy is stationary in P, so dE/dR is the
// derivative at fixed P).  Only the two-centre terms depend on the geometry; each pair's share
//   E_AB = 2 sum_{mu on B, nu on A} P_mu,nu H_mu,nu            (H_AB = 1/2 (beta_mu + beta_nu) S)
//        + sum_{mu,nu on A} P_mu,nu V_A(mu nu) + sum_{la,si on B} P_la,si V_B(la si)   (VuvB)
//        + sum_{mu,nu on A} sum_{la,si on B} (mu nu|la si) (P_mu,nu P_la,si - 1/2 P_mu,la P_nu,si)
//        + E_rep(A, B)
// and, for a spin-polarised density (P = P^a + P^b, spin density Ps = P^a - P^b), the exchange
// with each spin's own density, - sum (mu nu|la si) (P^a_mu,la P^a_nu,si + P^b_mu,la P^b_nu,si),
// = the term above - 1/2 sum (mu nu|la si) Ps_mu,la Ps_nu,si (NU/TwoElectronMatrix.cpp:168-241
// differentiate the same spin-resolved exchange).
// is a function of R_B - R_A alone and is differentiated by central differences in R_B
// (MOPAC's diatomic-fragment derivatives), evaluated with the same device functions as the
// energy: one wave per (pair, axis, sign) evaluates E_AB(R_B - R_A +- h e_axis).  Reference: the analytic derivatives of TwoElectronMatrix::addDerivatives,
// OneElectronMatrix::addDerivatives and the repulsion classes (NU/TwoElectronMatrix.cpp:168-241,
// NU/OneElectronMatrix.cpp:136-229).
// ---------------------------------------------------------------------------------------------
__device__ inline double waveSum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

struct PairGradLds {
  double PA[81], PB[81], PAB[81];  // P_AA (na x na), P_BB (nb x nb), P(mu on A, la on B) (na x nb), row-major
  double PsAB[81];                 // the spin density's (mu on A, la on B) block (spin-polarised densities)
  double H[81];                    // H(mu on B, nu on A) at the displaced geometry, row-major nb x na
  double red[2];
};

template <int LA, int LB>
__global__ __launch_bounds__(256) void k_pair_grad(KTables t, KBatch b, const double* __restrict__ P,
                                                   const double* __restrict__ Ps, double h, double* __restrict__ epair) {
  using Cfg = TwoCenterCfg<LA, LB>;
  constexpr int DB = Cfg::DB, W = Cfg::W;
  constexpr int cls = LA * 3 + LB;
  __shared__ TwoCenterLds<LA, LB> S;
  __shared__ PairGradLds G[W];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int item = blockIdx.x * W + w;  // (pair of the class, axis, sign)
  const int idx = b.classStart[cls] + item / 6, sub = item - (item / 6) * 6, axis = sub >> 1, sgn = sub & 1;
  const bool active = (w < W) && idx < b.classStart[cls + 1];
  const int wi = w < W ? w : 0;
  PairDev pr{};
  AtomDev A{}, B{};
  ElementDev const* eA = nullptr;
  ElementDev const* eB = nullptr;
  int na = 0, nb = 0;
  long long n = 0, off = 0;
  if (active) {
    pr = b.pairs[b.classPairs[idx]];
    A = b.atoms[pr.a];
    B = b.atoms[pr.b];
    eA = &b.elems[A.elem];
    eB = &b.elems[B.elem];
    na = A.nao;
    nb = B.nao;
    const MolDev m = b.mols[pr.mol];
    n = m.n;
    off = m.matOff;
    for (int k = lane; k < na * na; k += 64) G[wi].PA[k] = P[off + (long long)(A.ao + k % na) * n + A.ao + k / na];
    for (int k = lane; k < nb * nb; k += 64) G[wi].PB[k] = P[off + (long long)(B.ao + k % nb) * n + B.ao + k / nb];
    for (int k = lane; k < na * nb; k += 64) G[wi].PAB[k] = P[off + (long long)(B.ao + k % nb) * n + A.ao + k / nb];
    if (Ps)
      for (int k = lane; k < na * nb; k += 64) G[wi].PsAB[k] = Ps[off + (long long)(B.ao + k % nb) * n + A.ao + k / nb];
  }
  double e = 0.0;
  {
    double X = 0, Y = 0, Z = 0;
    if (active) {
      const double d = sgn == 0 ? h : -h;
      X = B.x - A.x + (axis == 0 ? d : 0.0);
      Y = B.y - A.y + (axis == 1 ? d : 0.0);
      Z = B.z - A.z + (axis == 2 ? d : 0.0);
    }
    // overlap and off-diagonal H: four lanes per (shell on B, shell on A), each a quarter of the
    // primitive pairs, summed across the group (all lanes of the wave take part in the shuffles)
    {
      const int sp = lane >> 2, part = lane & 3, sh = sp / 3, sl = sp - sh * 3;
      bool valid = false;
      int offH = 0, offL = 0;
      if (active && sp < 9 && sh < eB->nShell && sl < eA->nShell) {
        offH = eB->shellL[sh] == 0 ? 0 : eB->shellL[sh] == 1 ? 1 : 4;
        offL = eA->shellL[sl] == 0 ? 0 : eA->shellL[sl] == 1 ? 1 : 4;
        valid = offH < nb && offL < na;
      }
      double res[6][6];
      shellPairOverlapPart(valid ? *eB : b.elems[0], valid ? sh : 0, valid ? *eA : b.elems[0], valid ? sl : 0, -X, -Y, -Z,
                           part, 4, valid, res);
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          res[i][j] += __shfl_xor(res[i][j], 1, 64);
          res[i][j] += __shfl_xor(res[i][j], 2, 64);
        }
      if (valid && part == 0) {
        double o[5][5];
        cartesianToSpherical(eB->shellL[sh], eA->shellL[sl], res, o);
        for (int k = 0; k < 2 * eB->shellL[sh] + 1; ++k)
          for (int l = 0; l < 2 * eA->shellL[sl] + 1; ++l) {
            const int ah = offH + k, al = offL + l;
            const double bh = ah < 1 ? eB->beta[0] : ah < 4 ? eB->beta[1] : eB->beta[2];
            const double bl = al < 1 ? eA->beta[0] : al < 4 ? eA->beta[1] : eA->beta[2];
            G[wi].H[ah * na + al] = 0.5 * (bh + bl) * o[k][l];
          }
      }
    }
    twoCenterPair<LA, LB>(t, eA, eB, X, Y, Z, active, wi, lane, S);  // contains block barriers
    if (active) {
      double v = 0.0;
      for (int k = lane; k < nb * na; k += 64) {  // 2 sum P(mu on B, nu on A) H(mu, nu)
        const int mu = k / na, nu = k - mu * na;
        v += 2.0 * G[wi].PAB[nu * nb + mu] * G[wi].H[k];
      }
      for (int k = lane; k < na * na; k += 64) {  // VuvB on A's diagonal block
        const int Pi = t.pairIndex[(k / na) * 9 + k % na];
        if (Pi != kNoPair) v += G[wi].PA[k] * S.vuvbA[wi][Pi];
      }
      for (int k = lane; k < nb * nb; k += 64) {  // VuvB on B's diagonal block
        const int Qi = t.pairIndex[(k / nb) * 9 + k % nb];
        if (Qi != kNoPair) v += G[wi].PB[k] * S.vuvbB[wi][Qi];
      }
      // two-electron terms, one (mu, nu) on A and (la, si) on B per lane step
      const int nab = na * na * nb * nb;
      for (int k = lane; k < nab; k += 64) {
        const int ab = k / (nb * nb), ls = k - ab * (nb * nb);
        const int mu = ab / na, nu = ab - mu * na, la = ls / nb, si = ls - la * nb;
        const int Pi = t.pairIndex[mu * 9 + nu], Qi = t.pairIndex[la * 9 + si];
        if (Pi == kNoPair || Qi == kNoPair) continue;
        const double g = S.L[wi][Pi * DB + Qi];
        double x = G[wi].PAB[mu * nb + la] * G[wi].PAB[nu * nb + si];
        if (Ps) x += G[wi].PsAB[mu * nb + la] * G[wi].PsAB[nu * nb + si];
        v += g * (G[wi].PA[mu * na + nu] * G[wi].PB[la * nb + si] - 0.5 * x);
      }
      if (lane == 0) v += pairRepulsion(b, B.elem, A.elem, sqrt(X * X + Y * Y + Z * Z));
      e = waveSum(v);
    }
  }
  if (active && lane == 0) epair[6LL * b.classPairs[idx] + sub] = e;
}

// Gradient of every atom (3 per atom): the central differences of its pairs' energies in partner
// order, + for the pair's second atom, - for its first.
__global__ __launch_bounds__(256) void k_grad_atoms(KBatch b, const double* __restrict__ epair, double h,
                                                    double* __restrict__ grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 3 * b.natoms) return;
  const int a = k / 3, axis = k - a * 3;
  const AtomDev A = b.atoms[a];
  const MolDev m = b.mols[A.mol];
  const int la = a - m.atom0, nat = m.natoms;
  double g = 0.0;
  for (int o = 0; o < nat; ++o) {
    if (o == la) continue;
    const int i = la < o ? la : o, j = la < o ? o : la;
    const long long pid = m.pair0 + (long long)i * nat - (long long)i * (i + 1) / 2 + (j - i - 1);
    const double d = (epair[6 * pid + 2 * axis] - epair[6 * pid + 2 * axis + 1]) / (2.0 * h);
    g += la < o ? -d : d;
  }
  grad[k] = g;
}

// S := identity of every molecule block (the atom-diagonal blocks of S are the identity and
// shells without primitives leave their overlap at zero).
__global__ __launch_bounds__(256) void k_identity(KBatch b, double* __restrict__ M) {
  const MolDev m = b.mols[b.molList[blockIdx.y]];
  const long long n = m.n, nn = n * n;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nn; k += (long long)gridDim.x * blockDim.x) {
    long long c = k / n, r = k - c * n;
    M[m.matOff + k] = r == c ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_mol_pair_sum(KBatch b, const double* __restrict__ v, double* __restrict__ out) {
  const int id = b.molList[blockIdx.x];
  const MolDev m = b.mols[id];
  double s = 0.0;
  for (int p = threadIdx.x; p < m.npairs; p += blockDim.x) s += v[m.pair0 + p];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[id] = red[0];
}

// NDDODensityGuess::calculateGuess (NU/NDDODensityGuess.cpp:31-66): P = S*Nel/(2n) off the
// atom-diagonal blocks (where S is the identity), Z_core/nAO on the diagonal.
__global__ __launch_bounds__(256) void k_density_guess(KBatch b, const double* __restrict__ S, const double* __restrict__ nel,
                                                       double* __restrict__ P) {
  const int id = b.molList[blockIdx.y];
  const MolDev m = b.mols[id];
  const long long n = m.n;
  const long long nn = n * n;
  const double f = nel[id] / (2.0 * (double)n);
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nn; k += (long long)gridDim.x * blockDim.x) {
    long long c = k / n, r = k - c * n;
    if (r != c) P[m.matOff + k] = S[m.matOff + k] * f;  // the diagonal is written below
  }
  // diagonal: core charge / number of AOs of the owning atom
  for (int a = threadIdx.x + blockIdx.x * blockDim.x; a < m.natoms; a += gridDim.x * blockDim.x) {
    const AtomDev A = b.atoms[m.atom0 + a];
    const ElementDev& E = b.elems[A.elem];
    for (int k = 0; k < A.nao; ++k) P[m.matOff + (long long)(A.ao + k) * n + A.ao + k] = E.core / E.nao;
  }
}

}  // namespace

void launchTwoCenter(const KTables& t, const KBatch& b, double* ints, double* vuvb, hipStream_t s) {
  launchClass<0, 0>(t, b, ints, vuvb, s);
  launchClass<0, 1>(t, b, ints, vuvb, s);
  launchClass<0, 2>(t, b, ints, vuvb, s);
  launchClass<1, 0>(t, b, ints, vuvb, s);
  launchClass<1, 1>(t, b, ints, vuvb, s);
  launchClass<1, 2>(t, b, ints, vuvb, s);
  launchClass<2, 0>(t, b, ints, vuvb, s);
  launchClass<2, 1>(t, b, ints, vuvb, s);
  launchClass<2, 2>(t, b, ints, vuvb, s);
}

template <int LA, int LB>
static void launchPairOverlapClass(const KBatch& b, double* S, double* H, double* erepPair, hipStream_t s) {
  constexpr int cls = LA * 3 + LB;
  const long long count = b.classStart[cls + 1] - b.classStart[cls];
  if (count <= 0) return;
  const long long threads = count * (LA + 1) * (LB + 1);
  hipLaunchKernelGGL((k_pair_overlap<LA, LB>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, b, S, H, erepPair);
}

void launchPairOverlap(const KBatch& b, double* S, double* H, double* erepPair, hipStream_t s) {
  if (b.nMolList > 0) hipLaunchKernelGGL(k_identity, dim3(64, b.nMolList), dim3(256), 0, s, b, S);
  if (b.classStart[9] == 0) return;
  launchPairOverlapClass<0, 0>(b, S, H, erepPair, s);
  launchPairOverlapClass<0, 1>(b, S, H, erepPair, s);
  launchPairOverlapClass<0, 2>(b, S, H, erepPair, s);
  launchPairOverlapClass<1, 0>(b, S, H, erepPair, s);
  launchPairOverlapClass<1, 1>(b, S, H, erepPair, s);
  launchPairOverlapClass<1, 2>(b, S, H, erepPair, s);
  launchPairOverlapClass<2, 0>(b, S, H, erepPair, s);
  launchPairOverlapClass<2, 1>(b, S, H, erepPair, s);
  launchPairOverlapClass<2, 2>(b, S, H, erepPair, s);
}

void launchAtomHcore(const KTables& t, const KBatch& b, const double* vuvb, double* H, hipStream_t s) {
  if (b.nAtomList == 0) return;
  hipLaunchKernelGGL(k_atom_hcore, dim3(b.nAtomList), dim3(64), 0, s, t, b, vuvb, H);
}

void launchMirrorLower(const KBatch& b, double* S, double* H, hipStream_t s) {
  if (b.nMolList > 0) hipLaunchKernelGGL(k_mirror_lower, dim3(256, b.nMolList), dim3(256), 0, s, b, S, H);
}

// one workgroup per segment: a copy of `len` consecutive ints
__global__ __launch_bounds__(256) void k_gather_segments(const int* __restrict__ src, int* __restrict__ dst,
                                                         const int* __restrict__ segs) {
  const int q = blockIdx.x;
  const int from = segs[3 * q], to = segs[3 * q + 1], len = segs[3 * q + 2];
  for (int i = threadIdx.x; i < len; i += blockDim.x) dst[to + i] = src[from + i];
}

void launchGatherSegments(const int* src, int* dst, const int* segs, int nsegs, hipStream_t s) {
  if (nsegs > 0) hipLaunchKernelGGL(k_gather_segments, dim3(nsegs), dim3(256), 0, s, src, dst, segs);
}

void launchMolPairSum(const KBatch& b, const double* perPair, double* perMol, hipStream_t s) {
  if (b.nMolList > 0) hipLaunchKernelGGL(k_mol_pair_sum, dim3(b.nMolList), dim3(256), 0, s, b, perPair, perMol);
}

template <int LA, int LB>
static void launchPairGradClass(const KTables& t, const KBatch& b, const double* P, const double* Ps, double h, double* epair,
                                hipStream_t s) {
  constexpr int cls = LA * 3 + LB;
  const long long items = 6LL * (b.classStart[cls + 1] - b.classStart[cls]);
  if (items <= 0) return;
  constexpr int W = TwoCenterCfg<LA, LB>::W;
  hipLaunchKernelGGL((k_pair_grad<LA, LB>), dim3((unsigned)((items + W - 1) / W)), dim3(64 * W), 0, s, t, b, P, Ps, h, epair);
}

void launchGradients(const KTables& t, const KBatch& b, const double* P, const double* Ps, double h, double* epair,
                     double* grad, hipStream_t s) {
  launchPairGradClass<2, 2>(t, b, P, Ps, h, epair, s);
  launchPairGradClass<2, 1>(t, b, P, Ps, h, epair, s);
  launchPairGradClass<1, 2>(t, b, P, Ps, h, epair, s);
  launchPairGradClass<2, 0>(t, b, P, Ps, h, epair, s);
  launchPairGradClass<0, 2>(t, b, P, Ps, h, epair, s);
  launchPairGradClass<1, 1>(t, b, P, Ps, h, epair, s);
  launchPairGradClass<1, 0>(t, b, P, Ps, h, epair, s);
  launchPairGradClass<0, 1>(t, b, P, Ps, h, epair, s);
  launchPairGradClass<0, 0>(t, b, P, Ps, h, epair, s);
  if (b.natoms > 0) hipLaunchKernelGGL(k_grad_atoms, dim3((3 * b.natoms + 255) / 256), dim3(256), 0, s, b, epair, h, grad);
}

void launchDensityGuess(const KBatch& b, const double* S, const double* nElectrons, double* P, hipStream_t s) {
  if (b.nMolList > 0) hipLaunchKernelGGL(k_density_guess, dim3(64, b.nMolList), dim3(256), 0, s, b, S, nElectrons, P);
}

}  // namespace sparrow_amd
