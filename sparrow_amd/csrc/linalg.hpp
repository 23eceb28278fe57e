// Batched dense linear algebra launchers (linalg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "model.hpp"

namespace sparrow_amd {

// One symmetric eigenproblem of the batch.  All pointers are device pointers into per-molecule
// slices of the workspace.
struct EigJob {
  int n, nvec;          // order, number of wanted eigenvectors (lowest first)
  double* A;            // n x n column-major; destroyed (holds the reflectors on exit)
  double *d, *e, *tau;  // length n
  double *v, *w;        // length 2n scratch each
  double* wval;         // all n eigenvalues, ascending
  double* Z;            // n x nvec row-major: eigenvectors of the original matrix on exit
  double* work;         // 5 * n * nvec
  unsigned char* ipiv;  // n * nvec
};

// C = alpha * A * B^T + beta * D, element strides given explicitly (see linalg.hip).
// If ab is non-null, alpha = ab[0] and beta = ab[1] are read on the device.
struct GemmJob {
  int M, N, K, pad;
  const double* A;
  const double* B;
  double* C;
  const double* D;
  const double* ab;  // if set, alpha = ab[0] and beta = ab[1] are read on the device
  double* part;      // if set, part[2t] = trace and part[2t+1] = squared norm of output tile t
  long long sAi, sAk, sBj, sBk, sCi, sCj, sDi, sDj;
  double alpha, beta;
};

// Per-molecule state of the SP2 spectral projection (see linalg.hip).
struct Sp2Dev {
  double nocc, t1, t2, ePrev1, ePrev2, emin, emax;
  double ab[2];  // alpha, beta of the next step: (1,0) X^2, (-1,2) 2X - X^2, (0,1) copy
  int n, step, maxSteps, done, converged;
  int doneStep;        // steps taken when done was set
  int poly, prevPoly;  // polynomial of the last and the previous step (0 X^2, 1 2X - X^2, 2 copy)
};

// One SP2 step Y = alpha X X + beta X of a symmetric n x n X (column-major); alpha and beta
// come from the molecule's controller state.
struct Sp2StepJob {
  int n, pad;
  const double* X;
  double* Y;
};

void launchSytrd(EigJob* jobs, int njobs, hipStream_t s);
void launchStebz(EigJob* jobs, int njobs, int maxN, hipStream_t s);
void launchStein(EigJob* jobs, int njobs, int maxNvec, hipStream_t s);
void launchOrmtr(EigJob* jobs, int njobs, int maxNvec, hipStream_t s);
// tile: 32 or 64 (tileStart must have been computed for the same tile size)
void launchGemm(const GemmJob* jobs, const int* tileStart, int njobs, int ntiles, int tile, hipStream_t s);
// E = FP - (FP)^T into DIIS history slot `slot`; out[mol*slots + s] = <E, E_s> (part: scratch of
// ntiles32 * slots doubles)
void launchCommutatorDots(const MolDev* mols, const int* tiles32, int nmols, int ntiles32, const double* FP, double* hist,
                          int slots, int slot, double* part, double* out, hipStream_t s);
void launchDiisExtrapolate(const MolDev* mols, int nmols, const double* hist, int slots, const double* coef, double* dst,
                           hipStream_t s);
// copy n x n matrices of every molecule; *Hist selects histOff + slot*n*n instead of matOff
void launchCopyMats(const MolDev* mols, int nmols, const double* src, bool srcHist, double* dst, bool dstHist, int slot,
                    hipStream_t s);
// SP2 start: Gershgorin range of F (lo, hi: length-n vectors at vecOff), initial controller
// state, X0 into the X buffers of the even step jobs and the partial traces of X0.
void launchSp2Init(const MolDev* mols, int nmols, long long totalCols, const double* F, const Sp2StepJob* jobs,
                   const int* tileStart, int ntiles, double* lo, double* hi, double* part, Sp2Dev* state, int maxSteps,
                   hipStream_t s);
// controller decision on the X whose partials are in partIn (what the next step would decide)
void launchSp2Decide(const int* tileStart, int nmols, const double* partIn, const Sp2Dev* stateIn, Sp2Dev* stateOut,
                     hipStream_t s);
// One SP2 step of every molecule: each workgroup first takes the controller decision from the
// partials of X (partIn, stateIn -> stateOut), then computes its tile of Y and its partials
// (partOut).  tileStart: prefix sums of tn (tn + 1) / 2 with tn = ceil(n / 32) (tiles on and
// above the diagonal); partials: 2 doubles per tile (trace, squared norm of the tile's share).
void launchSp2Step(const Sp2StepJob* jobs, const int* tileStart, int njobs, int ntiles, const double* partIn,
                   double* partOut, const Sp2Dev* stateIn, Sp2Dev* stateOut, hipStream_t s);
void launchScaleMats(const MolDev* mols, int nmols, const double* src, double a, double* dst, hipStream_t s);
// rmsd[mol] = sqrt(mean((P - Pold)^2)); Pold <- P (part: scratch of ntiles32 doubles)
void launchRmsdUpdate(const MolDev* mols, const int* tiles32, int nmols, int ntiles32, const double* P, double* Pold,
                      double* part, double* rmsd, hipStream_t s);

}  // namespace sparrow_amd
