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

// C = alpha * A * B^T + beta * C, element strides given explicitly (see linalg.hip)
struct GemmJob {
  int M, N, K, pad;
  const double* A;
  const double* B;
  double* C;
  long long sAi, sAk, sBj, sBk, sCi, sCj;
  double alpha, beta;
};

void launchSytrd(EigJob* jobs, int njobs, hipStream_t s);
void launchStebz(EigJob* jobs, int njobs, int maxN, hipStream_t s);
void launchStein(EigJob* jobs, int njobs, int maxNvec, hipStream_t s);
void launchOrmtr(EigJob* jobs, int njobs, int maxNvec, hipStream_t s);
void launchGemm(const GemmJob* jobs, const int* tileStart, int njobs, int ntiles, hipStream_t s);
void launchCommutator(const MolDev* mols, int nmols, const double* FP, double* dst, int slot, hipStream_t s);
void launchDiisDots(const MolDev* mols, int nmols, const double* hist, int slots, int newSlot, double* out, hipStream_t s);
void launchDiisExtrapolate(const MolDev* mols, int nmols, const double* hist, int slots, const double* coef, double* dst,
                           hipStream_t s);
// copy n x n matrices of every molecule; *Hist selects histOff + slot*n*n instead of matOff
void launchCopyMats(const MolDev* mols, int nmols, const double* src, bool srcHist, double* dst, bool dstHist, int slot,
                    hipStream_t s);
void launchRmsdUpdate(const MolDev* mols, int nmols, const double* P, double* Pold, double* rmsd, hipStream_t s);

}  // namespace sparrow_amd
