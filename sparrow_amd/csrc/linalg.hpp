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
struct GemmJob {
  int M, N, K, pad;
  const double* A;
  const double* B;
  double* C;
  const double* D;
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
  int lowp;            // products in split-BF16 x3 (1) or FP64 (0)
  int lowpSteps;       // steps taken in split-BF16 x3
  int tail;            // FP64 purification steps left after the split-BF16 x3 steps (-1: none)
  int xSplit;          // the step's input X is held by its split copies only (X = Xh + Xl exactly)
};

// One SP2 step Y = alpha X X + beta X of a symmetric n x n X (column-major); alpha and beta
// come from the molecule's controller state.
struct Sp2StepJob {
  int n, mol;  // mol: index of the molecule's controller state
  const double* X;
  double* Y;
  // split-BF16 copies X = Xh + Xl (bf16 bit patterns, ldp x ldp column-major, ldp = n rounded up
  // to 32, zero outside n x n); null when the molecule never uses the low-precision steps
  const unsigned short* Xh;
  const unsigned short* Xl;
  unsigned short* Yh;
  unsigned short* Yl;
  int ldp;
  int tile;  // edge of the molecule's SP2 tiles: 32 (k split over 16 waves), 64 or 128 (whole k per wave)
};

void launchSytrd(EigJob* jobs, int njobs, hipStream_t s);
// one job of order n, the trailing updates spread over many workgroups (2 launches per column;
// needs J.work of at least ceil((n - 1) / 64) n doubles)
void launchSytrdWide(EigJob* job, int n, hipStream_t s);
void launchStebz(EigJob* jobs, int njobs, int maxN, hipStream_t s);
void launchStein(EigJob* jobs, int njobs, int maxNvec, hipStream_t s);
void launchOrmtr(EigJob* jobs, int njobs, int maxN, int maxNvec, hipStream_t s);
// one job (J: host copy of *dJ): Z <- Q Z by blocks of 64 reflectors, two FP64-MFMA GEMMs per
// block; J.work must hold 3 n^2 doubles; hG / dG: ormtrGemmJobs(n) GemmJob slots on the host
// (kept alive until the stream has run) and on the device; dZero: one device int holding 0
int ormtrGemmJobs(int n);
void launchOrmtrGemm(const EigJob& J, EigJob* dJ, GemmJob* hG, GemmJob* dG, const int* dZero, hipStream_t s);
// tile: 32 or 64 (tileStart must have been computed for the same tile size)
void launchGemm(const GemmJob* jobs, const int* tileStart, int njobs, int ntiles, int tile, hipStream_t s);
// E = FP - (FP)^T into the DIIS history slot of the molecule's current iteration,
// (scf.iter - 1) % slots, and out[id*slots + s] = <E, E_s> (part: scratch of ntiles32 * slots
// doubles; id = MolDev::id)
void launchCommutatorDots(const MolDev* mols, const int* tiles32, int nmols, int ntiles32, const double* FP, double* hist,
                          int slots, const ScfMolDev* scf, double* part, double* out, hipStream_t s);
// DIIS coefficients coef[id*slots + s] from the new error overlaps dots[id*slots + s]; Bmat:
// slots x slots per molecule, kept across iterations
void launchDiisSolve(const MolDev* mols, int nmols, const ScfMolDev* scf, const double* dots, double* Bmat, double* coef,
                     int slots, hipStream_t s);
// T = F P (F, P symmetric n x n at matOff) for every molecule of the list: the first nmols32 over
// their 32 x 32 tiles with the k range split over 16 waves (tiles32: prefix sums of the tile
// counts, ntiles32 tiles for those molecules; nmols32 < 0: all molecules), the others (molecules
// with 64-wide SP2 tiles) over 64 x 64 tiles, whole k per wave (tiles64: prefix sums of ceil(n/64)^2
// over mols + nmols32, ntiles64 in all).
void launchFPProduct(const MolDev* mols, const int* tiles32, int nmols, int ntiles32, const double* F, const double* P,
                     double* T, hipStream_t s, int nmols32 = -1, const int* tiles64 = nullptr, int ntiles64 = 0);
// EDIIS coefficients into coef (blend: EDIIS+DIIS mixing with the DIIS coefficients already
// there).  histP/histF: densities and Fock matrices of the history (slot of the current
// iteration filled); eHist (slots per molecule): energies, the current one copied from eMol;
// part: scratch of ntiles32 * 2 slots doubles; pf: 2 slots per molecule; Mmat: slots^2 per
// molecule, kept across iterations; dots: the DIIS error overlaps of this iteration.
void launchEdiis(const MolDev* mols, const int* tiles32, int nmols, int ntiles32, const double* histP, const double* histF,
                 int slots, const ScfMolDev* scf, const double* eMol, double* eHist, double* part, double* pf,
                 const double* dots, double* Mmat, double* coef, bool blend, hipStream_t s);
// convergence test / final energy / next iteration number for the molecules in molList;
// mixed-precision molecules: low-precision densities from |dE| < lowpStart until |dE| < lowpTol
void launchScfCheck(const int* molList, int nmols, ScfMolDev* scf, const double* eMol, const double* rmsd, double energyTol,
                    double rmsdTol, double lowpStart, double lowpTol, int maxIterations, hipStream_t s);
void launchDiisExtrapolate(const MolDev* mols, int nmols, int maxN, const double* hist, int slots, const double* coef, double* dst,
                           hipStream_t s);
// copy n x n matrices of every molecule; *Hist selects the history slot of the molecule's current
// iteration, histOff + ((scf.iter - 1) % slots) n^2, instead of matOff (scf may be null otherwise)
void launchCopyMats(const MolDev* mols, int nmols, int maxN, const double* src, bool srcHist, double* dst, bool dstHist,
                    const ScfMolDev* scf, int slots, hipStream_t s);
// SP2 start: Gershgorin range of F (lo, hi: length-n vectors at vecOff), initial controller
// state, X0 into the X buffers of the step jobs and the partial traces of X0.
// scf: per-molecule SCF state; a molecule with scf.lowp set starts in split-BF16 x3 (may be null)
// lz, lzCoef (both or neither): a short Lanczos run narrows the range of each molecule of
// max(lzMinN, 32) to 8000 AOs (lz: 5 n doubles per molecule at 5 vecOff; lzCoef: sp2LanczosCoefs()
// doubles per molecule id); whether a molecule's range is narrowed depends on its size only
void launchSp2Init(const MolDev* mols, int nmols, int maxN, const double* F, const Sp2StepJob* jobs, const int* tileStart,
                   int ntiles, double* lo, double* hi, double* part, Sp2Dev* state, int maxSteps, const ScfMolDev* scf,
                   hipStream_t s, double* lz = nullptr, double* lzCoef = nullptr, int lzMinN = 0);
int sp2LanczosCoefs();
// controller decision on the X whose partials are in partIn (what the next step would decide)
void launchSp2Decide(const Sp2StepJob* jobs, const int* tileStart, int njobs, const double* partIn, const Sp2Dev* stateIn,
                     Sp2Dev* stateOut, hipStream_t s);
// P = 2 X_final, X_s in buffer X0 (even s) or X1 (odd s)
void launchPFromX(const MolDev* mols, int nmols, const Sp2Dev* state, const double* X0, const double* X1, double* P,
                  hipStream_t s);
// One SP2 step of every molecule: each workgroup first takes the controller decision from the
// partials of X (partIn, stateIn -> stateOut), then computes its tile of Y and its partials
// (partOut).  tileStart: prefix sums of tn (tn + 1) / 2 with tn = ceil(n / tile) (tiles on and
// above the diagonal); partials: 2 doubles per tile (trace, squared norm of the tile's share).
// The jobs are ordered by tile edge: nt32 tiles of 32-wide molecules first, then nt64 and nt128.
void launchSp2Step(const Sp2StepJob* jobs, const int* tileStart, int njobs, int nt32, int nt64, int nt128,
                   const double* partIn, double* partOut, const Sp2Dev* stateIn, Sp2Dev* stateOut, hipStream_t s);
// rmsd[mol] = sqrt(mean((P - Pold)^2)); Pold <- P (part: scratch of ntiles32 doubles)
void launchRmsdUpdate(const MolDev* mols, const int* tiles32, int nmols, int ntiles32, const double* P, double* Pold,
                      double* part, double* rmsd, hipStream_t s);

}  // namespace sparrow_amd
