// NDDO model data shared by the host set-up code and the HIP kernels.
//
// Orbital order per atom: s, x, y, z, x2-y2, xz, z2, yz, xy  (IE/GeneralTypes.h:38 of the reference).
// Charge distributions ("orbital pairs") per atom: 1 (s), 10 (sp), 40 (spd) in the order of
// twoElIntegral_t (IE/GeneralTypes.h:44-55); index 100 = pair without a multipole.
// MultipolePair order: sp1 pd1 pp2 sd2 dd2 ss0 pp0 dd0 (IE/multipoleTypes.h:25).
#pragma once
#include <cstdint>

namespace sparrow_amd {

constexpr int kMaxAO = 9;
constexpr int kNoPair = 100;
constexpr int kOneCenterBlock = 81 * 81;  // doubles per element in the one-centre coefficient table

enum MultipolePairIdx { kSp1 = 0, kPd1, kPp2, kSd2, kDd2, kSs0, kPp0, kDd0 };

// One element type of the loaded parameter set, in runtime units (Hartree, bohr).
struct ElementDev {
  int Z, nao, l, nShell;
  int n1c, pcoreSpecified, nGauss;
  int w1c;  // index of the element's block in its parameter set's one-centre coefficient table
  double core, pcore, alpha, zcbrt;
  double U[3], beta[3];
  double D[5], rho[8];
  double oneC[58];
  double gauss[4][3];  // (a [Eh*bohr], b [bohr^-2], c [bohr]) ; a already divided by eV/Hartree
  int shellL[3];
  int shellPad;
  double shE[3][6], shC[3][6];  // STO-6G exponents / primitive-normalised coefficients
};

// Diatomic PM6 core-core parameters for an (element type, element type) pair.
struct DiatomicDev {
  double alpha, x;
  int kind, pad;  // 0 standard, 1 X-H exponential in R^2, 2 C-C, 3 Si-O
};

// Atom of a batch.
struct AtomDev {
  int elem;   // element type index
  int mol;    // molecule index in the batch
  int ao;     // first AO index inside its molecule
  int nao;
  double x, y, z;  // bohr
  long long jOff;  // offset (doubles) of the atom's Coulomb partials from its partners (see PairDev)
};

// Atom pair (a < b, same molecule) of a batch.
struct PairDev {
  int a, b;          // global atom indices
  int mol, cls;      // cls = lA*3 + lB
  long long intOff;  // offset (doubles) of the dA x dB two-centre block
  long long vecOff;  // offset (doubles) of the dA + dB per-pair vector scratch
  // copies of the atoms' and the molecule's layout, so that the Fock pair kernels reach the
  // density and Fock blocks without two further dependent loads
  int aoA, aoB;      // first AO of atoms a and b inside the molecule
  int n, pad;        // AOs of the molecule
  long long matOff;  // offset (doubles) of the molecule's n x n matrices
  // Coulomb partial vectors of the Fock build, stored atom-major: atom k's partials from its
  // nat - 1 partners (d_k doubles each, partners in index order) are contiguous from its jOff,
  // and the molecule's atoms follow each other from the molecule's first pair vecOff (the same
  // sum_pairs (d_A + d_B) doubles).  jA: where this pair's partial for atom a goes (partner b at
  // position b - 1 of a's list), jB: for atom b (partner a at position a).
  long long jA, jB;
};

// Sets AtomDev::jOff and PairDev::jA, jB of one molecule: nat atoms with charge-distribution
// counts d[k], its pairs in triangular order (i < j, j fastest), partials from base.
inline void setCoulombPartialOffsets(AtomDev* atoms, PairDev* pairs, int nat, const int* d, long long base) {
  long long o = base;
  for (int k = 0; k < nat; ++k) {
    atoms[k].jOff = o;
    o += (long long)(nat - 1) * d[k];
  }
  long long q = 0;
  for (int i = 0; i < nat; ++i)
    for (int j = i + 1; j < nat; ++j, ++q) {
      pairs[q].jA = atoms[i].jOff + (long long)(j - 1) * d[i];
      pairs[q].jB = atoms[j].jOff + (long long)i * d[j];
    }
}

// Molecule of a batch.
struct MolDev {
  int n, nocc;          // AOs, doubly occupied orbitals
  int atom0, natoms;    // atom range
  int pair0, npairs;    // pair range
  int id, pad1;         // id: index of the molecule in its batch (per-molecule arrays)
  long long matOff;     // offset (doubles) of this molecule's n x n matrices
  long long vecOff;     // offset (doubles) of this molecule's length-n vectors
  long long histOff;    // offset (doubles) of this molecule's DIIS history (slots x n x n)
  long long pad2;
};

// Per-molecule SCF progress, kept on the device between iterations.
struct ScfMolDev {
  double Eold;     // energy of the previous iteration
  double Eel;      // electronic energy at the converged density (valid when haveEnergy)
  int iter;        // number of the molecule's current iteration (1-based)
  int converged;   // convergence test passed; P and F_extrap are frozen from then on
  int convIter;    // iteration in which it passed
  int haveEnergy;  // Eel recorded (from the Fock build of the iteration after convergence)
  int lowp;        // densities from split-BF16 x3 SP2 steps (mixed precision: 1e-5 <= |dE| < 1e-1)
  int fp64Start;   // first iteration with FP64 densities after the low-precision window
  int stopped;     // iteration cap reached without convergence; P frozen like a converged one
  int lowpArmed;   // mixed precision: FP64 densities until |dE| first falls below 1e-1
};

// Local-integral evaluation tables for one (lA, lB) class.
//   instance  = one multipole-multipole interaction (m1 on A with pair type p1, m2 on B with p2),
//               a sum over point charges;
//   entry     = one local integral (i|j) = sum_k f_k * instance_k.
struct LocalClassDev {
  int nInst, nEntry, instCharge0, entryTerm0;  // offsets into the shared arrays
  int inst0, entry0, dA, dB;
};
struct MMInstanceDev {
  int p1, p2;            // MultipolePair indices on A / B
  int charge0, nCharge;  // range in the charge-pair array
};
struct ChargePairDev {
  double q;                   // product of the two point charges
  double x1, y1, z1;          // offset of charge 1 in units of D_A[p1]
  double x2, y2, z2;          // offset of charge 2 in units of D_B[p2]
};
struct EntryDev {
  int i, j, term0, nTerm;
};
struct EntryTermDev {
  double f;
  int inst, pad;
};

}  // namespace sparrow_amd
