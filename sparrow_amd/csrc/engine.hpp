// Host side of the NDDO SCF on one GPU: device context, batch layout and the SCF driver.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "linalg.hpp"
#include "params.hpp"
#include "tables.hpp"

namespace sparrow_amd {

struct ScfSettings {
  double energyTol = 1e-7;   // self_consistence_criterion
  double rmsdTol = 1e-5;     // density_rmsd_criterion
  int maxIterations = 100;   // max_scf_iterations
  int mixer = 1;             // 0 no mixer, 1 DIIS, 2 EDIIS, 3 EDIIS+DIIS
  int diisSize = 8;
  int densityBuilder = 0;    // 0 SP2 spectral projection (GEMMs), 1 diagonalisation every iteration
  int sp2MaxSteps = 100;
  int precision = 0;         // 0 FP64; 1 mixed: split-BF16 x3 SP2 products while lowpTol <= |dE| < lowpStart
  double lowpStart = 1e-1;
  double lowpTol = 1e-5;
  bool graphs = false;       // replay repeated iteration launch sequences from hipGraphs
  bool batched = false;      // batched / serving form (SP2 tiles for throughput) or one molecule (latency)
};

struct MoleculeInput {
  std::vector<int> Z;
  std::vector<double> xyz;          // bohr, natoms x 3
  int charge = 0;
  const double* densityGuess = nullptr;  // optional n x n column-major (warm start)
};

struct MoleculeResult {
  double energy = 0, electronicEnergy = 0, repulsionEnergy = 0;
  int iterations = 0;
  bool converged = false;
  int nao = 0, nocc = 0;
  std::vector<double> P, eps, C;  // n x n (col-major), n, n x n (col-major, column k = MO k)
};

struct PhaseTimes {
  double setup = 0, fock = 0, diis = 0, eig = 0, density = 0, total = 0;
  int iterations = 0;
};

// Work counters of the last solve(): lock-step iterations, SP2 steps issued, molecule-iterations
// whose SP2 projection did not converge and were diagonalised instead, batched eigensolves, the
// floating-point operations of the SP2 products that changed X (symmetric count n^2 (n + 1) each),
// and those of them done in split-BF16 x3, iteration segments replayed from hipGraphs and graph
// captures.
struct SolveCounters {
  long long iterations = 0, sp2Steps = 0, sp2Fallbacks = 0, eigensolves = 0, sp2Flops = 0, sp2LowpFlops = 0;
  long long graphLaunches = 0, graphCaptures = 0;
};

// Optional per-kernel timing with HIP events recorded on the engine's stream around the
// launches of the hot kernels (enabled through sa_set_profiling).
class KernelProfiler {
 public:
  ~KernelProfiler();
  void setEnabled(bool on) { enabled_ = on; }
  bool enabled() const { return enabled_; }
  void begin(const char* name, hipStream_t s);
  void end(hipStream_t s);
  void collect();  // after a stream synchronisation: accumulate elapsed times
  void reset();
  bool query(const std::string& name, double* totalMs, long long* count) const;
  std::vector<std::string> names() const;

 private:
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  bool enabled_ = false;
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  std::string current_;
  hipEvent_t curStart_ = nullptr;
  std::map<std::string, std::pair<double, long long>> acc_;
  hipEvent_t get();
};

// Growing device buffer that is reused across calls.
class DeviceArena {
 public:
  ~DeviceArena();
  void reserve(size_t bytes);
  char* base() const { return base_; }
  size_t capacity() const { return cap_; }

 private:
  char* base_ = nullptr;
  size_t cap_ = 0;
};

struct CisState;  // device-resident data of one CIS problem

// One molecule's density-independent part resident on the device (the operator-level handle
// sa_mol_*): layout tables, two-centre blocks, VuvB vectors, S, H, E_rep and the standard
// density guess are computed once per geometry; each Fock build then only runs the G-build.
struct MolState {
  ParameterSet* ps = nullptr;
  int n = 0, natoms = 0;
  double nel = 0, erep = 0;
  char* own = nullptr;  // device block owned by a handle (null: the engine's arena)
  KBatch b{};
  int* aoAtom = nullptr;
  double *ints = nullptr, *vuvb = nullptr, *jvec = nullptr, *S = nullptr, *H = nullptr, *guess = nullptr;
  double *P = nullptr, *F = nullptr, *erepPair = nullptr, *ePair = nullptr, *eAtom = nullptr, *eMol = nullptr;
  double *nelDev = nullptr, *erepDev = nullptr;
  // spin-polarised build: P doubles as P^alpha
  double *Pb = nullptr, *PT = nullptr, *FT = nullptr, *Fa1 = nullptr, *Fb1 = nullptr, *Fa = nullptr, *Fb = nullptr;
  double* E = nullptr;
};

// Size of the static device scratch array of the allocation-free Fock build (noalloc.cpp).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
// Device address of that array on `device` (which must be current); looked up once per device.
char* scratchAddress(int device);

class Engine {
 public:
  Engine(int device, const std::string& paramDir);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  ParameterSet& method(const std::string& name);
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  const KTables& tables() const { return ktab_; }
  // Edge of the SP2 tiles of a molecule with n basis functions: the first entry of the form's
  // table with n <= nmax (the last entry's tile beyond); molecules keep it for all their steps.
  // One molecule alone favours many small workgroups (latency), batches fewer larger ones.
  int sp2Tile(int n, bool batched) const;
  // Replaces a form's table (nmax ascending, tiles 32, 64 or 128); affects later molecules.
  void setSp2Tiles(bool batched, const std::vector<std::pair<int, int>>& table);
  // Device copy of the parameter set's one-centre coefficient table (uploaded once per method).
  const double* oneCenterTable(ParameterSet& ps);
  const double* oneCenterExchangeTable(ParameterSet& ps);
  const ModelTables& hostTables() const { return tables_; }

  // Full restricted SCF for a batch of molecules (one method).  Results in the order of `mols`
  // (densities only when wantDensity).
  std::vector<MoleculeResult> solve(const std::string& methodName, const std::vector<MoleculeInput>& mols,
                                    const ScfSettings& settings, bool wantOrbitals, bool wantEps,
                                    PhaseTimes* times = nullptr, bool wantDensity = true);
  const SolveCounters& counters() const { return counters_; }

  // Operator-level entry points (one molecule): S, H (full symmetric n x n), E_rep; F(P) and E_el.
  void densityIndependent(const std::string& methodName, const MoleculeInput& mol, double* S, double* H, double* Erep);
  void fock(const std::string& methodName, const MoleculeInput& mol, const double* P, double* F, double* Eel);
  // The same build with every device buffer carved from the static scratch array of this GPU
  // (no device allocation; molecules whose buffers exceed kNoAllocScratchBytes are rejected).
  void fockNoAlloc(const std::string& methodName, const MoleculeInput& mol, const double* P, double* F, double* Eel);
  // Molecule handle (FockMatrix::calculateDensityIndependentPart once, then
  // calculateDensityDependentPart per density; FockMatrix.cpp:40-56).  Open shells are allowed
  // (the unrestricted build); molCreate throws for invalid input.
  MolState* molCreate(const std::string& methodName, const MoleculeInput& mol);
  void molDestroy(MolState* m);
  void molDensityIndependent(MolState* m, double* S, double* H, double* Erep);
  void molDensityGuess(MolState* m, double* P);
  void molFock(MolState* m, const double* P, double* F, double* Eel);
  void molFockUnrestricted(MolState* m, const double* Pa, const double* Pb, double* Fa, double* Fb, double* Eel);
  // Unrestricted (spin-polarised) F^alpha, F^beta and E_el for full symmetric Pa, Pb.
  void fockUnrestricted(const std::string& methodName, const MoleculeInput& mol, const double* Pa, const double* Pb,
                        double* Fa, double* Fb, double* Eel);
  // Nuclear gradient dE/dR (Hartree/bohr, natoms x 3 row-major) at the converged density P
  // (n x n column-major): derivatives of the pair energies at fixed P, central differences of
  // step h bohr.  Ps (may be null): the spin density P^a - P^b of a converged unrestricted
  // solution with P = P^a + P^b (any electron count).
  void gradients(const std::string& methodName, const MoleculeInput& mol, const double* P, double* grad,
                 double h = 1e-4, const double* Ps = nullptr);
  // CIS sigma vectors of one molecule (the reference's CISSigmaVectorEvaluator).  cisCreate
  // keeps the two-centre integrals, the orbitals C (n x n column-major, columns ascending in
  // energy, the first nocc occupied) and the orbital-energy differences on the device;
  // cisSigma maps nvec trial vectors X (nconf = nvir * nocc each, virtual index fastest) to
  // sigma = (e_a - e_i) x + vec(C_vir^T F(D) C_occ), D = C_occ x^T C_vir^T, F = 2 c1 J(D) - c2 K(D^T)
  // (singlet) or -c2 K(D^T) (triplet).  deviceMs, when given, receives the device time of the
  // call without the host transfers.  deviceBuffers: X and S are device memory of this engine's
  // GPU (no host transfers; the caller's buffers are read and written in stream order).
  CisState* cisCreate(const std::string& methodName, const MoleculeInput& mol, const double* C, const double* eps, int nocc,
                      bool triplet, double c1, double c2);
  // Unrestricted reference: orbitals and occupations per spin; a trial vector is the alpha part
  // (nvir_a * nocc_a) followed by the beta part, sigma^s = (e_a - e_i) x^s + vec(C_vir^sT F^s C_occ^s),
  // F^s = c1 J(D^alpha + D^beta) - c2 K(D^sT).
  CisState* cisCreateUnrestricted(const std::string& methodName, const MoleculeInput& mol, const double* Ca,
                                  const double* Cb, const double* epsA, const double* epsB, int noccA, int noccB,
                                  double c1, double c2);
  void cisSigma(CisState* st, int nvec, const double* X, double* S, double* deviceMs = nullptr,
                bool deviceBuffers = false);
  void cisDestroy(CisState* st);
  int cisConfigurations(const CisState* st) const;
  // Algorithmic work of one molecule for the measurement: {floating-point operations of its
  // two-centre integral blocks, bytes of one G-build, atom pairs, basis functions} (see
  // sa_work_counts in include/sparrow_amd.h for the counting rules).
  std::array<double, 4> workCounts(const std::string& methodName, const std::vector<int>& Z);
  // Two-centre block of one atom pair (d_A x d_B row-major) as the kernels compute it.
  void twoCenterBlock(const std::string& methodName, int ZA, int ZB, const double* Rab, double* out);
  // Batched symmetric eigensolver / GEMM exposed for testing.
  void eigh(int n, int nvec, const double* A, double* w, double* V);
  // Spectral projector X onto the nocc lowest eigenvectors of the symmetric F (FP64 SP2 on the
  // device, the SCF's own kernels and bounds); false when the projection did not converge (a
  // degenerate HOMO/LUMO), X then holds the last iterate.  steps: purification steps taken.
  bool densityProjector(int n, int nocc, const double* F, double* X, int* steps);
  void gemmNT(int M, int N, int K, const double* A, const double* B, double* C);

  KernelProfiler& profiler() { return prof_; }

 private:
  // F(P) and E_el of one molecule in device memory at base (capacity cap; base null: the growing
  // arena); in scratch mode the one-centre coefficients of the present elements are copied there.
  void fockImpl(const std::string& methodName, const MoleculeInput& mol, const double* P, double* F, double* Eel,
                char* base, size_t cap);
  const std::vector<double>& oneCenterHost(ParameterSet& ps);
  // Lays out and computes a molecule's density-independent part in device memory: an own block
  // (own = true) or the arena.
  void molBuild(MolState& m, ParameterSet& ps, const MoleculeInput& in, bool own);
  // launch() bracketed by HIP events when kernel timing is on (sa_set_profiling)
  template <class Launch>
  void timed(const char* name, Launch&& launch);
  std::map<const ParameterSet*, std::vector<double>> oneCWHost_;
  char* scratch_ = nullptr;  // this device's copy of the allocation-free scratch array
  CisState* cisCreateSpins(const std::string& methodName, const MoleculeInput& mol, int nspin, const double* const* C,
                           const double* const* eps, const int* nocc, bool triplet, double c1, double c2);
  KernelProfiler prof_;
  SolveCounters counters_;
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  std::string paramDir_;
  std::map<std::string, ParameterSet> methods_;
  ModelTables tables_;
  KTables ktab_{};
  std::vector<void*> tableAllocs_;
  std::map<const ParameterSet*, const double*> oneCW_;
  std::map<const ParameterSet*, const double*> oneCK_;
  std::vector<std::pair<int, int>> sp2TilesSingle_{{2047, 32}, {1 << 30, 128}};
  std::vector<std::pair<int, int>> sp2TilesBatched_{{511, 32}, {1 << 30, 128}};
  DeviceArena arena_;
};

}  // namespace sparrow_amd
