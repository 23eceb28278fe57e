/*
 * sparrow_amd: C ABI of the MI355X-native NDDO (MNDO / AM1 / PM3 / RM1 / PM6) restricted SCF.
 *
 * Conventions
 *   - Plain C types only; every buffer is caller-owned host memory.
 *   - Coordinates in bohr, row-major natoms x 3.  Matrices are n x n column-major and full
 *     symmetric (the reference stores H/G/F as lower triangles; both triangles are filled here).
 *   - Energies in Hartree.
 *   - Every function returns 0 on success and a negative value on error; the message is
 *     available from sa_last_error() (thread-local).  No C++ exception crosses this boundary.
 *   - A context owns one GPU stream and the device memory; calls on one context must be
 *     externally serialised (like a reference calculator instance; independent contexts, like
 *     cloned calculators, may run concurrently).
 *
 * Reference interfaces replaced (paths relative to src/Sparrow/Sparrow/Implementations/ of
 * qcscine/sparrow 5.1.0):
 *   sa_scf / sa_scf_batch  <- GenericMethodWrapper::calculate   GenericMethodWrapper.cpp:58-99
 *                             with NDDOMethodWrapper::applySettings Nddo/NDDOMethodWrapper.cpp:68-103
 *                             and the Utils::ScfMethod loop it drives (Nddo/Pm6/Wrapper/PM6MethodWrapper.cpp:87-89)
 *   sa_density_independent <- FockMatrix::calculateDensityIndependentPart  Nddo/Utils/FockMatrix.cpp:40-47
 *                             + OverlapMatrix::calculateOverlap  Nddo/Utils/IntegralsEvaluationUtils/OverlapMatrix.cpp:32-64
 *                             + PM6RepulsionEnergy::getRepulsionEnergy  Nddo/Pm6/PM6RepulsionEnergy.cpp:69-78
 *   sa_fock                <- FockMatrix::calculateDensityDependentPart + getMatrix + calculateElectronicEnergy
 *                             Nddo/Utils/FockMatrix.cpp:49-56,108-145
 *   sa_mol_*               <- the same operator kept alive over an SCF: calculateDensityIndependentPart once
 *                             (FockMatrix.cpp:40-47), calculateDensityDependentPart per density (:49-56)
 *   sa_fock_unrestricted   <- the spin-polarised branches of TwoElectronMatrix (calculateSameAtomBlock,
 *                             calculateDifferentAtomsBlock) Nddo/Utils/TwoElectronMatrix.cpp:82-108,142-160
 *                             + NDDOElectronicEnergyCalculator::unrestrictedEnergy
 *                             Nddo/Utils/NDDOElectronicEnergyCalculator.cpp:69-89
 *   sa_gradients           <- the first-derivative path of NDDOMethodWrapper (Utils::Derivative::First):
 *                             TwoElectronMatrix::addDerivatives  Nddo/Utils/TwoElectronMatrix.cpp:168-241,
 *                             OneElectronMatrix::addDerivatives  Nddo/Utils/OneElectronMatrix.cpp:136-229,
 *                             the repulsion classes' derivatives (Nddo/Pm6/PM6RepulsionEnergy.cpp)
 *   sa_cis_create / sa_cis_create_unrestricted / sa_cis_sigma / sa_cis_sigma_device / sa_cis_destroy
 *                          <- CISSigmaVectorEvaluator (constructor and evaluate, restricted and unrestricted reference)
 *                             Nddo/TimeDependent/LinearResponse/CISSigmaVectorEvaluator.cpp:18-165 with
 *                             CISMatrixAOFockBuilder::buildFock  .../CISMatrixAOFockBuilder.cpp:48-248 and
 *                             CISPseudoDensityBuilder::getPseudoDensityMatrix  .../CISPseudoDensityBuilder.cpp:87-105
 *   sa_two_center_block    <- TwoCenterIntegralContainer::get / Global2c2eMatrix::getGlobalMatrix
 *                             Nddo/Utils/IntegralsEvaluationUtils/TwoCenterIntegralContainer.h:73-76
 *   sa_eigh                <- the symmetric eigensolver of the (absent) Utils::ScfMethod
 *   sa_gemm_nt             <- the density build P = 2 C_occ C_occ^T of the (absent) Utils::ScfMethod
 *                             (inside sa_scf the density comes from an SP2 projection by default:
 *                             the same projector from GEMMs, see density_builder)
 */
#ifndef SPARROW_AMD_H
#define SPARROW_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sa_context sa_context;

/* Settings keys of the reference (Nddo/NDDOMethodWrapper.cpp:71-78). */
typedef struct sa_scf_settings {
  double self_consistence_criterion; /* |dE| threshold, Hartree (reference default 1e-7) */
  double density_rmsd_criterion;     /* RMSD(dP) threshold (reference default 1e-5) */
  int max_scf_iterations;            /* reference default 100 */
  int scf_mixer;                     /* 0 = no_mixer, 1 = diis (reference default), 2 = ediis,
                                        3 = ediis_diis (the reference's mixer names) */
  int diis_size;                     /* DIIS subspace size (at most 16) */
  int density_builder;               /* 0 = SP2 spectral projection on the matrix cores (default),
                                        1 = diagonalisation every iteration (the reference's route);
                                        both give the same density, P = 2 C_occ C_occ^T */
  int precision;                     /* 0 = FP64; 1 = mixed: FP64 densities until |dE| < 0.1 Eh,
                                        then SP2 products in split-BF16 x3 (FP32 accumulation)
                                        until |dE| < 1e-5 Eh, then FP64 densities for at least two
                                        iterations before the convergence test */
  int graph_replay;                  /* 0 = every kernel launched from the host (default); 1 = an
                                        iteration's device work replayed from a hipGraph once its
                                        launch sequence repeats (same results) */
} sa_scf_settings;

typedef struct sa_scf_result {
  double energy;            /* E_el + E_rep */
  double electronic_energy;
  double repulsion_energy;
  int iterations;
  int converged;            /* SuccessfulCalculation (non-convergence is not an error) */
  int nao;
  int nocc;
} sa_scf_result;

const char* sa_last_error(void);

/* Create a context on HIP device `device`; `param_dir` holds <method>.txt parameter tables. */
int sa_create(sa_context** ctx, int device, const char* param_dir);
int sa_destroy(sa_context* ctx);

/* Number of atomic orbitals of a structure under `method` ("MNDO","AM1","PM3","RM1","PM6"). */
int sa_nao(sa_context* ctx, const char* method, int natoms, const int* Z, int* nao);

/* Restricted closed-shell SCF of one molecule.  density_guess (n x n) may be NULL (standard
 * guess, Nddo/Utils/NDDODensityGuess.cpp:31-66) or a previous density (warm start,
 * GenericMethodWrapper::loadState GenericMethodWrapper.cpp:268-278).  P (n x n), eps (n) and
 * C (n x n, column k = MO k) may each be NULL. */
int sa_scf(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr, int charge,
           const sa_scf_settings* settings, const double* density_guess, sa_scf_result* result, double* P,
           double* eps, double* C);

/* Batched SCF: nmol molecules solved concurrently.  Atom arrays are concatenated; natoms[i]
 * gives each molecule's count.  density_guess, P, eps and C are arrays of nmol pointers (the
 * arrays themselves and any entry may be NULL).  phase_seconds (may be NULL) receives host
 * wall times {setup, first part of the iterations (Fock build, mixer, SP2 start and the first
 * chunk of SP2 steps), 0, eigensolves, rest of the iterations (further SP2 steps, P, density
 * change, convergence test), total}. */
int sa_scf_batch(sa_context* ctx, const char* method, int nmol, const int* natoms, const int* Z, const double* xyz_bohr,
                 const int* charge, const sa_scf_settings* settings, const double* const* density_guess,
                 sa_scf_result* results, double* const* P, double* const* eps, double* const* C, double* phase_seconds);

/* Density-independent part: S and H (n x n) and the core-core repulsion energy. */
int sa_density_independent(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr,
                           double* S, double* H, double* repulsion_energy);

/* F = H + G(P) (n x n) and E_el = 1/2 tr[P (H + F)] for a full symmetric density P. */
int sa_fock(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr, const double* P,
            double* F, double* electronic_energy);

/* ---- Molecule handle: the operator-level boundary.  The reference computes a geometry's
 * density-independent part once (FockMatrix::calculateDensityIndependentPart, Nddo/Utils/
 * FockMatrix.cpp:40-47: overlap, one-electron matrix, two-centre integrals, core repulsion) and the
 * density-dependent part per SCF iteration (calculateDensityDependentPart + getMatrix, :49-56,
 * 108-145).  sa_mol_create does the first on the device and keeps it there (two-centre blocks,
 * VuvB vectors, S, H, E_rep and the standard density guess of NDDODensityGuess.cpp:31-66);
 * sa_mol_fock / sa_mol_fock_unrestricted then only run the G-build.  Open shells (odd electron
 * counts) are accepted.  Destroy a handle before its context; a handle is used on its context's
 * stream and shares its serialisation rule. */
typedef struct sa_molecule sa_molecule;
int sa_mol_create(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr, int charge,
                  sa_molecule** mol);
/* Basis size, electron count (core charges - charge) and core-core repulsion energy (any may be NULL). */
int sa_mol_info(const sa_molecule* mol, int* nao, double* n_electrons, double* repulsion_energy);
/* S and H (n x n, may be NULL) and E_rep, as sa_density_independent. */
int sa_mol_density_independent(sa_molecule* mol, double* S, double* H, double* repulsion_energy);
/* The standard guess P0 = S N / (2 n), diagonal Z_core(A) / nAO(A) (NDDODensityGuess::calculateGuess). */
int sa_mol_density_guess(sa_molecule* mol, double* P);
/* As sa_fock and sa_fock_unrestricted, without recomputing the density-independent part. */
int sa_mol_fock(sa_molecule* mol, const double* P, double* F, double* electronic_energy);
int sa_mol_fock_unrestricted(sa_molecule* mol, const double* Pa, const double* Pb, double* Fa, double* Fb,
                             double* electronic_energy);
int sa_mol_destroy(sa_molecule* mol);

/* sa_fock without any device allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that forbid hipMalloc).
 * Same arguments and results as sa_fock (the same kernels, bit-identical F and E_el), with every
 * device buffer taken from one static device array of 64 MiB per GPU, held from library load until
 * the process ends; its address on each GPU is looked up when a context is created there.  A
 * molecule whose buffers do not fit (about 300 PM6 atoms of C/H/N/O) is rejected with an error.
 * Calls on one GPU share the array: calls that overlap on the same GPU (two contexts on one device
 * used from two threads) race. */
int sa_fock_noalloc(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr,
                    const double* P, double* F, double* electronic_energy);

/* Unrestricted (spin-polarised) Fock matrices F^alpha = H + J(Pa + Pb) - K(Pa), F^beta likewise
 * (n x n, full symmetric, column-major) and E_el = 1/2 sum [P H + Pa F^alpha + Pb F^beta] for
 * full symmetric spin densities Pa, Pb. */
int sa_fock_unrestricted(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr,
                         const double* Pa, const double* Pb, double* Fa, double* Fb, double* electronic_energy);

/* Nuclear gradient dE/dR (Hartree/bohr, natoms x 3 row-major) at a converged density P (n x n,
 * full symmetric, from sa_scf): the derivatives of every pair's energy at fixed P (the SCF energy
 * is stationary in P), by central differences of step h bohr (h <= 0 selects 1e-4). */
int sa_gradients(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr, const double* P,
                 double h, double* grad);

/* Nuclear gradient of a converged unrestricted solution (spin densities Pa, Pb, n x n full
 * symmetric; any electron count): the same pair-energy derivatives with each spin's exchange,
 * i.e. the restricted terms of P = Pa + Pb plus -1/2 sum (mu nu|la si) Ps_mu,la Ps_nu,si of the
 * spin density Ps = Pa - Pb.  Replaces the spin-resolved branch of the reference's derivative
 * path, NU/TwoElectronMatrix.cpp:168-241 (TwoElectronMatrix::addDerivatives, unrestricted). */
int sa_gradients_unrestricted(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr,
                              int charge, const double* Pa, const double* Pb, double h, double* grad);

/* Global two-centre two-electron block (d_A x d_B, row-major; d = 1/10/40 for s/sp/spd) of an
 * atom pair with Rab = R_B - R_A (bohr). */
int sa_two_center_block(sa_context* ctx, const char* method, int ZA, int ZB, const double* Rab, double* out);

/* Symmetric eigensolver on the GPU: all n eigenvalues (ascending) and the lowest nvec
 * eigenvectors (V: n x nvec column-major, may be NULL). */
int sa_eigh(sa_context* ctx, int n, int nvec, const double* A, double* w, double* V);

/* Density of an aufbau occupation without orbitals: the spectral projector X (n x n, column-major)
 * onto the nocc lowest eigenvectors of the symmetric F, by FP64 SP2 purification on the GPU (the
 * SCF's own kernels and spectral bounds); equal to C_occ C_occ^T of an eigensolve, i.e. one spin's
 * density, or half the closed-shell P.  Replaces the per-iteration eigensolve + aufbau density of
 * the reference's SCF (the absent Utils::ScfMethod, called from PM6MethodWrapper.cpp:87-89) for an
 * SCF loop kept by the caller, e.g. each spin of an unrestricted one.  *converged (may be NULL) = 0
 * when the projection did not reach an idempotent X of trace nocc (a degenerate HOMO/LUMO); X
 * then holds the last iterate and an eigensolve (sa_eigh) must be used instead. */
int sa_density_projector(sa_context* ctx, int n, int nocc, const double* F, double* X, int* converged);

/* C (M x N) = A (M x K) * B (N x K)^T, all column-major, on the FP64 matrix cores. */
int sa_gemm_nt(sa_context* ctx, int M, int N, int K, const double* A, const double* B, double* C);

/* Kernel timing with HIP events on the context's stream (no counterpart in the reference, whose
 * only timing is the CLI's wall clock around calculate(), App/CalculationHandler.cpp:204-208).
 * Names: twocenter pair_overlap atom_hcore fock_pairs fock_atoms gemm_fp diis_dots sp2_init
 * gemm_sp2 sytrd stebz stein ormtr gemm_p rmsd.  sa_kernel_time returns the accumulated milliseconds and launch count. */
int sa_set_profiling(sa_context* ctx, int enabled);
int sa_reset_profile(sa_context* ctx);
int sa_kernel_time(sa_context* ctx, const char* name, double* total_ms, long long* count);

/* SP2 tile edge by molecule size, per form: batched = 0 for sa_scf (one molecule: latency),
 * 1 for sa_scf_batch and the server (throughput).  A molecule with n basis functions uses
 * tile[i] for the first i with n <= nmax[i] (tile[count - 1] beyond); tiles are 32 (k split
 * over 16 waves: many small workgroups), 64 or 128 (whole k per wave).  nmax ascending.
 * Defaults: one molecule {2047: 32, above: 128}, batched {511: 32, above: 128}.  Affects
 * molecules submitted afterwards; the tables are what sparrow_amd/autotune.py writes.  No
 * counterpart in the reference (its density comes from a diagonalisation). */
int sa_set_sp2_tiles(sa_context* ctx, int batched, int count, const int* nmax, const int* tile);

/* Work counters of the context's last SCF call; the first min(n, 8) of {lock-step iterations,
 * SP2 steps issued, molecule-iterations whose SP2 projection did not converge and were
 * diagonalised instead, batched eigensolves, FLOPs of the SP2 products at the symmetric count
 * (n^2 (n + 1) each: the product is symmetric, so one triangle with the diagonal is the work),
 * FLOPs of those done in split-BF16 x3, iteration segments replayed from hipGraphs, hipGraph
 * captures} are written to out.  No counterpart in the reference. */
int sa_solve_counters(sa_context* ctx, long long* out, int n);

/* Algorithmic work of one molecule, the counts the measurement divides by kernel time (no
 * counterpart in the reference): out[0] = floating-point operations of all its two-centre integral
 * blocks (per atom pair: 16 per point-charge pair of the multipole instances of its (l_A, l_B)
 * class, 2 per local-integral term, 5 per term of the two rotation contractions); out[1] = bytes of
 * one G-build, 8 (sum over pairs of d_A d_B + sum over atoms of their unique one-centre integrals,
 * 1 / 6 / 58) + 8 n (n + 1) (the lower triangles of P read and F written); out[2] = atom pairs;
 * out[3] = n. */
int sa_work_counts(sa_context* ctx, const char* method, int natoms, const int* Z, double* out);

/* ---- Serving loop (no counterpart in the reference, whose calculators solve one structure per
 * calculate() call, GenericMethodWrapper.cpp:58-99; this is the continuous-batching form of the
 * same SCF).  A server keeps up to max_molecules molecules on the GPU in workspace pools sized from
 * hbm_budget_bytes (7/8 of it; 1/8 holds the warm-start density cache).  Requests queue on
 * submit; each sa_server_step admits queued requests in arrival order while the pools have room,
 * runs one lock-step SCF iteration for every molecule on the GPU and retires the finished ones.
 * A request with warm_key >= 0 starts from the cached final density of an earlier request with
 * that key and the same basis size (e.g. a perturbed geometry), and leaves its own final
 * density under that key.  Results are those of sa_scf with the same settings (bit-identical
 * for cold starts).  A server is not thread-safe and shares its context's stream. */
typedef struct sa_server sa_server;

int sa_server_create(sa_context* ctx, const char* method, const sa_scf_settings* settings, double hbm_budget_bytes,
                     int max_molecules, sa_server** server);
int sa_server_destroy(sa_server* server);
/* Validates the molecule (errors as sa_scf) and queues it; ticket identifies the request. */
int sa_server_submit(sa_server* server, int natoms, const int* Z, const double* xyz_bohr, int charge, long long warm_key,
                     long long* ticket);
/* Admission + one SCF iteration; on_gpu / queued (may be NULL) receive the counts after it. */
int sa_server_step(sa_server* server, int* on_gpu, int* queued);
/* done = 1 if the request has finished: result and P (n x n, may be NULL) are filled and the
 * request is forgotten; done = 0 if it is still queued or on the GPU. */
int sa_server_poll(sa_server* server, long long ticket, int* done, sa_scf_result* result, double* P);
/* Number of requests admitted with a cached warm-start density so far. */
int sa_server_warm_hits(sa_server* server, long long* hits);

/* CIS sigma vectors of one molecule with a restricted closed-shell reference.  sa_cis_create
 * keeps the two-centre integrals, the orbitals and the orbital-energy differences on the device:
 * C (n x n, column-major, column k = orbital k, ascending in energy, the first nocc occupied, as
 * sa_scf returns them), eps (n orbital energies), triplet 0 for the singlet block, 1 for the
 * triplet block, c1, c2 the reference's excited-states scaling (1, 1 without a parameter file).
 * sa_cis_sigma maps nvec trial vectors X (nconf = (n - nocc) nocc doubles each, consecutive,
 * virtual index fastest: element (a, i) at a + i (n - nocc)) to
 *   sigma = (e_a - e_i) x + vec(C_vir^T F C_occ),  D = C_occ x^T C_vir^T,
 *   F = 2 c1 J(D) - c2 K(D^T) (singlet) or -c2 K(D^T) (triplet),
 * in the configuration order of the reference's standard order (its energy ordering is a
 * permutation the caller applies).  Trial vectors go through in blocks of up to 64. */
typedef struct sa_cis sa_cis;
int sa_cis_create(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr, const double* C,
                  const double* eps, int nocc, int triplet, double c1, double c2, sa_cis** out);
/* Unrestricted reference (CISSigmaVectorEvaluator<Unrestricted>): orbitals, energies and
 * occupations per spin; a trial vector is its alpha part ((n - nocc_alpha) nocc_alpha, virtual
 * index fastest) followed by its beta part, and
 *   sigma^s = (e_a - e_i) x^s + vec(C_vir^sT F^s C_occ^s),  F^s = c1 J(D^alpha + D^beta) - c2 K(D^sT). */
int sa_cis_create_unrestricted(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr,
                               const double* C_alpha, const double* C_beta, const double* eps_alpha,
                               const double* eps_beta, int nocc_alpha, int nocc_beta, double c1, double c2,
                               sa_cis** out);
/* Number of configurations: (n - nocc) nocc, or the sum over both spins. */
int sa_cis_configurations(const sa_cis* cis, int* nconf);
int sa_cis_sigma(sa_cis* cis, int nvec, const double* X, double* sigma);
/* The same with X and sigma in device memory of the context's GPU (a trial subspace kept in HBM:
 * no host transfers); both buffers must hold nvec * nconf doubles (the library only sees the
 * pointers and cannot check their sizes).  The call returns when sigma is written.  device_ms
 * (may be NULL) receives the GPU time of the call. */
int sa_cis_sigma_device(sa_cis* cis, int nvec, const double* X_device, double* sigma_device, double* device_ms);
/* Frees the evaluator's device memory; call it before sa_destroy of its context. */
int sa_cis_destroy(sa_cis* cis);

#ifdef __cplusplus
}
#endif

#endif /* SPARROW_AMD_H */
