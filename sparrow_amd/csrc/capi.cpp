// extern "C" boundary of libsparrow_amd (declarations and reference citations: include/sparrow_amd.h).
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "../../include/sparrow_amd.h"
#include "engine.hpp"
#include "server.hpp"

struct sa_context {
  std::unique_ptr<sparrow_amd::Engine> engine;
};

struct sa_server {
  sa_context* ctx;
  std::unique_ptr<sparrow_amd::Server> server;
};

struct sa_cis {
  sa_context* ctx;
  sparrow_amd::CisState* state;
};

struct sa_molecule {
  sa_context* ctx;
  sparrow_amd::MolState* state;
};

namespace {
thread_local std::string gLastError;

template <class Fn>
int guard(Fn&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    gLastError = e.what();
    return -1;
  } catch (...) {
    gLastError = "unknown error";
    return -2;
  }
}

sparrow_amd::MoleculeInput makeInput(int natoms, const int* Z, const double* xyz, int charge) {
  if (natoms <= 0) throw std::runtime_error("empty molecular structure");
  if (!Z || !xyz) throw std::runtime_error("null element or coordinate array");
  sparrow_amd::MoleculeInput m;
  m.Z.assign(Z, Z + natoms);
  m.xyz.assign(xyz, xyz + 3 * (size_t)natoms);
  m.charge = charge;
  return m;
}

sparrow_amd::ScfSettings makeSettings(const sa_scf_settings* s) {
  sparrow_amd::ScfSettings st;
  if (s) {
    st.energyTol = s->self_consistence_criterion;
    st.rmsdTol = s->density_rmsd_criterion;
    st.maxIterations = s->max_scf_iterations;
    st.mixer = s->scf_mixer;
    st.diisSize = s->diis_size > 0 ? s->diis_size : 8;
    if (st.diisSize > 16) throw std::runtime_error("diis_size must be at most 16");
    st.densityBuilder = s->density_builder;
    st.precision = s->precision;
    st.graphs = s->graph_replay != 0;
    if (s->precision != 0 && s->precision != 1) throw std::runtime_error("precision must be 0 (FP64) or 1 (mixed)");
  }
  if (st.densityBuilder != 0 && st.densityBuilder != 1)
    throw std::runtime_error("density_builder must be 0 (SP2 projection) or 1 (diagonalisation)");
  if (st.maxIterations < 1) throw std::runtime_error("max_scf_iterations must be >= 1");
  if (st.mixer < 0 || st.mixer > 3)
    throw std::runtime_error("scf_mixer must be 0 (no_mixer), 1 (diis), 2 (ediis) or 3 (ediis_diis)");
  return st;
}

void fillResult(const sparrow_amd::MoleculeResult& r, sa_scf_result* out, double* P, double* eps, double* C) {
  if (out) {
    out->energy = r.energy;
    out->electronic_energy = r.electronicEnergy;
    out->repulsion_energy = r.repulsionEnergy;
    out->iterations = r.iterations;
    out->converged = r.converged ? 1 : 0;
    out->nao = r.nao;
    out->nocc = r.nocc;
  }
  if (P && !r.P.empty()) std::memcpy(P, r.P.data(), r.P.size() * sizeof(double));
  if (eps && !r.eps.empty()) std::memcpy(eps, r.eps.data(), r.eps.size() * sizeof(double));
  if (C && !r.C.empty()) std::memcpy(C, r.C.data(), r.C.size() * sizeof(double));
}
}  // namespace

extern "C" {

const char* sa_last_error(void) { return gLastError.c_str(); }

int sa_create(sa_context** ctx, int device, const char* param_dir) {
  return guard([&] {
    if (!ctx || !param_dir) throw std::runtime_error("sa_create: null argument");
    auto* c = new sa_context();
    try {
      c->engine = std::make_unique<sparrow_amd::Engine>(device, param_dir);
    } catch (...) {
      delete c;
      throw;
    }
    *ctx = c;
  });
}

int sa_destroy(sa_context* ctx) {
  return guard([&] { delete ctx; });
}

int sa_nao(sa_context* ctx, const char* method, int natoms, const int* Z, int* nao) {
  return guard([&] {
    if (!ctx || !method || !nao) throw std::runtime_error("sa_nao: null argument");
    auto& ps = ctx->engine->method(method);
    int n = 0;
    for (int i = 0; i < natoms; ++i) n += ps.element(Z[i], ctx->engine->hostTables()).nao;
    *nao = n;
  });
}

int sa_scf(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr, int charge,
           const sa_scf_settings* settings, const double* density_guess, sa_scf_result* result, double* P, double* eps,
           double* C) {
  return guard([&] {
    if (!ctx || !method) throw std::runtime_error("sa_scf: null argument");
    auto in = makeInput(natoms, Z, xyz_bohr, charge);
    in.densityGuess = density_guess;
    auto res = ctx->engine->solve(method, {in}, makeSettings(settings), C != nullptr, eps != nullptr, nullptr, P != nullptr);
    fillResult(res[0], result, P, eps, C);
  });
}

int sa_scf_batch(sa_context* ctx, const char* method, int nmol, const int* natoms, const int* Z, const double* xyz_bohr,
                 const int* charge, const sa_scf_settings* settings, const double* const* density_guess,
                 sa_scf_result* results, double* const* P, double* const* eps, double* const* C, double* phase_seconds) {
  return guard([&] {
    if (!ctx || !method || !natoms || nmol <= 0) throw std::runtime_error("sa_scf_batch: bad arguments");
    std::vector<sparrow_amd::MoleculeInput> ins;
    size_t off = 0;
    bool wantC = false, wantEps = false, wantP = false;
    for (int i = 0; i < nmol; ++i) {
      auto in = makeInput(natoms[i], Z + off, xyz_bohr + 3 * off, charge ? charge[i] : 0);
      if (density_guess) in.densityGuess = density_guess[i];
      ins.push_back(std::move(in));
      off += natoms[i];
      if (C && C[i]) wantC = true;
      if (eps && eps[i]) wantEps = true;
      if (P && P[i]) wantP = true;
    }
    sparrow_amd::PhaseTimes pt;
    sparrow_amd::ScfSettings st = makeSettings(settings);
    st.batched = true;
    auto res = ctx->engine->solve(method, ins, st, wantC, wantEps, &pt, wantP);
    for (int i = 0; i < nmol; ++i)
      fillResult(res[i], results ? &results[i] : nullptr, P ? P[i] : nullptr, eps ? eps[i] : nullptr, C ? C[i] : nullptr);
    if (phase_seconds) {
      phase_seconds[0] = pt.setup;
      phase_seconds[1] = pt.fock;
      phase_seconds[2] = pt.diis;
      phase_seconds[3] = pt.eig;
      phase_seconds[4] = pt.density;
      phase_seconds[5] = pt.total;
    }
  });
}

int sa_density_independent(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr,
                           double* S, double* H, double* repulsion_energy) {
  return guard([&] {
    if (!ctx || !method) throw std::runtime_error("sa_density_independent: null argument");
    ctx->engine->densityIndependent(method, makeInput(natoms, Z, xyz_bohr, 0), S, H, repulsion_energy);
  });
}

int sa_fock(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr, const double* P,
            double* F, double* electronic_energy) {
  return guard([&] {
    if (!ctx || !method || !P) throw std::runtime_error("sa_fock: null argument");
    ctx->engine->fock(method, makeInput(natoms, Z, xyz_bohr, 0), P, F, electronic_energy);
  });
}

int sa_mol_create(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr, int charge,
                  sa_molecule** out) {
  return guard([&] {
    if (!ctx || !method || !out) throw std::runtime_error("sa_mol_create: null argument");
    *out = nullptr;
    auto* m = new sa_molecule{ctx, nullptr};
    try {
      m->state = ctx->engine->molCreate(method, makeInput(natoms, Z, xyz_bohr, charge));
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int sa_mol_info(const sa_molecule* mol, int* nao, double* n_electrons, double* repulsion_energy) {
  return guard([&] {
    if (!mol || !mol->state) throw std::runtime_error("sa_mol_info: null molecule");
    if (nao) *nao = mol->state->n;
    if (n_electrons) *n_electrons = mol->state->nel;
    if (repulsion_energy) *repulsion_energy = mol->state->erep;
  });
}

int sa_mol_density_independent(sa_molecule* mol, double* S, double* H, double* repulsion_energy) {
  return guard([&] {
    if (!mol || !mol->state) throw std::runtime_error("sa_mol_density_independent: null molecule");
    mol->ctx->engine->molDensityIndependent(mol->state, S, H, repulsion_energy);
  });
}

int sa_mol_density_guess(sa_molecule* mol, double* P) {
  return guard([&] {
    if (!mol || !mol->state || !P) throw std::runtime_error("sa_mol_density_guess: null argument");
    mol->ctx->engine->molDensityGuess(mol->state, P);
  });
}

int sa_mol_fock(sa_molecule* mol, const double* P, double* F, double* electronic_energy) {
  return guard([&] {
    if (!mol || !mol->state || !P) throw std::runtime_error("sa_mol_fock: null argument");
    mol->ctx->engine->molFock(mol->state, P, F, electronic_energy);
  });
}

int sa_mol_fock_unrestricted(sa_molecule* mol, const double* Pa, const double* Pb, double* Fa, double* Fb,
                             double* electronic_energy) {
  return guard([&] {
    if (!mol || !mol->state || !Pa || !Pb) throw std::runtime_error("sa_mol_fock_unrestricted: null argument");
    mol->ctx->engine->molFockUnrestricted(mol->state, Pa, Pb, Fa, Fb, electronic_energy);
  });
}

int sa_mol_destroy(sa_molecule* mol) {
  return guard([&] {
    if (!mol) return;
    sparrow_amd::MolState* st = mol->state;
    sa_context* ctx = mol->ctx;
    delete mol;
    if (st) ctx->engine->molDestroy(st);
  });
}

int sa_fock_noalloc(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr,
                    const double* P, double* F, double* electronic_energy) {
  return guard([&] {
    if (!ctx || !method || !P) throw std::runtime_error("sa_fock_noalloc: null argument");
    ctx->engine->fockNoAlloc(method, makeInput(natoms, Z, xyz_bohr, 0), P, F, electronic_energy);
  });
}

int sa_fock_unrestricted(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr,
                         const double* Pa, const double* Pb, double* Fa, double* Fb, double* electronic_energy) {
  return guard([&] {
    if (!ctx || !method || !Pa || !Pb) throw std::runtime_error("sa_fock_unrestricted: null argument");
    ctx->engine->fockUnrestricted(method, makeInput(natoms, Z, xyz_bohr, 0), Pa, Pb, Fa, Fb, electronic_energy);
  });
}

int sa_gradients(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr, const double* P,
                 double h, double* grad) {
  return guard([&] {
    if (!ctx || !method || !P || !grad) throw std::runtime_error("sa_gradients: null argument");
    ctx->engine->gradients(method, makeInput(natoms, Z, xyz_bohr, 0), P, grad, h > 0.0 ? h : 1e-4);
  });
}

int sa_gradients_unrestricted(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr,
                              int charge, const double* Pa, const double* Pb, double h, double* grad) {
  return guard([&] {
    if (!ctx || !method || !Pa || !Pb || !grad) throw std::runtime_error("sa_gradients_unrestricted: null argument");
    auto in = makeInput(natoms, Z, xyz_bohr, charge);
    auto& ps = ctx->engine->method(method);
    int n = 0;
    for (int i = 0; i < natoms; ++i) n += ps.element(Z[i], ctx->engine->hostTables()).nao;
    const size_t nn = (size_t)n * n;
    std::vector<double> P(nn), Ps(nn);
    for (size_t k = 0; k < nn; ++k) {
      P[k] = Pa[k] + Pb[k];
      Ps[k] = Pa[k] - Pb[k];
    }
    ctx->engine->gradients(method, in, P.data(), grad, h > 0.0 ? h : 1e-4, Ps.data());
  });
}

int sa_cis_create(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr, const double* C,
                  const double* eps, int nocc, int triplet, double c1, double c2, sa_cis** out) {
  return guard([&] {
    if (!ctx || !method || !C || !eps || !out) throw std::runtime_error("sa_cis_create: null argument");
    sparrow_amd::CisState* st =
        ctx->engine->cisCreate(method, makeInput(natoms, Z, xyz_bohr, 0), C, eps, nocc, triplet != 0, c1, c2);
    *out = new sa_cis{ctx, st};
  });
}

int sa_cis_create_unrestricted(sa_context* ctx, const char* method, int natoms, const int* Z, const double* xyz_bohr,
                               const double* C_alpha, const double* C_beta, const double* eps_alpha,
                               const double* eps_beta, int nocc_alpha, int nocc_beta, double c1, double c2,
                               sa_cis** out) {
  return guard([&] {
    if (!ctx || !method || !C_alpha || !C_beta || !eps_alpha || !eps_beta || !out)
      throw std::runtime_error("sa_cis_create_unrestricted: null argument");
    sparrow_amd::CisState* st = ctx->engine->cisCreateUnrestricted(method, makeInput(natoms, Z, xyz_bohr, 0), C_alpha, C_beta,
                                                                   eps_alpha, eps_beta, nocc_alpha, nocc_beta, c1, c2);
    *out = new sa_cis{ctx, st};
  });
}

int sa_cis_configurations(const sa_cis* cis, int* nconf) {
  return guard([&] {
    if (!cis || !nconf) throw std::runtime_error("sa_cis_configurations: null argument");
    *nconf = cis->ctx->engine->cisConfigurations(cis->state);
  });
}

int sa_cis_sigma(sa_cis* cis, int nvec, const double* X, double* sigma) {
  return guard([&] {
    if (!cis || (nvec > 0 && (!X || !sigma))) throw std::runtime_error("sa_cis_sigma: null argument");
    cis->ctx->engine->cisSigma(cis->state, nvec, X, sigma);
  });
}

int sa_cis_sigma_device(sa_cis* cis, int nvec, const double* X_device, double* sigma_device, double* device_ms) {
  return guard([&] {
    if (!cis || (nvec > 0 && (!X_device || !sigma_device))) throw std::runtime_error("sa_cis_sigma_device: null argument");
    cis->ctx->engine->cisSigma(cis->state, nvec, X_device, sigma_device, device_ms, /*deviceBuffers*/ true);
  });
}

int sa_cis_destroy(sa_cis* cis) {
  return guard([&] {
    if (!cis) return;
    cis->ctx->engine->cisDestroy(cis->state);
    delete cis;
  });
}

int sa_two_center_block(sa_context* ctx, const char* method, int ZA, int ZB, const double* Rab, double* out) {
  return guard([&] {
    if (!ctx || !method || !Rab || !out) throw std::runtime_error("sa_two_center_block: null argument");
    ctx->engine->twoCenterBlock(method, ZA, ZB, Rab, out);
  });
}

int sa_eigh(sa_context* ctx, int n, int nvec, const double* A, double* w, double* V) {
  return guard([&] {
    if (!ctx || !A || !w) throw std::runtime_error("sa_eigh: null argument");
    ctx->engine->eigh(n, nvec, A, w, V);
  });
}

int sa_density_projector(sa_context* ctx, int n, int nocc, const double* F, double* X, int* converged) {
  return guard([&] {
    if (!ctx || !F || !X) throw std::runtime_error("sa_density_projector: null argument");
    int steps = 0;
    const bool ok = ctx->engine->densityProjector(n, nocc, F, X, &steps);
    if (converged) *converged = ok ? 1 : 0;
  });
}

int sa_gemm_nt(sa_context* ctx, int M, int N, int K, const double* A, const double* B, double* C) {
  return guard([&] {
    if (!ctx || !A || !B || !C) throw std::runtime_error("sa_gemm_nt: null argument");
    ctx->engine->gemmNT(M, N, K, A, B, C);
  });
}

int sa_set_profiling(sa_context* ctx, int enabled) {
  return guard([&] {
    if (!ctx) throw std::runtime_error("sa_set_profiling: null context");
    ctx->engine->profiler().setEnabled(enabled != 0);
  });
}

int sa_reset_profile(sa_context* ctx) {
  return guard([&] {
    if (!ctx) throw std::runtime_error("sa_reset_profile: null context");
    ctx->engine->profiler().reset();
  });
}

int sa_kernel_time(sa_context* ctx, const char* name, double* total_ms, long long* count) {
  return guard([&] {
    if (!ctx || !name || !total_ms || !count) throw std::runtime_error("sa_kernel_time: null argument");
    ctx->engine->profiler().query(name, total_ms, count);
  });
}

int sa_set_sp2_tiles(sa_context* ctx, int batched, int count, const int* nmax, const int* tile) {
  return guard([&] {
    if (!ctx || count <= 0 || !nmax || !tile) throw std::runtime_error("sa_set_sp2_tiles: bad arguments");
    std::vector<std::pair<int, int>> t;
    for (int i = 0; i < count; ++i) t.emplace_back(nmax[i], tile[i]);
    ctx->engine->setSp2Tiles(batched != 0, t);
  });
}

int sa_work_counts(sa_context* ctx, const char* method, int natoms, const int* Z, double* out) {
  return guard([&] {
    if (!ctx || !method || !Z || !out || natoms <= 0) throw std::runtime_error("sa_work_counts: bad arguments");
    const auto w = ctx->engine->workCounts(method, std::vector<int>(Z, Z + natoms));
    for (int k = 0; k < 4; ++k) out[k] = w[k];
  });
}

int sa_solve_counters(sa_context* ctx, long long* out, int n) {
  return guard([&] {
    if (!ctx || (!out && n > 0)) throw std::runtime_error("sa_solve_counters: null argument");
    const auto& c = ctx->engine->counters();
    const long long v[8] = {c.iterations, c.sp2Steps,     c.sp2Fallbacks,   c.eigensolves,
                            c.sp2Flops,   c.sp2LowpFlops, c.graphLaunches, c.graphCaptures};
    for (int k = 0; k < n && k < 8; ++k) out[k] = v[k];
  });
}

int sa_server_create(sa_context* ctx, const char* method, const sa_scf_settings* settings, double hbm_budget_bytes,
                     int max_molecules, sa_server** out) {
  return guard([&] {
    if (!ctx || !method || !out) throw std::runtime_error("sa_server_create: null argument");
    auto* sv = new sa_server();
    try {
      sv->ctx = ctx;
      sparrow_amd::ScfSettings st = makeSettings(settings);
      st.batched = true;
      sv->server = std::make_unique<sparrow_amd::Server>(*ctx->engine, method, st, hbm_budget_bytes, max_molecules);
    } catch (...) {
      delete sv;
      throw;
    }
    *out = sv;
  });
}

int sa_server_destroy(sa_server* server) {
  return guard([&] { delete server; });
}

int sa_server_submit(sa_server* server, int natoms, const int* Z, const double* xyz_bohr, int charge, long long warm_key,
                     long long* ticket) {
  return guard([&] {
    if (!server || !ticket) throw std::runtime_error("sa_server_submit: null argument");
    *ticket = server->server->submit(makeInput(natoms, Z, xyz_bohr, charge), warm_key);
  });
}

int sa_server_step(sa_server* server, int* on_gpu, int* queued) {
  return guard([&] {
    if (!server) throw std::runtime_error("sa_server_step: null argument");
    const int g = server->server->step();
    if (on_gpu) *on_gpu = g;
    if (queued) *queued = server->server->queued();
  });
}

int sa_server_poll(sa_server* server, long long ticket, int* done, sa_scf_result* result, double* P) {
  return guard([&] {
    if (!server || !done) throw std::runtime_error("sa_server_poll: null argument");
    sparrow_amd::MoleculeResult r;
    *done = server->server->poll(ticket, &r) ? 1 : 0;
    if (*done) fillResult(r, result, P, nullptr, nullptr);
  });
}

int sa_server_warm_hits(sa_server* server, long long* hits) {
  return guard([&] {
    if (!server || !hits) throw std::runtime_error("sa_server_warm_hits: null argument");
    *hits = server->server->warmHits();
  });
}

}  // extern "C"
