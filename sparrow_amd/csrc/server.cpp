// Serving loop: continuous batching of SCF requests on one GPU (see server.hpp).
#include "server.hpp"

#include <stdexcept>

#include "util.hpp"

namespace sparrow_amd {

namespace {

// Pool capacities for a matrix pool of `mat` elements.  Ratios per n^2 element, for molecules
// of C/H/N/O/S with about 2.5 AOs per atom: atom pairs ~ n^2/12.5 (capacity n^2/6), integrals
// ~ 30 doubles per pair (capacity 4 n^2), per-pair vectors ~ 11 per pair (capacity 2 n^2); the
// split-BF16 copies use ldp^2 >= n^2 (capacity 2 n^2).
WorkspaceCaps capsForMat(long long mat, int maxMols, int slots, bool mixed) {
  WorkspaceCaps c;
  c.mols = maxMols;
  c.mat = mat;
  c.hist = (long long)slots * mat;
  c.lowp = mixed ? 2 * mat : 0;
  c.ints = 4 * mat;
  c.pvec = 2 * mat;
  c.pairs = mat / 6 + 4096;
  c.atoms = mat / 100 + 8192;
  c.vecN = mat / 20 + 8192;
  return c;
}

}  // namespace

Server::Server(Engine& engine, const std::string& method, const ScfSettings& st, double hbmBudgetBytes, int maxMolecules)
    : eng_(engine), ps_(engine.method(method)), st_(st) {
  if (maxMolecules < 1) throw std::runtime_error("max_molecules must be >= 1");
  if (hbmBudgetBytes < 64.0 * (1 << 20)) throw std::runtime_error("HBM budget below 64 MiB");
  const int slots = std::max(2, st.diisSize);
  const bool mixed = st.precision == 1;
  // 7/8 of the budget for the workspace, 1/8 for the warm-start density cache
  const double wsBudget = hbmBudgetBytes * 0.875;
  long long lo = 1 << 12, hi = (long long)(wsBudget / 8);
  if (Workspace::bytesFor(capsForMat(lo, maxMolecules, slots, mixed), slots) > wsBudget)
    throw std::runtime_error("HBM budget too small for the requested number of molecule slots");
  while (lo < hi) {  // largest matrix pool that fits the budget (bytes grow linearly with it)
    const long long mid = lo + (hi - lo + 1) / 2;
    if (Workspace::bytesFor(capsForMat(mid, maxMolecules, slots, mixed), slots) <= wsBudget) lo = mid;
    else hi = mid - 1;
  }
  const WorkspaceCaps caps = capsForMat(lo, maxMolecules, slots, mixed);
  const size_t bytes = Workspace::bytesFor(caps, slots);
  arena_.reserve(bytes);
  ws_ = std::make_unique<Workspace>(eng_, ps_, st_, caps, arena_.base(), arena_.capacity());
  const long long cacheDoubles = (long long)((hbmBudgetBytes - wsBudget) / 8);
  cacheArena_.reserve((size_t)cacheDoubles * 8);
  cacheAlloc_.reset(cacheDoubles);
}

Server::~Server() = default;

long long Server::submit(const MoleculeInput& in, long long warmKey) {
  const MolDemand d = ws_->demand(in);  // throws for invalid input
  (void)d;
  Request r;
  r.ticket = nextTicket_++;
  r.in = in;
  r.in.densityGuess = nullptr;
  r.warmKey = warmKey;
  queue_.push_back(std::move(r));
  return queue_.back().ticket;
}

const double* Server::cacheFind(long long key, int n) {
  auto it = cache_.find(key);
  if (it == cache_.end() || it->second.n != n) return nullptr;
  lru_.erase(it->second.lru);
  lru_.push_front(key);
  it->second.lru = lru_.begin();
  return reinterpret_cast<const double*>(cacheArena_.base()) + it->second.off;
}

void Server::cacheStore(long long key, const double* devP, int n) {
  const long long len = (long long)n * n;
  if (len > cacheAlloc_.capacity()) return;
  auto old = cache_.find(key);
  if (old != cache_.end()) {
    cacheAlloc_.release(old->second.off, (long long)old->second.n * old->second.n);
    lru_.erase(old->second.lru);
    cache_.erase(old);
  }
  long long off = 0;
  while (!cacheAlloc_.alloc(len, &off)) {  // drop least recently used entries until it fits
    if (lru_.empty()) return;
    const long long victim = lru_.back();
    lru_.pop_back();
    auto v = cache_.find(victim);
    cacheAlloc_.release(v->second.off, (long long)v->second.n * v->second.n);
    cache_.erase(v);
  }
  double* dst = reinterpret_cast<double*>(cacheArena_.base()) + off;
  HIPCHK(hipMemcpyAsync(dst, devP, (size_t)len * sizeof(double), hipMemcpyDeviceToDevice, eng_.stream()));
  lru_.push_front(key);
  cache_[key] = CacheEntry{off, n, lru_.begin()};
}

int Server::step() {
  HIPCHK(hipSetDevice(eng_.device()));
  // admission in arrival order while the pools have room
  while (!queue_.empty()) {
    Request& r = queue_.front();
    const double* warm = nullptr;
    if (r.warmKey >= 0) warm = cacheFind(r.warmKey, ws_->demand(r.in).n);
    const int slot = warm ? ws_->admitWarm(r.in, warm) : ws_->admit(r.in);
    if (slot < 0) {
      if (!ws_->busy())
        throw std::runtime_error("request " + std::to_string(r.ticket) + " does not fit the serving budget");
      break;
    }
    if (warm) ++warmHits_;
    slotOf_[slot] = {r.ticket, r.warmKey};
    queue_.pop_front();
  }
  if (!ws_->busy()) return 0;
  ws_->iterate();
  std::vector<int> fin = ws_->finishedSlots();
  if (!fin.empty()) {
    for (int slot : fin) {
      const long long key = slotOf_.at(slot).second;
      if (key >= 0) cacheStore(key, ws_->densityOf(slot), ws_->nao(slot));
    }
    std::vector<MoleculeResult> res = ws_->collect(fin, false, false);
    for (size_t k = 0; k < fin.size(); ++k) {
      done_[slotOf_.at(fin[k]).first] = std::move(res[k]);
      slotOf_.erase(fin[k]);
    }
  }
  return onGpu();
}

bool Server::poll(long long ticket, MoleculeResult* out) {
  auto it = done_.find(ticket);
  if (it == done_.end()) return false;
  if (out) *out = std::move(it->second);
  done_.erase(it);
  return true;
}

}  // namespace sparrow_amd
