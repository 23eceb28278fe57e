// Serving loop: continuous batching of SCF requests on one GPU.
//
// Requests are queued on submit().  Each step() admits queued requests in arrival order while
// the workspace pools (sized from an HBM budget) have room, runs one lock-step SCF iteration
// for every molecule on the GPU (newly admitted ones start with their set-up), and moves the
// molecules that finished into the result table.  A request may name a warm-start key: the final
// density of a finished request with that key stays on the device (least recently used entries
// are dropped when the cache budget is full) and is the starting density of the next request
// with the same key and basis size.
#pragma once
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <unordered_map>

#include "workspace.hpp"

namespace sparrow_amd {

class Server {
 public:
  Server(Engine& engine, const std::string& method, const ScfSettings& st, double hbmBudgetBytes, int maxMolecules);
  ~Server();
  // Validates the input (throws std::runtime_error) and queues it; returns a ticket.
  long long submit(const MoleculeInput& in, long long warmKey);
  // Admission + one iteration; returns the number of molecules on the GPU after it.
  int step();
  // If the request has finished, moves its result into *out and returns true.
  bool poll(long long ticket, MoleculeResult* out);
  int queued() const { return (int)queue_.size(); }
  int onGpu() const { return (int)slotOf_.size(); }
  long long warmHits() const { return warmHits_; }

 private:
  struct Request {
    long long ticket;
    MoleculeInput in;
    std::vector<int> Z;
    std::vector<double> xyz;
    long long warmKey;
  };
  struct CacheEntry {
    long long off;  // doubles, in the cache pool
    int n;
    std::list<long long>::iterator lru;
  };
  void cacheStore(long long key, const double* devP, int n);
  const double* cacheFind(long long key, int n);

  Engine& eng_;
  ParameterSet& ps_;
  ScfSettings st_;
  DeviceArena arena_, cacheArena_;
  std::unique_ptr<Workspace> ws_;
  std::deque<Request> queue_;
  std::map<int, std::pair<long long, long long>> slotOf_;  // slot -> (ticket, warm key)
  std::unordered_map<long long, MoleculeResult> done_;
  RangeAlloc cacheAlloc_;
  std::unordered_map<long long, CacheEntry> cache_;
  std::list<long long> lru_;  // most recent first
  long long nextTicket_ = 1, warmHits_ = 0;
};

}  // namespace sparrow_amd
