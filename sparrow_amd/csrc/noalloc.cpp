// Device scratch memory of the allocation-free entry point (sa_fock_noalloc).
//
// `scratch` is one statically sized array in the library's device code object: every GPU the
// library is loaded on holds its own copy from module load until the process ends, so a call
// that works inside it never calls hipMalloc / hipFree.  Host code never uses the array's name as
// an address (on the host it is only a stand-in symbol): its device address, different on each
// GPU, is looked up once per device with hipGetSymbolAddress when a context is created on that
// device (sa_create, with the device current) and kept per device.  All calls on one GPU share
// the array, which keeps the previous call's contents; calls that overlap on one GPU (two
// contexts on the same device used from two threads) race on it.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "engine.hpp"

namespace sparrow_amd {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex gScratchMutex;
std::map<int, char*> gScratchAddress;  // device -> address of that device's copy of `scratch`
}  // namespace

char* scratchAddress(int device) {
  std::lock_guard<std::mutex> lock(gScratchMutex);
  auto it = gScratchAddress.find(device);
  if (it != gScratchAddress.end()) return it->second;
  int current = -1;
  if (hipGetDevice(&current) != hipSuccess || current != device)
    throw std::runtime_error("scratch lookup: device " + std::to_string(device) + " is not the current device");
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw std::runtime_error(std::string("hipGetSymbolAddress of the device scratch array failed: ") + hipGetErrorString(e));
  gScratchAddress[device] = static_cast<char*>(p);
  return static_cast<char*>(p);
}

}  // namespace sparrow_amd
