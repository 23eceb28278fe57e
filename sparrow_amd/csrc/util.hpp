// Small host-side helpers shared by the engine sources.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <stdexcept>
#include <string>

#define HIPCHK(x)                                                                                               \
  do {                                                                                                          \
    hipError_t e_ = (x);                                                                                        \
    if (e_ != hipSuccess)                                                                                       \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " (" #x ") at " __FILE__ ":" + \
                               std::to_string(__LINE__));                                                       \
  } while (0)

namespace sparrow_amd {

inline double nowSeconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Number of charge distributions of an atom with shells up to l (s, sp, spd).
inline int nChargeDist(int l) { return l == 0 ? 1 : l == 1 ? 10 : 40; }

// Bump allocator over a device block, 256-byte aligned.  With base == nullptr it only counts
// (the returned pointers are not to be used), which gives the size of a layout.
struct Carver {
  char* base;
  size_t off = 0, cap = 0;
  template <class T>
  T* take(size_t count) {
    const size_t bytes = std::max<size_t>(count * sizeof(T), 8);
    off = (off + 255) & ~size_t(255);
    T* p = reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(base) + off);
    off += bytes;
    if (base && off > cap) throw std::runtime_error("device workspace too small");
    return p;
  }
};

// Device memory owned by one scope (freed with it; hipFree waits for the device).
struct DeviceBuffer {
  char* ptr = nullptr;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() {
    if (ptr) (void)hipFree(ptr);
  }
  void allocate(size_t bytes) {
    if (ptr) {
      HIPCHK(hipFree(ptr));
      ptr = nullptr;
    }
    HIPCHK(hipMalloc(reinterpret_cast<void**>(&ptr), bytes));
  }
};

// Page-locked host array (device-to-host copies into it are true asynchronous DMA and may be
// recorded in a graph).
template <class T>
class PinnedBuf {
 public:
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p_) (void)hipHostFree(p_);
  }
  void assign(size_t n, const T& v) {
    if (n != n_) {
      if (p_) HIPCHK(hipHostFree(p_));
      p_ = nullptr;
      n_ = n;
      if (n) HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault));
    }
    for (size_t i = 0; i < n; ++i) p_[i] = v;
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  T& operator[](size_t i) { return p_[i]; }
  const T& operator[](size_t i) const { return p_[i]; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace sparrow_amd
