// Device staging buffer of a handle for its host-array entry points: grown
// on demand and kept, so repeated calls of the same size allocate nothing.
// The mutex serialises host-array calls on one handle (they share it).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

namespace mf {

struct HostStage {
  std::mutex mu;
  void *dev = nullptr;
  size_t bytes = 0;
  HostStage() = default;
  HostStage(const HostStage &) = delete;
  HostStage &operator=(const HostStage &) = delete;
  ~HostStage() {
    if (dev) (void)hipFree(dev);
  }
  // at least `need` bytes of device memory (the caller holds mu)
  hipError_t reserve(size_t need) {
    if (need <= bytes) return hipSuccess;
    if (dev) {
      const hipError_t e = hipFree(dev);
      dev = nullptr;
      bytes = 0;
      if (e != hipSuccess) return e;
    }
    const hipError_t e = hipMalloc(&dev, need);
    if (e == hipSuccess) bytes = need;
    return e;
  }
};

}  // namespace mf
