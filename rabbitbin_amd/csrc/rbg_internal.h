// Internal helpers shared by the librbg translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace rbg {

// Owning device allocation that only grows (reused across pipeline runs).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  hipError_t ensure(size_t nbytes);
  void release();
};

// Takes the device address of the no-allocation entry point's scratch array
// on `device` (current), once per device (rbg_noalloc.hip).
hipError_t noalloc_setup(int device);

}  // namespace rbg
