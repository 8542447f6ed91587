// Shared device/host definitions for the gfx950 state-vector backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <set>
#include <string>
#include <utility>

#include "../../include/aleph_hip.h"
#include "device_types.hpp"

namespace aleph {

// Thrown by host helpers; the C ABI converts it into a status code.
struct Failure {
  int32_t code;
  std::string message;
  uint32_t qubit = 0, num_qubits = 0, expected = 0, got = 0;
  double value = 0.0;
  std::string kind;
};

#define ALEPH_HIP_CHECK(expr)                                                         \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ::aleph::Failure f_{ALEPH_HIP_ERR_DEVICE_FAILURE,                               \
                          std::string(#expr) + ": " + hipGetErrorString(e_)};         \
      if (e_ == hipErrorOutOfMemory) f_.code = ALEPH_HIP_ERR_OUT_OF_MEMORY;           \
      throw f_;                                                                       \
    }                                                                                 \
  } while (0)

// Raises kernel `fn`'s dynamic LDS limit to the CU's 160 KiB, once per
// (device, kernel).  Thread-safe: plans are built and launched from several
// host threads, and a process may drive more than one device.
inline void allow_full_lds(const void* fn, int bytes = 160 * 1024) {
  static std::mutex mu;
  static std::set<std::pair<int, const void*>> done;
  int dev = 0;
  ALEPH_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if (!done.insert({dev, fn}).second) return;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) {
    done.erase({dev, fn});
    (void)hipGetLastError();  // a failed call is also the thread's last error: do not leak it into later checks
    ALEPH_HIP_CHECK(e);
  }
}

}  // namespace aleph
