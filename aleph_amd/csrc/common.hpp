// Shared device/host definitions for the gfx950 state-vector backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/aleph_hip.h"

namespace aleph {

// Amplitude storage: interleaved (re, im), 16-byte aligned for FP64 so a
// lane moves one amplitude with a single dwordx4 access (8-byte for FP32).
template <typename T>
struct alignas(2 * sizeof(T)) C2 {
  T x, y;
};

template <typename T>
__host__ __device__ inline C2<T> cmul(C2<T> a, C2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Insert a zero bit at position p of x (bits at and above p shift up).
__host__ __device__ inline uint64_t insert_zero(uint64_t x, uint32_t p) {
  const uint64_t low = (1ull << p) - 1ull;
  return (x & low) | ((x & ~low) << 1);
}

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

}  // namespace aleph
