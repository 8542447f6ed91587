// Device-side value types shared by the compiled kernels and by the pass
// kernels generated and compiled at plan time (which see only this file,
// plan.hpp and fused_chunk.hpp).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <cstdint>

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

}  // namespace aleph
