// Shared device/host definitions for the gfx950 state-vector backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

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

}  // namespace aleph
