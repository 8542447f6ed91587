// Host lowering of a program and its noise channels into trajectory records.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/aleph_hip.h"
#include "engine.hpp"
#include "noise.hpp"

namespace aleph {

struct TrajHost {
  std::vector<TrajOp> ops;
  std::vector<double> reals;
  std::vector<uint64_t> masks;
  std::vector<uint32_t> index;
  std::vector<double> readout;   // n x 2 flip probabilities
  uint64_t readout_mask = 0;
};

// Validates channels and readout errors (throws Failure, kind
// "invalid-channel" or "unsupported") and lowers gates, diagonal phases and
// channels in program order.  Channels must be sorted by `after`.
TrajHost build_trajectories(const std::vector<ProgInstr>& prog, uint32_t n, const aleph_hip_channel* channels,
                            uint32_t n_channels, const double* readout, uint64_t readout_mask);

}  // namespace aleph
