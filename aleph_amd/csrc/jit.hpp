// Fused-pass kernels generated per pass at plan time.
//
// The compiled k_fused_pass reads each op's kind and register bits at run
// time, so its op loop is a chain of branches whose merges cost register
// copies of the whole chunk slice (about 96 64-bit moves per 4x4 op in FP64).
// A generated kernel has the pass's stages and ops written out as one
// straight-line body with every register bit and control flag fixed at
// compile time; it is compiled with hiprtc for the device's architecture,
// and identical pass structures share one module.  The matrices, masks and
// phase terms stay in the plan's device payload, so the source depends only
// on the pass's structure.
#pragma once

#include <string>

#include "engine.hpp"

namespace aleph {

// Source of the generated kernel (`aleph_pass_f64` / `aleph_pass_f32`) for
// fused pass `pass` of `p`.
std::string jit_pass_source(const Plan& p, uint32_t pass);

// Compile (or take from the process-wide cache) and load a kernel for every
// fused pass of `p` on the current device.  Passes with many phase-polynomial
// ops and passes whose chunk does not fit 64 KiB of static LDS keep the
// compiled kernel.
void jit_compile_plan(Plan& p);

// Launch step `i` of `p` (a fused pass that has a generated kernel).
void launch_jit_pass(hipStream_t s, const Plan& p, size_t i, void* state);

}  // namespace aleph
