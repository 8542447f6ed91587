// Monte-Carlo noise trajectories on the device (quantum-jump unravelling).
//
// Semantics follow crates/aleph-sv/src/noise: after every gate the channels
// attached to it fire in order (execution.rs:102-135); a Pauli channel picks
// one Pauli string from its fixed weights with one uniform draw and applies
// it (apply.rs:46-66); a one-qubit Kraus channel computes p_i = |K_i psi|^2,
// draws one uniform, applies K_i / sqrt(p_i) and leaves the state alone
// when p_i < 1e-300 (apply.rs:68-133); the terminal outcome is one
// cumulative-inversion draw (sampling.rs:61-91), then one draw per qubit
// with a readout error flips that bit with probability m[t][1-t]
// (apply.rs:135-158).  Every trajectory draws from its own counter-based
// stream keyed by the reference's splitmix64 shot seed (apply.rs:160-180),
// so shot s's outcome depends only on (seed, s), never on how shots are
// spread over workgroups or calls.
//
// Device layout: one workgroup owns one trajectory at a time.  Up to 13
// qubits (FP64, 128 KiB) the state lives in LDS and never touches HBM;
// wider states live in a per-workgroup slice of a device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace aleph {

// One trajectory-program record.
struct TrajOp {
  enum Kind : uint32_t { DENSE = 0, DIAG = 1, PHASE = 2, PAULI = 3, KRAUS = 4 };
  uint32_t kind = DENSE;
  uint32_t k = 0;          // DENSE / DIAG: target count (DENSE <= 3, DIAG <= 6)
  uint32_t q[6] = {0};     // targets, q[0] = most significant bit of the matrix index
  uint32_t count = 0;      // PHASE: terms; PAULI / KRAUS: branches
  uint32_t off = 0;        // doubles into `reals`: DENSE matrix (2*4^k), DIAG entries (2*2^k),
                           // PHASE angles, PAULI weights, KRAUS matrices (8 per branch)
  uint32_t off2 = 0;       // PHASE: u64 conds into `masks`; PAULI: (x, z) mask pairs into `masks`
  uint32_t off3 = 0;       // PHASE: u32 CSR offsets (count + 1) into `index`
  uint64_t cmask = 0;      // DENSE / DIAG: control qubits
};
static_assert(sizeof(TrajOp) == 56, "TrajOp layout");

constexpr uint32_t kNoiseLdsMaxQubits = 13;     // 2^13 complex128 = 128 KiB of LDS
constexpr uint32_t kNoiseMaxQubits = 24;        // per-workgroup device slices beyond 13 qubits
constexpr uint32_t kNoiseMaxKraus = 64;         // branches of one Kraus channel
constexpr uint32_t kNoiseThreadsLds = 256;
constexpr uint32_t kNoiseThreadsWide = 1024;

struct TrajProgram {
  const TrajOp* ops = nullptr;
  uint32_t n_ops = 0;
  uint32_t n = 0;
  const double* reals = nullptr;
  const uint64_t* masks = nullptr;
  const uint32_t* index = nullptr;
  const double* readout = nullptr;   // n x 2: P(flip | true 0), P(flip | true 1)
  uint64_t readout_mask = 0;         // qubits with a readout error
  uint64_t seed = 0;
  uint64_t shot0 = 0;                // absolute index of the first shot of this launch
  uint32_t shots = 0;
};

// Runs `p.shots` trajectories; outcome of shot i -> out[i].  `slices` is
// null for n <= kNoiseLdsMaxQubits, else `workgroups` * 2^n complex128.
void launch_trajectories(hipStream_t s, const TrajProgram& p, uint32_t workgroups, void* slices, uint64_t* out);

// Workgroups and device bytes the launch for an n-qubit program uses
// (bytes = 0 when the state fits in LDS).  `budget` bounds the slices.
void trajectory_geometry(uint32_t n, uint32_t shots, size_t budget, uint32_t* workgroups, size_t* bytes);

}  // namespace aleph
