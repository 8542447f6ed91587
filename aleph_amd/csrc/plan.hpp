// Device-visible descriptors of an execution plan.
//
// A plan is a list of full-state sweeps ("passes").  A fused pass owns a
// set S of m physical qubits (the chunk qubits); each workgroup streams one
// chunk of 2^m amplitudes (all other index bits fixed) HBM -> registers,
// applies every op of the pass, and writes the chunk back once.  Inside a
// pass the chunk moves through "stages": in a stage every thread holds 2^G
// amplitudes in registers spanning G of the chunk qubits (the register
// qubits); dense gates act on register qubits only, diagonal ops and
// controls act on any qubit through the amplitude's global index.  Between
// stages the chunk is re-distributed through LDS.
#pragma once

#include <cstdint>

namespace aleph {

constexpr int kRegBits = 4;                 // default G: amplitudes per thread = 2^G
constexpr int kMaxRegBits = 5;              // G is 4 or 5
constexpr int kMaxLocalBits = 13;           // chunk <= 2^13 amplitudes
constexpr int kMaxThreadBits = 10;          // slots for thread-id bits (m - G <= 9)

enum OpType : uint32_t {
  OP_1Q = 1,     // 2x2 on register bit a; optional control mask (global index)
  OP_2Q = 2,     // 4x4 on register bits (a = matrix MSB, b = LSB), a > b
  OP_DIAG = 3,   // diagonal on up to two qubits given as global masks ma (MSB), mb
  OP_SIGN = 4,   // negate where (idx & mask) == mask; thread part in `cmask`, register part in `a`
  OP_PHASE = 5,  // multiply by exp(i*phi): terms[term_off, +n_terms) of PhaseSlots classes, then `b` general terms
};

struct alignas(16) DevOp {
  uint32_t type;
  uint32_t a, b;
  uint32_t n_terms;
  uint64_t cmask;     // OP_1Q/2Q/DIAG: controls (global); OP_SIGN: thread-side mask
  uint64_t ma, mb;    // OP_DIAG target masks
  uint32_t term_off;
  uint32_t pad;
  double m[32];       // OP_1Q: 2x2, OP_2Q: 4x4, OP_DIAG: 4 entries (interleaved complex)
};

// Accumulator slots of OP_PHASE for G register bits: 0 constant, then one
// per register bit (linear), then one per register pair (g < h).
template <int G>
struct PhaseSlots {
  static constexpr int kCount = 1 + G + G * (G - 1) / 2;
  static constexpr int linear(int g) { return 1 + g; }
  static constexpr int pair(int g, int h) { return 1 + G + g * G - g * (g + 1) / 2 + (h - g - 1); }
};

// Phase term: contributes `angle` when (thread_index & mask_out) == mask_out
// and (reg & rmask) == rmask.  `cls` is its PhaseSlots accumulator; terms
// on three or more register bits are evaluated per amplitude instead.
struct alignas(16) PhaseTerm {
  uint64_t mask_out;
  double angle;
  uint32_t cls;
  uint32_t rmask;
  double c, s;       // cos(angle), sin(angle): the term's unit factor
  uint64_t pad;
};

struct alignas(16) StageDev {
  uint32_t op_begin, op_end;
  uint32_t n_thread_bits;
  uint32_t pad;
  uint64_t tg[kMaxThreadBits];   // physical mask of thread-id bit j
  uint32_t tl[kMaxThreadBits];   // chunk-local mask of thread-id bit j
  uint32_t pad2[2];
  uint64_t rg[kMaxRegBits];      // physical mask of register bit g
  uint32_t rl[kMaxRegBits];      // chunk-local mask of register bit g
  uint32_t pad3;
  // Uniform parts of the addresses of register slot r, precomputed: the
  // global offset (OR of rg over the bits of r) and the swizzled LDS slot
  // of the register part (swz of the OR of rl), so a stage's address maps
  // cost the kernel one scalar load instead of per-slot arithmetic.
  uint64_t go[1 << kMaxRegBits];
  uint32_t ls[1 << kMaxRegBits];
  uint32_t lb[1 << kMaxRegBits];   // ls[r] * 8: byte offset in an LDS array of 8-byte scalars
};

// Op families a pass uses; the launcher picks the kernel variant compiled
// for exactly those, keeping register pressure of unused paths out.
constexpr uint32_t kFeat2Q = 1;
constexpr uint32_t kFeatPhase = 2;
constexpr uint32_t kFeatNontemporal = 4;   // stream the chunk with non-temporal loads/stores

struct alignas(16) PassDev {
  uint32_t m;                    // chunk qubits
  uint32_t n_stages;
  uint32_t stage_begin;
  uint32_t features;             // kFeat* bits
  uint32_t local_pos[14];        // ascending physical positions of chunk qubits
  uint32_t reg_bits;             // G of this pass
  uint32_t pad2;
};

// ---- batched execution of many small plans ----

// One unit of work of a batch: chunks [chunk_begin, chunk_begin +
// chunk_count) of one fused pass of one instance.  The instance's done
// counter (chunks finished over all its passes) must reach dep_target
// before the item may start; the queue lists every pass's items after the
// items of the instance's earlier passes.
struct alignas(16) BatchItem {
  uint64_t state_off;            // first amplitude of the instance in the pool
  uint32_t pass;                 // global PassDev index
  uint32_t inst;
  uint32_t chunk_begin, chunk_count;
  uint32_t dep_target;
  uint32_t pad;
};

// Initial state of one instance: amplitude x of the 2^n at state_off is
// amp when x has no bit outside plus (|0..0> for plus = 0), else 0.
struct alignas(16) BatchPrep {
  uint64_t state_off;
  uint64_t plus;
  double amp;
  uint32_t n;
  uint32_t pad;
};

// Queue head and failure flag of the persistent executor.
struct alignas(16) BatchCtl {
  uint32_t head;
  uint32_t error;                // set when a dependency wait exceeded its spin limit
  uint32_t pad[2];
};

// MaxCut cost of one instance: adj[u] = neighbour mask of qubit u (n <= 32);
// ll[r] = cut among edges inside qubits 0..3 for the 16 settings r of
// those bits, 4 bits each.
struct alignas(16) CutDesc {
  uint64_t state_off;
  uint32_t n;
  uint32_t block_begin;          // first partial-sum slot of this instance
  uint64_t ll;
  uint32_t adj[32];
};

}  // namespace aleph
