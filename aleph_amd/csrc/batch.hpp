// Host side of batched execution: many small circuits of uneven qubit
// count, each planned into fused passes, packed into one set of device
// arrays and run by one persistent launch (or one launch per pass round),
// captured once into a hipGraph and replayed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "engine.hpp"
#include "kernels.hpp"

namespace aleph {

enum BatchMode : uint32_t { BATCH_PERSISTENT = 0, BATCH_ROUNDS = 1 };

struct BatchConfig {
  uint32_t mode = BATCH_PERSISTENT;
  uint32_t local_qubits = 0;     // chunk qubits per pass (0: 12)
  uint32_t chunks_per_item = 0;  // 0: 16
  uint32_t groups_per_cu = 0;    // persistent workgroups per CU (0: occupancy limit)
  uint32_t max_ops = 0;          // pass budget (0: planner default)
  uint32_t spin_limit = 0;       // dependency wait bound (0: default)
  bool nontemporal = false;
  bool unreachable_deps = false; // testing aid: no dependency wait can succeed
  uint32_t solo_qubits = 0;      // instances of at least this many qubits run as their own plans
                                 // with generated pass kernels (0: kDefaultSoloQubits; > 40: never)
};

// Instances this large fill the device by themselves (2^9 workgroups per
// pass and more): they run as their own plans with kernels generated per
// pass, which stream a pass 1.5-2x faster than the batch executors' op loop.
constexpr uint32_t kDefaultSoloQubits = 21;

// A QAOA MaxCut instance kept as its graph and angles; its program is
// rebuilt whenever the batch is planned.
struct QaoaSpec {
  std::vector<uint32_t> edges;   // pairs u, v
  std::vector<double> gammas, betas;
};

struct Batch {
  uint32_t precision = ALEPH_HIP_F32;
  std::vector<uint32_t> nq;
  std::vector<std::vector<ProgInstr>> progs;   // generic instances, or
  std::vector<QaoaSpec> qaoa;                  // QAOA instances
  BatchConfig cfg;

  uint32_t size() const { return (uint32_t)nq.size(); }

  // packed plans
  std::vector<PassDev> passes;
  std::vector<StageDev> stages;
  std::vector<DevOp> ops;
  std::vector<PhaseTerm> terms;
  std::vector<BatchItem> items;
  std::vector<uint32_t> round_begin;   // items of pass round r: [round_begin[r], round_begin[r+1])
  std::vector<uint64_t> offs;          // first amplitude of each instance
  // Qubits whose first instruction was an uncontrolled H, removed from the
  // plan: the reset prepares them in |+> (every amplitude with no bit set
  // outside the mask equals 2^(-|mask|/2), the rest 0).
  std::vector<uint64_t> plus_mask;
  std::vector<uint32_t> inst_passes;   // passes per instance
  // Instances run as their own plans (generated kernels), after the packed ones.
  std::vector<std::unique_ptr<Plan>> solo;
  std::vector<uint32_t> solo_inst;
  uint64_t solo_generation = 0;        // bumped whenever plan() replaces the solo plans
  uint64_t total_amps = 0;
  uint32_t total_passes = 0;
  double plan_ms = 0.0;
  double sweep_bytes = 0.0;            // sum over passes of 2 * 2^n * bytes per amplitude
  BatchLaunch launch;
  uint32_t grid = 0;

  // device
  void* pool = nullptr;
  size_t pool_bytes = 0;
  PassDev* d_passes = nullptr;
  StageDev* d_stages = nullptr;
  DevOp* d_ops = nullptr;
  PhaseTerm* d_terms = nullptr;
  BatchItem* d_items = nullptr;
  uint64_t* d_offs = nullptr;
  BatchPrep* d_prep = nullptr;
  uint32_t* d_done = nullptr;
  BatchCtl* d_ctl = nullptr;
  std::vector<BatchPrep> prep;
  size_t cap_passes = 0, cap_stages = 0, cap_ops = 0, cap_terms = 0, cap_items = 0, cap_prep = 0;
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_src = nullptr;
  // What the captured graph baked in: launch geometry, per-round item
  // ranges and device pointers.  A replay with any of them changed would
  // run the old launches over the new plans, so run() re-captures instead.
  struct GraphKey {
    uint32_t mode = 0, grid = 0, threads = 0, n_items = 0;
    uint64_t solo_generation = 0;
    size_t lds_bytes = 0;
    bool nontemporal = false;
    std::vector<uint32_t> round_begin;
    std::vector<const void*> ptrs;
    bool operator==(const GraphKey& o) const {
      return mode == o.mode && grid == o.grid && threads == o.threads && n_items == o.n_items &&
             solo_generation == o.solo_generation &&
             lds_bytes == o.lds_bytes && nontemporal == o.nontemporal && round_begin == o.round_begin &&
             ptrs == o.ptrs;
    }
  };
  GraphKey graph_key;
  GraphKey current_key() const;

  ~Batch();
  // Plan every instance with `cfg` and pack the plans (host only).
  void plan();
  // Allocate the pool (first call) and upload the packed plans.
  void upload(hipStream_t s, int device);
  // Initialise every state to |0..0> and run all plans (graph replay).
  void run(hipStream_t s);
  // Error flag of the last persistent run (reads device memory; syncs).
  bool failed(hipStream_t s);
  void drop_graph();
};

// Remove every uncontrolled H that is the first instruction on its qubit
// (barriers aside) and return the mask of those qubits.  On |0..0> these
// gates only prepare |+> on the masked qubits, which the batch reset writes
// directly, saving the passes that applied them.
uint64_t strip_leading_hadamards(std::vector<ProgInstr>& prog, uint32_t n);

// Build the program of one QAOA MaxCut instance: H on every qubit, then per
// layer CNOT(u,v) Rz(gamma)(v) CNOT(u,v) on every edge and Rx(beta) on
// every qubit (the gate sequence of the seeded workload builder).
std::vector<ProgInstr> qaoa_program(uint32_t n, const uint32_t* edges, uint32_t n_edges, uint32_t p,
                                    const double* gammas, const double* betas);

// MaxCut descriptor of one instance (neighbour masks and the 4-qubit table).
CutDesc cut_desc(uint32_t n, const uint32_t* edges, uint32_t n_edges);

}  // namespace aleph
