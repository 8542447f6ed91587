// Host engine: program decoding, gate lowering, the rule-based pass planner
// and plan execution.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "gates.hpp"
#include "plan.hpp"

namespace aleph {

// One decoded program instruction.
struct ProgInstr {
  enum Type { GATE, DIAG_PHASE, BARRIER } type = GATE;
  DenseGate gate;                       // GATE
  std::vector<double> angles;           // DIAG_PHASE terms
  std::vector<std::vector<uint64_t>> conds;
};

// Parse the flat program encoding of aleph_hip_plan_create and validate
// every gate against an n-qubit state (throws Failure).
std::vector<ProgInstr> decode_program(const uint32_t* words, size_t n_words, const double* reals, size_t n_reals,
                                      uint32_t n);

// A gate lowered for the fused executor.
struct LGate {
  enum Kind { G1Q, G2Q, TERMS, SIGN, DIAGX, KQ, POLY } kind = G1Q;
  std::vector<uint32_t> nd;             // qubits acted on non-diagonally (must sit in registers)
  std::vector<uint32_t> dg;             // qubits acted on diagonally (controls, diagonal targets)
  uint32_t qa = 0, qb = 0;              // G1Q: qa; G2Q: qa = matrix MSB, qb = LSB; DIAGX: targets
  uint32_t n_diag_targets = 0;          // DIAGX: 1 or 2
  uint64_t cmask = 0;                   // controls
  uint64_t sign_mask = 0;               // SIGN
  double m[32] = {0};                   // G1Q/G2Q matrix; DIAGX entries
  std::vector<std::pair<uint64_t, double>> terms;  // TERMS: AND-mask phase terms
  // standalone sweeps
  uint32_t k = 0;
  std::vector<uint32_t> kq_qubits;      // KQ, MSB first
  std::vector<double> kq_mat;           // KQ interleaved row-major
  std::vector<double> poly_angles;      // POLY
  std::vector<uint64_t> poly_conds;
  std::vector<uint32_t> poly_off;
  double cost = 0.0;
  // The original uncontrolled 1- or 2-qubit unitary (dense_k = 1 or 2,
  // qubits MSB first), kept so neighbouring gates can be fused into one
  // 4x4 block; dense_k = 0 when the gate is controlled or wider.
  uint32_t dense_k = 0;
  uint32_t dense_q[2] = {0, 0};
  double dense[32] = {0};
  bool standalone() const { return kind == KQ || kind == POLY; }
};

std::vector<LGate> lower_program(const std::vector<ProgInstr>& prog, uint32_t n);

struct PlanOptions {
  uint32_t planner = ALEPH_HIP_PLANNER_RULE;
  uint32_t local_qubits = 0;
  uint32_t max_ops = 0;
  uint32_t coalesce_bits = 0;
  bool fuse_1q = true;
  uint32_t reg_bits = 0;      // 0: default (4)
  bool nontemporal = false;
  bool jit = false;           // generate and compile one kernel per fused pass at plan creation
  bool fuse_2q = true;        // fuse 1q neighbours into 2q gates (4x4 blocks)
  bool fuse_diag = true;      // fuse {diagonal, CNOT} runs into one DiagonalPhase
  uint32_t kq_width = 0;      // >= 3: fuse gates into dense blocks of up to this many qubits (standalone sweeps)
  bool renormalize = false;   // after the plan: FP64 norm, then psi /= sqrt(norm) on the device
  bool generic_phase = false; // generated kernels keep phase ops as the run-time term loop
};

// One device launch of a plan.
// Standard-layout so the host plan view (aleph_hip_plan_view_get) can expose it.
struct Step {
  enum Kind : uint32_t { FUSED, KQ, DIAG, POLY, G1Q, G2Q, SWAPQ, BITPERM };
  uint32_t kind = FUSED;
  uint32_t pass = 0;          // FUSED: index into passes
  uint32_t m = 0, n_stages = 0;
  // per-gate launches
  uint32_t k = 0;
  uint32_t qubits[8] = {0};
  uint64_t cmask = 0;
  double mat8[32] = {0};
  size_t dev_off = 0;         // offset of the step's device payload (bytes) in Plan::payload
  size_t dev_off2 = 0, dev_off3 = 0;
  uint32_t count = 0;
};

struct Plan {
  uint32_t n = 0;
  uint32_t precision = ALEPH_HIP_F64;
  PlanOptions opts;
  std::vector<PassDev> passes;
  std::vector<StageDev> stages;
  std::vector<DevOp> ops;
  std::vector<PhaseTerm> terms;
  std::vector<unsigned char> payload;  // standalone matrices / CSR, 256-byte aligned records
  std::vector<Step> steps;
  // device mirrors
  PassDev* d_passes = nullptr;
  StageDev* d_stages = nullptr;
  DevOp* d_ops = nullptr;
  PhaseTerm* d_terms = nullptr;
  unsigned char* d_payload = nullptr;
  aleph_hip_plan_stats stats{};
  bool uploaded = false;
  // Generated kernel per pass (null: the compiled k_fused_pass), filled by jit_compile_plan.
  std::vector<hipFunction_t> jit;
  ~Plan();
  void upload(hipStream_t s);
  void execute(hipStream_t s, void* state) const;
  // Records events[i] before launch i and events[n] after the last one.
  void execute_with_events(hipStream_t s, void* state, hipEvent_t* events) const;
};

std::unique_ptr<Plan> build_plan(const std::vector<ProgInstr>& prog, uint32_t n, uint32_t precision,
                                 const PlanOptions& opts);

// Greedy fusion of gates into dense blocks of at most `width` qubits.
std::vector<ProgInstr> fuse_dense_blocks(const std::vector<ProgInstr>& prog, uint32_t width);

// The cost model predicts ln(µs) of a fused launch scaled to a state of
// 2^kCostRefQubits amplitudes (a pass streams the whole state, so its time
// is proportional to 2^n at the sizes that matter); the planner multiplies
// by 2^(n - kCostRefQubits).
constexpr int kCostRefQubits = 24;
// Cost-model features of every fused-pass launch of `p` (kPlanFeatures
// floats per row, launch order); `launch` receives each row's step index.
constexpr int kPlanFeatures = 16;
void plan_features(const Plan& p, std::vector<float>& rows, std::vector<uint32_t>& launch);
// Estimated microseconds of the plan's launches that are not fused passes
// (each streams the state once at an assumed 4 TB/s).
double plan_standalone_us(const Plan& p);

// Apply one validated gate with the per-gate kernels (Backend::apply_gate).
// `scratch` must hold at least 2 * 4^6 doubles of device memory.
void apply_dense_gate(hipStream_t s, void* state, uint32_t n, uint32_t precision, const DenseGate& g,
                      void* scratch);

}  // namespace aleph
