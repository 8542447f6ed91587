/*
 * aleph_hip.h — C ABI of the MI355X (gfx950) dense state-vector backend.
 *
 * This is the boundary a Rust `aleph-hip` crate binds with `extern "C"`
 * to implement `aleph_backend::Backend` (crates/aleph-backend/src/lib.rs:86-163)
 * and a `CudaPlanBackend`-style plan hook (crates/aleph-cuda/src/execution.rs:197-208),
 * so that `--backend hip` / `backend="hip"` sits beside `sv` and `cuda`.
 * Every entry point takes plain pointers and sizes and returns an
 * `aleph_hip_status` whose values map one-to-one onto `BackendError`
 * variants (crates/aleph-backend/src/lib.rs:22-81).  Structured error
 * detail (qubit, limits, deviation, gate name) is read back with
 * aleph_hip_last_error().
 *
 * Conventions carried across the boundary (same as the reference):
 *  - amplitude index bit q is qubit q (qubit 0 = LSB), ADR 0004 as
 *    implemented by crates/aleph-sv/src/kernels/aos.rs:1-8;
 *  - multi-qubit matrices are row-major with qubits[0] as the MSB of the
 *    row/column index (crates/aleph-core/src/gate/kinds.rs:74-96);
 *  - complex numbers are interleaved (re, im) doubles on the host side;
 *  - one caller at a time per context (Backend methods take &mut self);
 *    launches are stream-ordered, readout calls synchronise;
 *  - randomness lives on the host: the context owns a seeded generator
 *    and draws one f64 per sampled shot and one f64 per genuine
 *    (non-degenerate) measurement, the consumption contract of
 *    crates/aleph-cuda/src/sv/sampling.rs:209 and
 *    crates/aleph-sv/src/measure.rs:79-90.
 */
#ifndef ALEPH_HIP_H
#define ALEPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (BackendError, crates/aleph-backend/src/lib.rs:22-81) ---- */
typedef int32_t aleph_hip_status;
enum {
  ALEPH_HIP_OK = 0,
  ALEPH_HIP_ERR_DEVICE_FAILURE = 1,         /* DeviceFailure(String)            */
  ALEPH_HIP_ERR_QUBIT_OUT_OF_RANGE = 2,     /* QubitOutOfRange{qubit,num_qubits} */
  ALEPH_HIP_ERR_DUPLICATE_QUBIT = 3,        /* DuplicateQubit{qubit}            */
  ALEPH_HIP_ERR_ARITY_MISMATCH = 4,         /* ArityMismatch{kind,expected,got} */
  ALEPH_HIP_ERR_UNSUPPORTED_GATE = 5,       /* UnsupportedGate{kind}            */
  ALEPH_HIP_ERR_UNSUPPORTED_INSTRUCTION = 6,/* UnsupportedInstruction{kind}     */
  ALEPH_HIP_ERR_SYMBOLIC_PARAM = 7,         /* SymbolicParam                    */
  ALEPH_HIP_ERR_NON_FINITE_PARAM = 8,       /* NonFiniteParam{kind}             */
  ALEPH_HIP_ERR_NON_UNITARY_MATRIX = 9,     /* NonUnitaryMatrix{deviation}      */
  ALEPH_HIP_ERR_EMPTY_CIRCUIT = 10,         /* EmptyCircuit                     */
  ALEPH_HIP_ERR_DEGENERATE_MEASUREMENT = 11,/* DegenerateMeasurement{qubit,p}   */
  ALEPH_HIP_ERR_TOO_MANY_QUBITS = 12,       /* TooManyQubits{requested,limit}   */
  ALEPH_HIP_ERR_INVALID_PAULI_STRING = 13,  /* InvalidPauliString{reason}       */
  ALEPH_HIP_ERR_INVALID_STATE = 14,         /* InvalidState{reason}             */
  ALEPH_HIP_ERR_OPTIMIZATION = 15,          /* Optimization(PassError)          */
  ALEPH_HIP_ERR_NO_DEVICE = 16,             /* aleph-cuda Error::NoDevice       */
  ALEPH_HIP_ERR_INVALID_ARGUMENT = 17,      /* null pointer / bad size: DeviceFailure */
  ALEPH_HIP_ERR_OUT_OF_MEMORY = 18          /* memory preflight rejection: DeviceFailure */
};

/* ---- gate kinds: order of the reference `Gate` enum (kinds.rs:13-115) ---- */
enum {
  ALEPH_HIP_GATE_H = 0, ALEPH_HIP_GATE_X, ALEPH_HIP_GATE_Y, ALEPH_HIP_GATE_Z,
  ALEPH_HIP_GATE_S, ALEPH_HIP_GATE_SDG, ALEPH_HIP_GATE_T, ALEPH_HIP_GATE_TDG,
  ALEPH_HIP_GATE_RX, ALEPH_HIP_GATE_RY, ALEPH_HIP_GATE_RZ, ALEPH_HIP_GATE_PHASE,
  ALEPH_HIP_GATE_U3,
  ALEPH_HIP_GATE_CNOT, ALEPH_HIP_GATE_CZ, ALEPH_HIP_GATE_SWAP, ALEPH_HIP_GATE_ISWAP,
  ALEPH_HIP_GATE_ISWAPDG, ALEPH_HIP_GATE_CRX, ALEPH_HIP_GATE_CRY, ALEPH_HIP_GATE_CRZ,
  ALEPH_HIP_GATE_TOFFOLI, ALEPH_HIP_GATE_CCZ,
  ALEPH_HIP_GATE_UNITARY1Q, ALEPH_HIP_GATE_UNITARY2Q, ALEPH_HIP_GATE_UNITARYKQ,
  ALEPH_HIP_GATE_UNITARY1QDIAG, ALEPH_HIP_GATE_UNITARY2QDIAG,
  ALEPH_HIP_GATE_COUNT
};

/* Program-only instruction records (Instruction enum, crates/aleph-ir/src/instruction.rs:13-36). */
enum {
  ALEPH_HIP_INSTR_DIAGONAL_PHASE = 64,
  ALEPH_HIP_INSTR_BARRIER = 65
};

enum { ALEPH_HIP_F64 = 64, ALEPH_HIP_F32 = 32 };

typedef struct aleph_hip_ctx aleph_hip_ctx;
typedef struct aleph_hip_state aleph_hip_state;
typedef struct aleph_hip_plan aleph_hip_plan;

typedef struct aleph_hip_error {
  int32_t code;
  uint32_t qubit;       /* QubitOutOfRange / DuplicateQubit / DegenerateMeasurement */
  uint32_t num_qubits;  /* QubitOutOfRange */
  uint32_t expected;    /* ArityMismatch.expected, TooManyQubits.limit */
  uint32_t got;         /* ArityMismatch.got, TooManyQubits.requested */
  double value;         /* NonUnitaryMatrix.deviation, DegenerateMeasurement.probability */
  char kind[48];        /* gate / instruction name, or the InvalidState reason */
  char message[256];    /* rendered message, device errors included */
} aleph_hip_error;

/* One gate placed on qubits: GateInstance (crates/aleph-core/src/gate/instance.rs:40-59).
 * `params` holds the gate's angles (U3: theta, phi, lambda); `matrix` the
 * interleaved payload of Unitary1q/2q/Kq (dim*dim complex, row-major) or of
 * Unitary1qDiag/2qDiag (dim complex).  `k` is the width of UnitaryKq. */
typedef struct aleph_hip_gate {
  uint32_t kind;
  uint32_t n_qubits;
  uint32_t n_controls;
  uint32_t k;
  const uint32_t* qubits;
  const uint32_t* controls;
  double params[3];
  const double* matrix;
} aleph_hip_gate;

/* ---- context: replaces CudaSvBackend::with_seed (crates/aleph-cuda/src/sv/backend.rs:88-129) ---- */
aleph_hip_status aleph_hip_device_count(int32_t* count);
aleph_hip_status aleph_hip_ctx_create(int32_t ordinal, uint64_t seed, aleph_hip_ctx** out);
void aleph_hip_ctx_destroy(aleph_hip_ctx* ctx);
/* with_qubit_cap (crates/aleph-sv/src/backend.rs:46-50); default 32 for F64, 33 for F32 */
aleph_hip_status aleph_hip_ctx_set_qubit_cap(aleph_hip_ctx* ctx, uint32_t cap);
aleph_hip_status aleph_hip_ctx_reseed(aleph_hip_ctx* ctx, uint64_t seed);
aleph_hip_status aleph_hip_synchronize(aleph_hip_ctx* ctx);
aleph_hip_status aleph_hip_last_error(const aleph_hip_ctx* ctx, aleph_hip_error* out);
/* The hipStream_t every launch of this context is ordered on (for event timing). */
aleph_hip_status aleph_hip_ctx_stream(aleph_hip_ctx* ctx, void** stream);

/* ---- state: Backend::allocate (lib.rs:95) / CudaSvState (crates/aleph-cuda/src/sv/state.rs) ---- */
aleph_hip_status aleph_hip_state_alloc(aleph_hip_ctx* ctx, uint32_t num_qubits, uint32_t precision,
                                       aleph_hip_state** out);
void aleph_hip_state_free(aleph_hip_state* st);
/* Resets to |0...0>. */
aleph_hip_status aleph_hip_state_reset(aleph_hip_ctx* ctx, aleph_hip_state* st);
/* Host interleaved complex128 in / out (FP32 states narrow / widen). */
aleph_hip_status aleph_hip_state_load(aleph_hip_ctx* ctx, aleph_hip_state* st, const double* amps);
aleph_hip_status aleph_hip_state_read(aleph_hip_ctx* ctx, aleph_hip_state* st, double* out);
/* Device-resident copy in / out in the state's own precision. */
aleph_hip_status aleph_hip_state_load_device(aleph_hip_ctx* ctx, aleph_hip_state* st, const void* dptr);
aleph_hip_status aleph_hip_state_read_device(aleph_hip_ctx* ctx, aleph_hip_state* st, void* dptr);
aleph_hip_status aleph_hip_state_info(const aleph_hip_state* st, uint32_t* num_qubits, uint32_t* precision,
                                      void** device_ptr);

/* ---- gates: Backend::apply_gate (crates/aleph-sv/src/backend.rs:74-151) ---- */
aleph_hip_status aleph_hip_apply_gate(aleph_hip_ctx* ctx, aleph_hip_state* st, const aleph_hip_gate* g);
/* Backend::apply_diagonal_phase (lib.rs:128-138; crates/aleph-sv/src/kernels/diagonal_phase.rs:26-45).
 * CSR terms: term t fires when popcount(conds[c] & x) is odd for every
 * c in [offsets[t], offsets[t+1]), adding angles[t] to the phase. */
aleph_hip_status aleph_hip_apply_diagonal_phase(aleph_hip_ctx* ctx, aleph_hip_state* st, uint32_t n_terms,
                                                const double* angles, const uint64_t* conds,
                                                const uint32_t* offsets);
/* Backend::unpermute_state (lib.rs:156-163; crates/aleph-core/src/perm.rs:25-40), perm[logical] = physical. */
aleph_hip_status aleph_hip_unpermute(aleph_hip_ctx* ctx, aleph_hip_state* st, const uint32_t* perm);

/* Low-level kernels (CUDA peers in crates/aleph-cuda/src/sv/kernels.cu, diag.cu), no validation
 * beyond range checks.  Masks are u64. */
aleph_hip_status aleph_hip_apply_1q(aleph_hip_ctx* ctx, aleph_hip_state* st, uint32_t target,
                                    uint64_t ctrl_mask, const double* m /* 8 doubles */);
aleph_hip_status aleph_hip_apply_kq(aleph_hip_ctx* ctx, aleph_hip_state* st, uint32_t k, const uint32_t* qubits,
                                    uint64_t ctrl_mask, const double* m /* 2*4^k doubles */);
aleph_hip_status aleph_hip_apply_diag(aleph_hip_ctx* ctx, aleph_hip_state* st, uint32_t k, const uint32_t* qubits,
                                      uint64_t ctrl_mask, const double* d /* 2*2^k doubles */);
aleph_hip_status aleph_hip_apply_cnot(aleph_hip_ctx* ctx, aleph_hip_state* st, uint32_t control, uint32_t target);
aleph_hip_status aleph_hip_apply_swap(aleph_hip_ctx* ctx, aleph_hip_state* st, uint32_t a, uint32_t b);

/* ---- programs: CudaExecutionPlan::compile / execute (crates/aleph-cuda/src/execution.rs:78-194) ----
 * A program is a flat instruction stream.  `words` per record:
 *   gate:           kind, n_qubits, n_controls, k, qubits..., controls...
 *   diagonal phase: 64, n_terms, then per term: n_conds, (lo32, hi32) per cond
 *   barrier:        65, n, qubits...
 * `reals` per record, in the same order: the gate's params (1 for the
 * one-angle kinds, 3 for U3, none otherwise) followed by its matrix payload
 * (Unitary* kinds); a diagonal phase contributes one angle per term.
 */
enum { ALEPH_HIP_PLANNER_UNFUSED = 0, ALEPH_HIP_PLANNER_RULE = 1, ALEPH_HIP_PLANNER_LEARNED = 2 };

typedef struct aleph_hip_plan_opts {
  uint32_t planner;         /* ALEPH_HIP_PLANNER_* */
  uint32_t local_qubits;    /* qubits per LDS chunk (0 = 12) */
  uint32_t max_ops_per_pass;/* compute budget per pass, in 1q-gate equivalents (0 = default) */
  uint32_t coalesce_bits;   /* low qubits kept in every chunk for coalescing (0 = default, 4) */
  uint32_t fuse_1q;         /* merge adjacent 1q gates on one qubit before planning */
  uint32_t reg_bits;        /* amplitudes per thread = 2^reg_bits, 4 or 5 (0 = default 4) */
  uint32_t flags;           /* ALEPH_HIP_PLAN_* bits */
  uint32_t kq_width;        /* 3..6: fuse gates into dense blocks of up to this many qubits, each one
                               sweep (widths >= 4 on the matrix cores); 0 = off */
} aleph_hip_plan_opts;
enum {
  ALEPH_HIP_PLAN_NONTEMPORAL = 1, /* stream chunks with non-temporal loads/stores */
  ALEPH_HIP_PLAN_NO_FUSE_2Q = 2,  /* keep 1q neighbours out of 2q gates (no 4x4 blocks) */
  ALEPH_HIP_PLAN_NO_FUSE_DIAG = 4, /* keep {diagonal, CNOT} runs gate by gate (no DiagonalPhase fusion) */
  ALEPH_HIP_PLAN_JIT = 8,         /* fused passes run as kernels generated per pass (hiprtc at plan creation) */
  ALEPH_HIP_PLAN_RENORMALIZE = 16,/* after every execution: FP64 norm, then psi /= sqrt(norm), on the device
                                     (FP32 states drift ~1e-8 per pass; see aleph_hip_state_normalize) */
  ALEPH_HIP_PLAN_GENERIC_PHASE = 32 /* generated kernels keep phase ops as the run-time term loop, so that
                                     circuits differing only in their phase terms (QAOA instances of one
                                     size) share kernels */
};

typedef struct aleph_hip_plan_stats {
  double plan_ms;           /* host planning wall time */
  uint32_t input_gates;
  uint32_t n_passes;        /* full-state sweeps */
  uint32_t n_stages;        /* register stages summed over passes */
  uint32_t n_ops;           /* device ops after lowering */
  uint32_t n_launches;
  uint32_t planner_used;    /* ALEPH_HIP_PLANNER_* that produced the plan (LEARNED falls back to RULE without a model) */
  double sweep_bytes;       /* algorithmic HBM bytes over all passes */
  /* LEARNED: the candidate options the cost model chose and its prediction */
  uint32_t chosen_local_qubits, chosen_max_ops, chosen_flags, n_candidates;
  double predicted_us;
  uint32_t chosen_kq_width;  /* LEARNED: dense-block width of the chosen candidate (0 = LDS passes) */
  uint32_t chosen_rule;      /* LEARNED: 1 when the rule planner's own options were kept */
  double rule_predicted_us;  /* LEARNED: predicted time of the rule planner's plan */
  uint32_t cancelled_gates;  /* gates removed as adjacent exact-inverse pairs (CancelInversePairs) */
  uint32_t reserved2;
} aleph_hip_plan_stats;

aleph_hip_status aleph_hip_plan_create(aleph_hip_ctx* ctx, uint32_t num_qubits, uint32_t precision,
                                       const uint32_t* words, size_t n_words, const double* reals,
                                       size_t n_reals, const aleph_hip_plan_opts* opts, aleph_hip_plan** out);
aleph_hip_status aleph_hip_plan_execute(aleph_hip_ctx* ctx, aleph_hip_plan* plan, aleph_hip_state* st);
/* Execute with a hipEvent pair around every launch; ms[i] receives the
 * device time of launch i (n_launches entries, see aleph_hip_plan_stats),
 * kinds[i] its kind (0 fused pass, 1 dense block, 2 diagonal, 3 phase
 * polynomial, 4 one-qubit, 5 two-qubit, 6 swap, 7 trailing-SWAP permutation
 * sweep).  This is the timing
 * recorder the plan cost models train on. */
aleph_hip_status aleph_hip_plan_execute_timed(aleph_hip_ctx* ctx, aleph_hip_plan* plan, aleph_hip_state* st,
                                             double* ms, uint32_t* kinds);
aleph_hip_status aleph_hip_plan_stats_get(const aleph_hip_plan* plan, aleph_hip_plan_stats* out);
void aleph_hip_plan_free(aleph_hip_plan* plan);
/* Diagnostics: plan on the host only (no device needed; such a plan cannot
 * be executed) and read the plan's descriptor arrays, whose layouts are
 * those of aleph_amd/csrc/plan.hpp (PassDev, StageDev, DevOp, PhaseTerm)
 * and engine.hpp (Step).  Used by planner analysis and host-side tests. */
typedef struct aleph_hip_plan_view {
  const void* steps; uint32_t n_steps; uint32_t step_bytes;
  const void* passes; uint32_t n_passes; uint32_t pass_bytes;
  const void* stages; uint32_t n_stages; uint32_t stage_bytes;
  const void* ops; uint32_t n_ops; uint32_t op_bytes;
  const void* terms; uint32_t n_terms; uint32_t term_bytes;
  const void* payload; uint64_t payload_bytes;   /* standalone-step matrices / CSR */
} aleph_hip_plan_view;
aleph_hip_status aleph_hip_plan_build_host(uint32_t num_qubits, uint32_t precision, const uint32_t* words,
                                           size_t n_words, const double* reals, size_t n_reals,
                                           const aleph_hip_plan_opts* opts, aleph_hip_plan** out,
                                           aleph_hip_error* err);
aleph_hip_status aleph_hip_plan_view_get(const aleph_hip_plan* plan, aleph_hip_plan_view* out);
/* Diagnostics (no counterpart in the reference): the source of the kernel
 * ALEPH_HIP_PLAN_JIT generates for fused pass `pass` (works on host-only
 * plans).  *len receives the length; out (if cap > 0) a NUL-terminated copy. */
aleph_hip_status aleph_hip_plan_jit_source(const aleph_hip_plan* plan, uint32_t pass, char* out, size_t cap,
                                           size_t* len);
/* plan_create + execute + free in one call (run_with_outcomes without measurements, lib.rs:211-248). */
aleph_hip_status aleph_hip_run_program(aleph_hip_ctx* ctx, aleph_hip_state* st, const uint32_t* words,
                                       size_t n_words, const double* reals, size_t n_reals,
                                       const aleph_hip_plan_opts* opts, aleph_hip_plan_stats* stats);

/* ---- readout (crates/aleph-sv/src/measure.rs; crates/aleph-cuda/src/sv/readout.rs) ---- */
/* Σ|ψ|² accumulated in FP64 with compensation (no validation). */
aleph_hip_status aleph_hip_norm_sqr(aleph_hip_ctx* ctx, aleph_hip_state* st, double* out);
/* validate_state: finite amplitudes and |Σ|ψ|² − 1| <= sqrt(2^n)·1e-10 (measure.rs:23-39). */
aleph_hip_status aleph_hip_validate(aleph_hip_ctx* ctx, aleph_hip_state* st);
aleph_hip_status aleph_hip_measure(aleph_hip_ctx* ctx, aleph_hip_state* st, uint32_t qubit, int32_t* outcome);
aleph_hip_status aleph_hip_measure_with_uniform(aleph_hip_ctx* ctx, aleph_hip_state* st, uint32_t qubit,
                                                double u, int32_t* outcome, int32_t* consumed);
aleph_hip_status aleph_hip_sample(aleph_hip_ctx* ctx, aleph_hip_state* st, uint32_t shots, uint64_t* out);
aleph_hip_status aleph_hip_sample_with_uniforms(aleph_hip_ctx* ctx, aleph_hip_state* st, const double* uniforms,
                                                uint32_t shots, uint64_t* out);
/* No counterpart in the reference: aleph_hip_sample_with_uniforms for callers that cannot
 * allocate device memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and the same results, but the block sums, uniforms and shot
 * indices live in a device array of the library (17 MiB per device, set up by
 * aleph_hip_ctx_create and held until the process ends): no allocation in the call.  States
 * of at most 32 qubits and at most 65536 shots per call, else ALEPH_HIP_ERR_INVALID_ARGUMENT.
 * Every call on one device shares that array, so calls that overlap on one device (from
 * several contexts or threads) race. */
aleph_hip_status aleph_hip_sample_with_uniforms_noalloc(aleph_hip_ctx* ctx, aleph_hip_state* st,
                                                        const double* uniforms, uint32_t shots, uint64_t* out);
/* paulis[i]: 0 = I, 1 = X, 2 = Y, 3 = Z on qubits[i]. */
aleph_hip_status aleph_hip_expectation(aleph_hip_ctx* ctx, aleph_hip_state* st, double coefficient,
                                       const uint32_t* qubits, const uint8_t* paulis, uint32_t n_terms,
                                       double* out);
/* sum_t coeffs[t] * <psi|P_t|psi> for n_terms Pauli strings; term t acts on
 * qubits[offsets[t] .. offsets[t+1]) with paulis[...] (codes as above).
 * Replaces Backend::expectation_pauli_sum (crates/aleph-backend/src/lib.rs:422-432):
 * one validation; terms sharing an X mask are evaluated together, up to 16 per read of the
 * state (a diagonal Hamiltonian of T terms costs ceil(T / 16) reads). */
aleph_hip_status aleph_hip_expectation_sum(aleph_hip_ctx* ctx, aleph_hip_state* st, uint32_t n_terms,
                                           const double* coeffs, const uint32_t* offsets, const uint32_t* qubits,
                                           const uint8_t* paulis, double* out);
/* Linear cross-entropy benchmark of the exact state, 2^n * sum_x |psi_x|^4 - 1, accumulated in
 * FP64 in one read of the state: benches/src/lib.rs:29 linear_xeb (computed there on the host
 * from the amplitudes). */
/* FP32 path renormalisation (the FP64 accumulation of crates/aleph-cuda/src/sv/readout_f32.rs:1-10
 * applied to the norm): psi /= sqrt(sum |psi|^2) with the sum in FP64 on the device; the norm
 * before scaling goes to *norm_before when it is not NULL (then the call synchronises). */
aleph_hip_status aleph_hip_state_normalize(aleph_hip_ctx* ctx, aleph_hip_state* st, double* norm_before);
/* Two states of one size, any precisions: out = {Re <a|b>, Im <a|b>, sum |a-b|^2, max |a-b|},
 * accumulated in FP64 on the device (fidelity |<a|b>|^2 of normalised states). */
aleph_hip_status aleph_hip_state_compare(aleph_hip_ctx* ctx, aleph_hip_state* a, aleph_hip_state* b, double* out);
aleph_hip_status aleph_hip_linear_xeb(aleph_hip_ctx* ctx, aleph_hip_state* st, double* out);
/* Marginal over `qubits` (qubits[0] = LSB of the output index), 2^n_qubits doubles. */
aleph_hip_status aleph_hip_probabilities(aleph_hip_ctx* ctx, aleph_hip_state* st, const uint32_t* qubits,
                                         uint32_t n_qubits, double* out);

/* ---- the learned planner's cost model ----
 * A perceptron (ALEPH_HIP_MODEL_INPUTS features -> 3 x 64 ReLU -> 1, FP32)
 * predicting ln(microseconds) of one fused-pass launch on a 2^24-amplitude
 * state from the pass's features (aleph_hip_plan_features; slot 0 is unused
 * and 0); the planner scales each prediction by 2^(n - 24).  It is trained on the device by the
 * library's own forward, gradient and Adam kernels, on launch times the
 * engine records (aleph_hip_plan_execute_timed).  With a model set on the
 * context, ALEPH_HIP_PLANNER_LEARNED plans every candidate (chunk qubits
 * x pass budget x 4x4 fusion), predicts each plan's time and keeps the
 * fastest; without one it plans with the rule planner.  No counterpart in
 * the reference, whose fusion widths are fixed rules
 * (crates/aleph-ir/src/passes/fuse_kq.rs:185-390, aleph-cuda/src/fusion.rs:48). */
#define ALEPH_HIP_MODEL_INPUTS 16
#define ALEPH_HIP_MODEL_PARAMS 9473
typedef struct aleph_hip_model aleph_hip_model;
/* params: ALEPH_HIP_MODEL_PARAMS floats, or NULL for a seeded He initialisation. */
aleph_hip_status aleph_hip_model_create(aleph_hip_ctx* ctx, const float* params, uint64_t seed,
                                        aleph_hip_model** out);
aleph_hip_status aleph_hip_model_get_params(aleph_hip_ctx* ctx, const aleph_hip_model* m, float* out);
aleph_hip_status aleph_hip_model_set_params(aleph_hip_ctx* ctx, aleph_hip_model* m, const float* params);
/* out[i] = model(X[i]) for `rows` rows of ALEPH_HIP_MODEL_INPUTS floats. */
aleph_hip_status aleph_hip_model_predict(aleph_hip_ctx* ctx, const aleph_hip_model* m, const float* X,
                                         uint32_t rows, float* out);
/* grads = d/dparams of mean((model(X) - y)^2) and *loss that mean (gradient check hook). */
aleph_hip_status aleph_hip_model_gradient(aleph_hip_ctx* ctx, const aleph_hip_model* m, const float* X,
                                          const float* y, uint32_t rows, float* grads, float* loss);
/* Adam on shuffled mini-batches, entirely on the device; *loss = full-data MSE after training. */
aleph_hip_status aleph_hip_model_train(aleph_hip_ctx* ctx, aleph_hip_model* m, const float* X, const float* y,
                                       uint32_t rows, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                                       float* loss);
void aleph_hip_model_free(aleph_hip_model* m);
/* The model LEARNED plans use on this context (copied; NULL clears it). */
aleph_hip_status aleph_hip_ctx_set_cost_model(aleph_hip_ctx* ctx, const aleph_hip_model* m);
/* One row of ALEPH_HIP_MODEL_INPUTS features per fused-pass launch, in launch order;
 * launch[i] receives the row's launch index (for pairing with plan_execute_timed). */
aleph_hip_status aleph_hip_plan_features(const aleph_hip_plan* plan, float* out, uint32_t* launch,
                                         uint32_t max_rows, uint32_t* rows);

/* ---- batches: many small circuits of uneven qubit count in one launch ----
 * No counterpart in the reference, which runs a parameter sweep as one
 * run_with_outcomes call per circuit (crates/aleph-backend/src/lib.rs:211-248)
 * and scores MaxCut with expectation_pauli_sum over maxcut_pauli_sum
 * (lib.rs:422-432, crates/aleph-ir/src/ansatz.rs).  Every instance is planned
 * into fused passes; the batch runs as one persistent launch (or one launch
 * per pass round), captured once into a hipGraph and replayed. */
typedef struct aleph_hip_batch aleph_hip_batch;
enum { ALEPH_HIP_BATCH_PERSISTENT = 0, ALEPH_HIP_BATCH_ROUNDS = 1 };
enum {
  ALEPH_HIP_BATCH_NONTEMPORAL = 1,
  /* Testing aid: every dependency wait of the persistent executor is made
   * unsatisfiable, so the bounded wait gives up and the run reports
   * DeviceFailure (every workgroup still exits). */
  ALEPH_HIP_BATCH_UNREACHABLE_DEPS = 2
};
typedef struct aleph_hip_batch_opts {
  uint32_t mode;             /* ALEPH_HIP_BATCH_* */
  uint32_t local_qubits;     /* chunk qubits per pass, 5..13 (0 = 12) */
  uint32_t chunks_per_item;  /* chunks one workgroup takes per claim (0 = 16) */
  uint32_t groups_per_cu;    /* persistent workgroups per CU (0 = occupancy limit) */
  uint32_t max_ops_per_pass; /* planner budget (0 = default) */
  uint32_t spin_limit;       /* bound on a dependency wait, in polls (0 = default) */
  uint32_t flags;            /* ALEPH_HIP_BATCH_NONTEMPORAL */
  uint32_t solo_qubits;      /* instances of at least this many qubits run as their own plans with
                                per-pass generated kernels after the packed ones (0 = 21, > 40 = never) */
} aleph_hip_batch_opts;
typedef struct aleph_hip_batch_stats {
  double plan_ms;
  double sweep_bytes;        /* algorithmic HBM bytes of one run (2 * 2^n * B per pass) */
  uint64_t pool_bytes;
  uint32_t n_instances, n_passes, n_items, n_rounds;
  uint32_t grid, threads, lds_bytes, mode;
  uint32_t local_qubits, chunks_per_item, groups_per_cu, solo_instances;
  uint32_t prepared_instances; /* instances whose leading H gates became the reset's |+> preparation */
  uint32_t prepared_qubits;    /* those H gates, summed over the instances */
} aleph_hip_batch_stats;
typedef struct aleph_hip_batch_tune {
  uint32_t mode, local_qubits, chunks_per_item, groups_per_cu;
  double ms;                 /* best of the timed runs */
} aleph_hip_batch_tune;
/* Generic batch: instance i is the program words[word_offsets[i] .. word_offsets[i+1]),
 * reals[real_offsets[i] .. real_offsets[i+1]) on num_qubits[i] qubits (5..32). */
aleph_hip_status aleph_hip_batch_create(aleph_hip_ctx* ctx, uint32_t n_instances, const uint32_t* num_qubits,
                                        uint32_t precision, const uint32_t* words, const uint64_t* word_offsets,
                                        const double* reals, const uint64_t* real_offsets,
                                        const aleph_hip_batch_opts* opts, aleph_hip_batch** out);
/* QAOA MaxCut batch: instance i has edges[2*edge_offsets[i] ..] (pairs u, v) and p layers of
 * gammas[i*p + l], betas[i*p + l]; program H^n, then per layer CNOT(u,v) Rz(gamma)(v) CNOT(u,v)
 * on every edge and Rx(beta) on every qubit. */
aleph_hip_status aleph_hip_batch_create_qaoa(aleph_hip_ctx* ctx, uint32_t n_instances, const uint32_t* num_qubits,
                                             const uint64_t* edge_offsets, const uint32_t* edges, uint32_t p,
                                             const double* gammas, const double* betas, uint32_t precision,
                                             const aleph_hip_batch_opts* opts, aleph_hip_batch** out);
/* New angles for a QAOA batch (same graphs and p): re-plans and re-uploads. */
aleph_hip_status aleph_hip_batch_set_qaoa_angles(aleph_hip_ctx* ctx, aleph_hip_batch* b, const double* gammas,
                                                 const double* betas);
/* Reset every instance to |0..0> and run every plan (stream-ordered graph replay). */
aleph_hip_status aleph_hip_batch_run(aleph_hip_ctx* ctx, aleph_hip_batch* b);
/* Same, synchronised, with the device time of the whole batch in *ms. */
aleph_hip_status aleph_hip_batch_run_timed(aleph_hip_ctx* ctx, aleph_hip_batch* b, double* ms);
/* <C>_i = sum_edges (1 - <Z_u Z_v>)/2 = sum_x |psi_x|^2 cut_i(x), accumulated in FP64. */
aleph_hip_status aleph_hip_batch_expect_maxcut(aleph_hip_ctx* ctx, aleph_hip_batch* b, const uint64_t* edge_offsets,
                                               const uint32_t* edges, double* out);
/* Instance i's amplitudes as interleaved complex128 (2 * 2^n doubles). */
aleph_hip_status aleph_hip_batch_read_state(aleph_hip_ctx* ctx, aleph_hip_batch* b, uint32_t instance,
                                            double* out);
aleph_hip_status aleph_hip_batch_stats_get(const aleph_hip_batch* b, aleph_hip_batch_stats* out);
/* Time candidate configurations (executor, chunk size, chunks per claim) on this batch,
 * `reps` timed runs each, keep the fastest; results[0..*n_results) lists every candidate. */
aleph_hip_status aleph_hip_batch_autotune(aleph_hip_ctx* ctx, aleph_hip_batch* b, uint32_t reps,
                                          aleph_hip_batch_tune* results, uint32_t max_results,
                                          uint32_t* n_results);
void aleph_hip_batch_free(aleph_hip_batch* b);
/* Diagnostics: plan a QAOA batch on the host only (no device; such a batch
 * cannot run) and view its packed arrays. */
aleph_hip_status aleph_hip_batch_build_host_qaoa(uint32_t n_instances, const uint32_t* num_qubits,
                                                 const uint64_t* edge_offsets, const uint32_t* edges, uint32_t p,
                                                 const double* gammas, const double* betas, uint32_t precision,
                                                 const aleph_hip_batch_opts* opts, aleph_hip_batch** out,
                                                 aleph_hip_error* err);
typedef struct aleph_hip_batch_view {
  const void* items; uint32_t n_items, item_bytes;          /* BatchItem records, queue order */
  const void* passes; uint32_t n_passes, pass_bytes;        /* PassDev records */
  const void* stages; uint32_t n_stages, stage_bytes;
  const void* ops; uint32_t n_ops, op_bytes;
  const void* terms; uint32_t n_terms, term_bytes;
  const uint64_t* offsets; uint32_t n_instances, n_rounds; /* first amplitude of each instance */
  const uint32_t* round_begin;                              /* n_rounds + 1 item indices */
  uint32_t threads, lds_bytes;
  uint64_t total_amps;
  const uint64_t* plus_masks;                               /* per instance: qubits prepared in |+> */
} aleph_hip_batch_view;
aleph_hip_status aleph_hip_batch_view_get(const aleph_hip_batch* b, aleph_hip_batch_view* out);

/* ---- distributed state on one device: DistStep::Exchange through LocalExchange
 * (crates/aleph-cuda/src/dist/exchange.rs:101-200; CPU form exchange_cpu, crates/aleph-sv/src/dist_ref.rs:78-104) ----
 * ranks[r] is rank r's slice (2^m amplitudes, all of one precision); global physical bit
 * global_bits[j] (m <= bit < m + log2(n_ranks)) trades places with local bit m-k+j: chunk c of
 * rank r moves to rank r with bits global_bits[j]-m set to c's bits, at the chunk index made of r's
 * old values of those bits.  In place, one kernel over every moved chunk pair. */
/* DeviceSv::alloc_rank (crates/aleph-cuda/src/dist/mod.rs:33-35): an m-qubit rank slice,
 * |0...0> for rank 0 and all-zero otherwise. */
aleph_hip_status aleph_hip_dist_alloc_rank(aleph_hip_ctx* ctx, uint32_t m, uint32_t precision, uint32_t rank,
                                           aleph_hip_state** out);
/* Gate::matrix (crates/aleph-core/src/gate/kinds.rs:213-399) of a validated gate on a
 * num_qubits-qubit register: out receives dim*dim interleaved complex, row-major, qubits[0] the
 * MSB (at most 2*4^6 doubles).  Host only, no context. */
aleph_hip_status aleph_hip_gate_matrix(const aleph_hip_gate* g, uint32_t num_qubits, double* out, uint32_t* dim);
aleph_hip_status aleph_hip_dist_exchange(aleph_hip_ctx* ctx, aleph_hip_state* const* ranks, uint32_t n_ranks,
                                         const uint32_t* global_bits, uint32_t k);

/* ---- noise trajectories: run_noisy / run_noisy_sparse (crates/aleph-sv/src/noise/execution.rs:261-400) ----
 * Monte-Carlo quantum-jump trajectories, one per shot, each from |0...0>.
 * A channel fires after the program record `after` (a gate), channels of one
 * record in array order (the all-qubit attachments first, then the
 * qubit-specific ones: crates/aleph-sv/src/noise/model.rs:60-80).
 * Pauli channels (PauliChannel, noise/error.rs:11-20) pick one Pauli string
 * from `weights`, `paulis` holding n_qubits codes per branch (0 I, 1 X, 2 Y,
 * 3 Z; code j acts on qubits[j]); Kraus channels (KrausChannel, one qubit)
 * hold n_branches 2x2 complex matrices, interleaved row-major.  Readout
 * errors (ReadoutError) give m[t][o] = P(read o | true t) as 4 doubles per
 * qubit for the qubits set in `readout_mask`.
 * Shot i of the call is absolute shot `shot_offset + i`: its outcome depends
 * only on (seed, shot_offset + i), so a run split over several calls equals
 * one call.  Outcomes are basis indices (bit q = qubit q).  States of up to
 * 13 qubits run in LDS; 14..24 qubits run in per-workgroup device slices of
 * at most `budget_bytes` (0: 4 GiB).  Errors: ALEPH_HIP_ERR_INVALID_ARGUMENT
 * with kind "invalid-channel" (malformed channel or readout error, the
 * reference's NoiseError::InvalidChannel) or "unsupported" (a dense gate on
 * more than 3 qubits), ALEPH_HIP_ERR_TOO_MANY_QUBITS beyond 24 qubits. */
enum { ALEPH_HIP_CHANNEL_PAULI = 0, ALEPH_HIP_CHANNEL_KRAUS = 1 };
typedef struct aleph_hip_channel {
  uint32_t after;
  uint32_t kind;
  uint32_t n_qubits;
  uint32_t n_branches;
  uint32_t qubits[2];
  const double* weights;
  const uint8_t* paulis;
  const double* kraus;
} aleph_hip_channel;
aleph_hip_status aleph_hip_run_noisy(aleph_hip_ctx* ctx, uint32_t num_qubits, const uint32_t* words, size_t n_words,
                                     const double* reals, size_t n_reals, const aleph_hip_channel* channels,
                                     uint32_t n_channels, const double* readout, uint64_t readout_mask,
                                     uint32_t shots, uint64_t seed, uint64_t shot_offset, size_t budget_bytes,
                                     uint64_t* outcomes);

#ifdef __cplusplus
}
#endif
#endif /* ALEPH_HIP_H */
