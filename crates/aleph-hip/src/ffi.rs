//! `extern "C"` declarations of `include/aleph_hip.h`, one for one.
//!
//! Struct layouts are `#[repr(C)]` copies of the header's; the library's
//! Python binding (`aleph_amd/_lib.py`) checks the same sizes
//! (`aleph_hip_error` 336 B, `aleph_hip_gate` 64 B, `aleph_hip_plan_opts` 32 B,
//! `aleph_hip_batch_opts` 32 B), and `tests::layouts` below pins them here.
#![allow(non_camel_case_types, dead_code)]

use std::os::raw::{c_char, c_void};

pub type aleph_hip_status = i32;

// Status codes: BackendError variants (crates/aleph-backend/src/lib.rs:22-81).
pub const ALEPH_HIP_OK: i32 = 0;
pub const ALEPH_HIP_ERR_DEVICE_FAILURE: i32 = 1;
pub const ALEPH_HIP_ERR_QUBIT_OUT_OF_RANGE: i32 = 2;
pub const ALEPH_HIP_ERR_DUPLICATE_QUBIT: i32 = 3;
pub const ALEPH_HIP_ERR_ARITY_MISMATCH: i32 = 4;
pub const ALEPH_HIP_ERR_UNSUPPORTED_GATE: i32 = 5;
pub const ALEPH_HIP_ERR_UNSUPPORTED_INSTRUCTION: i32 = 6;
pub const ALEPH_HIP_ERR_SYMBOLIC_PARAM: i32 = 7;
pub const ALEPH_HIP_ERR_NON_FINITE_PARAM: i32 = 8;
pub const ALEPH_HIP_ERR_NON_UNITARY_MATRIX: i32 = 9;
pub const ALEPH_HIP_ERR_EMPTY_CIRCUIT: i32 = 10;
pub const ALEPH_HIP_ERR_DEGENERATE_MEASUREMENT: i32 = 11;
pub const ALEPH_HIP_ERR_TOO_MANY_QUBITS: i32 = 12;
pub const ALEPH_HIP_ERR_INVALID_PAULI_STRING: i32 = 13;
pub const ALEPH_HIP_ERR_INVALID_STATE: i32 = 14;
pub const ALEPH_HIP_ERR_OPTIMIZATION: i32 = 15;
pub const ALEPH_HIP_ERR_NO_DEVICE: i32 = 16;
pub const ALEPH_HIP_ERR_INVALID_ARGUMENT: i32 = 17;
pub const ALEPH_HIP_ERR_OUT_OF_MEMORY: i32 = 18;

// Gate kinds: the order of the reference `Gate` enum (crates/aleph-core/src/gate/kinds.rs:13-115).
pub const ALEPH_HIP_GATE_H: u32 = 0;
pub const ALEPH_HIP_GATE_X: u32 = 1;
pub const ALEPH_HIP_GATE_Y: u32 = 2;
pub const ALEPH_HIP_GATE_Z: u32 = 3;
pub const ALEPH_HIP_GATE_S: u32 = 4;
pub const ALEPH_HIP_GATE_SDG: u32 = 5;
pub const ALEPH_HIP_GATE_T: u32 = 6;
pub const ALEPH_HIP_GATE_TDG: u32 = 7;
pub const ALEPH_HIP_GATE_RX: u32 = 8;
pub const ALEPH_HIP_GATE_RY: u32 = 9;
pub const ALEPH_HIP_GATE_RZ: u32 = 10;
pub const ALEPH_HIP_GATE_PHASE: u32 = 11;
pub const ALEPH_HIP_GATE_U3: u32 = 12;
pub const ALEPH_HIP_GATE_CNOT: u32 = 13;
pub const ALEPH_HIP_GATE_CZ: u32 = 14;
pub const ALEPH_HIP_GATE_SWAP: u32 = 15;
pub const ALEPH_HIP_GATE_ISWAP: u32 = 16;
pub const ALEPH_HIP_GATE_ISWAPDG: u32 = 17;
pub const ALEPH_HIP_GATE_CRX: u32 = 18;
pub const ALEPH_HIP_GATE_CRY: u32 = 19;
pub const ALEPH_HIP_GATE_CRZ: u32 = 20;
pub const ALEPH_HIP_GATE_TOFFOLI: u32 = 21;
pub const ALEPH_HIP_GATE_CCZ: u32 = 22;
pub const ALEPH_HIP_GATE_UNITARY1Q: u32 = 23;
pub const ALEPH_HIP_GATE_UNITARY2Q: u32 = 24;
pub const ALEPH_HIP_GATE_UNITARYKQ: u32 = 25;
pub const ALEPH_HIP_GATE_UNITARY1QDIAG: u32 = 26;
pub const ALEPH_HIP_GATE_UNITARY2QDIAG: u32 = 27;

// Program-only records (Instruction enum, crates/aleph-ir/src/instruction.rs:13-36).
pub const ALEPH_HIP_INSTR_DIAGONAL_PHASE: u32 = 64;
pub const ALEPH_HIP_INSTR_BARRIER: u32 = 65;

pub const ALEPH_HIP_F64: u32 = 64;
pub const ALEPH_HIP_F32: u32 = 32;

pub const ALEPH_HIP_PLANNER_UNFUSED: u32 = 0;
pub const ALEPH_HIP_PLANNER_RULE: u32 = 1;
pub const ALEPH_HIP_PLANNER_LEARNED: u32 = 2;
pub const ALEPH_HIP_PLAN_NONTEMPORAL: u32 = 1;
pub const ALEPH_HIP_PLAN_NO_FUSE_2Q: u32 = 2;
pub const ALEPH_HIP_PLAN_NO_FUSE_DIAG: u32 = 4;
pub const ALEPH_HIP_PLAN_JIT: u32 = 8;
pub const ALEPH_HIP_PLAN_RENORMALIZE: u32 = 16;
pub const ALEPH_HIP_PLAN_GENERIC_PHASE: u32 = 32;

pub const ALEPH_HIP_BATCH_PERSISTENT: u32 = 0;
pub const ALEPH_HIP_BATCH_ROUNDS: u32 = 1;
pub const ALEPH_HIP_BATCH_NONTEMPORAL: u32 = 1;

pub const ALEPH_HIP_CHANNEL_PAULI: u32 = 0;
pub const ALEPH_HIP_CHANNEL_KRAUS: u32 = 1;

pub const ALEPH_HIP_MODEL_INPUTS: usize = 16;
pub const ALEPH_HIP_MODEL_PARAMS: usize = 9473;

#[repr(C)] pub struct aleph_hip_ctx { _p: [u8; 0] }
#[repr(C)] pub struct aleph_hip_state { _p: [u8; 0] }
#[repr(C)] pub struct aleph_hip_plan { _p: [u8; 0] }
#[repr(C)] pub struct aleph_hip_model { _p: [u8; 0] }
#[repr(C)] pub struct aleph_hip_batch { _p: [u8; 0] }

#[repr(C)]
#[derive(Clone, Copy)]
pub struct aleph_hip_error {
    pub code: i32,
    pub qubit: u32,
    pub num_qubits: u32,
    pub expected: u32,
    pub got: u32,
    pub value: f64,
    pub kind: [c_char; 48],
    pub message: [c_char; 256],
}

/// `GateInstance` (crates/aleph-core/src/gate/instance.rs:40-59) as plain data.
#[repr(C)]
pub struct aleph_hip_gate {
    pub kind: u32,
    pub n_qubits: u32,
    pub n_controls: u32,
    pub k: u32,
    pub qubits: *const u32,
    pub controls: *const u32,
    pub params: [f64; 3],
    /// Interleaved (re, im) payload of Unitary* kinds: row-major, qubits[0] = MSB.
    pub matrix: *const f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct aleph_hip_plan_opts {
    pub planner: u32,
    pub local_qubits: u32,
    pub max_ops_per_pass: u32,
    pub coalesce_bits: u32,
    pub fuse_1q: u32,
    pub reg_bits: u32,
    pub flags: u32,
    pub kq_width: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct aleph_hip_plan_stats {
    pub plan_ms: f64,
    pub input_gates: u32,
    pub n_passes: u32,
    pub n_stages: u32,
    pub n_ops: u32,
    pub n_launches: u32,
    pub planner_used: u32,
    pub sweep_bytes: f64,
    pub chosen_local_qubits: u32,
    pub chosen_max_ops: u32,
    pub chosen_flags: u32,
    pub n_candidates: u32,
    pub predicted_us: f64,
    pub chosen_kq_width: u32,
    pub chosen_rule: u32,
    pub rule_predicted_us: f64,
    pub cancelled_gates: u32,
    pub reserved2: u32,
}

#[repr(C)]
pub struct aleph_hip_plan_view {
    pub steps: *const c_void, pub n_steps: u32, pub step_bytes: u32,
    pub passes: *const c_void, pub n_passes: u32, pub pass_bytes: u32,
    pub stages: *const c_void, pub n_stages: u32, pub stage_bytes: u32,
    pub ops: *const c_void, pub n_ops: u32, pub op_bytes: u32,
    pub terms: *const c_void, pub n_terms: u32, pub term_bytes: u32,
    pub payload: *const c_void, pub payload_bytes: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct aleph_hip_batch_opts {
    pub mode: u32,
    pub local_qubits: u32,
    pub chunks_per_item: u32,
    pub groups_per_cu: u32,
    pub max_ops_per_pass: u32,
    pub spin_limit: u32,
    pub flags: u32,
    pub solo_qubits: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct aleph_hip_batch_stats {
    pub plan_ms: f64,
    pub sweep_bytes: f64,
    pub pool_bytes: u64,
    pub n_instances: u32, pub n_passes: u32, pub n_items: u32, pub n_rounds: u32,
    pub grid: u32, pub threads: u32, pub lds_bytes: u32, pub mode: u32,
    pub local_qubits: u32, pub chunks_per_item: u32, pub groups_per_cu: u32, pub solo_instances: u32,
    pub prepared_instances: u32, pub prepared_qubits: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct aleph_hip_batch_tune {
    pub mode: u32,
    pub local_qubits: u32,
    pub chunks_per_item: u32,
    pub groups_per_cu: u32,
    pub ms: f64,
}

#[repr(C)]
pub struct aleph_hip_batch_view {
    pub items: *const c_void, pub n_items: u32, pub item_bytes: u32,
    pub passes: *const c_void, pub n_passes: u32, pub pass_bytes: u32,
    pub stages: *const c_void, pub n_stages: u32, pub stage_bytes: u32,
    pub ops: *const c_void, pub n_ops: u32, pub op_bytes: u32,
    pub terms: *const c_void, pub n_terms: u32, pub term_bytes: u32,
    pub offsets: *const u64, pub n_instances: u32, pub n_rounds: u32,
    pub round_begin: *const u32,
    pub threads: u32, pub lds_bytes: u32,
    pub total_amps: u64,
    pub plus_masks: *const u64,
}

/// One noise channel attached after a program record (noise trajectories).
#[repr(C)]
pub struct aleph_hip_channel {
    pub after: u32,
    pub kind: u32,
    pub n_qubits: u32,
    pub n_branches: u32,
    pub qubits: [u32; 2],
    pub weights: *const f64,
    pub paulis: *const u8,
    pub kraus: *const f64,
}

extern "C" {
    pub fn aleph_hip_device_count(count: *mut i32) -> aleph_hip_status;
    pub fn aleph_hip_ctx_create(ordinal: i32, seed: u64, out: *mut *mut aleph_hip_ctx) -> aleph_hip_status;
    pub fn aleph_hip_ctx_destroy(ctx: *mut aleph_hip_ctx);
    pub fn aleph_hip_ctx_set_qubit_cap(ctx: *mut aleph_hip_ctx, cap: u32) -> aleph_hip_status;
    pub fn aleph_hip_ctx_reseed(ctx: *mut aleph_hip_ctx, seed: u64) -> aleph_hip_status;
    pub fn aleph_hip_synchronize(ctx: *mut aleph_hip_ctx) -> aleph_hip_status;
    pub fn aleph_hip_last_error(ctx: *const aleph_hip_ctx, out: *mut aleph_hip_error) -> aleph_hip_status;
    pub fn aleph_hip_ctx_stream(ctx: *mut aleph_hip_ctx, stream: *mut *mut c_void) -> aleph_hip_status;
    pub fn aleph_hip_state_alloc(ctx: *mut aleph_hip_ctx, num_qubits: u32, precision: u32, out: *mut *mut aleph_hip_state) -> aleph_hip_status;
    pub fn aleph_hip_state_free(st: *mut aleph_hip_state);
    pub fn aleph_hip_state_reset(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state) -> aleph_hip_status;
    pub fn aleph_hip_state_load(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, amps: *const f64) -> aleph_hip_status;
    pub fn aleph_hip_state_read(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, out: *mut f64) -> aleph_hip_status;
    pub fn aleph_hip_state_load_device(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, dptr: *const c_void) -> aleph_hip_status;
    pub fn aleph_hip_state_read_device(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, dptr: *mut c_void) -> aleph_hip_status;
    pub fn aleph_hip_state_info(st: *const aleph_hip_state, num_qubits: *mut u32, precision: *mut u32, device_ptr: *mut *mut c_void) -> aleph_hip_status;
    pub fn aleph_hip_apply_gate(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, g: *const aleph_hip_gate) -> aleph_hip_status;
    pub fn aleph_hip_apply_diagonal_phase(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, n_terms: u32, angles: *const f64, conds: *const u64, offsets: *const u32) -> aleph_hip_status;
    pub fn aleph_hip_unpermute(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, perm: *const u32) -> aleph_hip_status;
    pub fn aleph_hip_apply_1q(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, target: u32, ctrl_mask: u64, m: *const f64) -> aleph_hip_status;
    pub fn aleph_hip_apply_kq(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, k: u32, qubits: *const u32, ctrl_mask: u64, m: *const f64) -> aleph_hip_status;
    pub fn aleph_hip_apply_diag(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, k: u32, qubits: *const u32, ctrl_mask: u64, d: *const f64) -> aleph_hip_status;
    pub fn aleph_hip_apply_cnot(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, control: u32, target: u32) -> aleph_hip_status;
    pub fn aleph_hip_apply_swap(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, a: u32, b: u32) -> aleph_hip_status;
    pub fn aleph_hip_plan_create(ctx: *mut aleph_hip_ctx, num_qubits: u32, precision: u32, words: *const u32, n_words: usize, reals: *const f64, n_reals: usize, opts: *const aleph_hip_plan_opts, out: *mut *mut aleph_hip_plan) -> aleph_hip_status;
    pub fn aleph_hip_plan_execute(ctx: *mut aleph_hip_ctx, plan: *mut aleph_hip_plan, st: *mut aleph_hip_state) -> aleph_hip_status;
    pub fn aleph_hip_plan_execute_timed(ctx: *mut aleph_hip_ctx, plan: *mut aleph_hip_plan, st: *mut aleph_hip_state, ms: *mut f64, kinds: *mut u32) -> aleph_hip_status;
    pub fn aleph_hip_plan_stats_get(plan: *const aleph_hip_plan, out: *mut aleph_hip_plan_stats) -> aleph_hip_status;
    pub fn aleph_hip_plan_free(plan: *mut aleph_hip_plan);
    pub fn aleph_hip_plan_build_host(num_qubits: u32, precision: u32, words: *const u32, n_words: usize, reals: *const f64, n_reals: usize, opts: *const aleph_hip_plan_opts, out: *mut *mut aleph_hip_plan, err: *mut aleph_hip_error) -> aleph_hip_status;
    pub fn aleph_hip_plan_view_get(plan: *const aleph_hip_plan, out: *mut aleph_hip_plan_view) -> aleph_hip_status;
    pub fn aleph_hip_plan_jit_source(plan: *const aleph_hip_plan, pass: u32, out: *mut c_char, cap: usize, len: *mut usize) -> aleph_hip_status;
    pub fn aleph_hip_run_program(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, words: *const u32, n_words: usize, reals: *const f64, n_reals: usize, opts: *const aleph_hip_plan_opts, stats: *mut aleph_hip_plan_stats) -> aleph_hip_status;
    pub fn aleph_hip_norm_sqr(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, out: *mut f64) -> aleph_hip_status;
    pub fn aleph_hip_validate(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state) -> aleph_hip_status;
    pub fn aleph_hip_measure(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, qubit: u32, outcome: *mut i32) -> aleph_hip_status;
    pub fn aleph_hip_measure_with_uniform(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, qubit: u32, u: f64, outcome: *mut i32, consumed: *mut i32) -> aleph_hip_status;
    pub fn aleph_hip_sample(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, shots: u32, out: *mut u64) -> aleph_hip_status;
    pub fn aleph_hip_sample_with_uniforms(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, uniforms: *const f64, shots: u32, out: *mut u64) -> aleph_hip_status;
    pub fn aleph_hip_sample_with_uniforms_noalloc(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, uniforms: *const f64, shots: u32, out: *mut u64) -> aleph_hip_status;
    pub fn aleph_hip_expectation(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, coefficient: f64, qubits: *const u32, paulis: *const u8, n_terms: u32, out: *mut f64) -> aleph_hip_status;
    pub fn aleph_hip_expectation_sum(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, n_terms: u32, coeffs: *const f64, offsets: *const u32, qubits: *const u32, paulis: *const u8, out: *mut f64) -> aleph_hip_status;
    pub fn aleph_hip_linear_xeb(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, out: *mut f64) -> aleph_hip_status;
    pub fn aleph_hip_probabilities(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, qubits: *const u32, n_qubits: u32, out: *mut f64) -> aleph_hip_status;
    pub fn aleph_hip_model_create(ctx: *mut aleph_hip_ctx, params: *const f32, seed: u64, out: *mut *mut aleph_hip_model) -> aleph_hip_status;
    pub fn aleph_hip_model_get_params(ctx: *mut aleph_hip_ctx, m: *const aleph_hip_model, out: *mut f32) -> aleph_hip_status;
    pub fn aleph_hip_model_set_params(ctx: *mut aleph_hip_ctx, m: *mut aleph_hip_model, params: *const f32) -> aleph_hip_status;
    pub fn aleph_hip_model_predict(ctx: *mut aleph_hip_ctx, m: *const aleph_hip_model, X: *const f32, rows: u32, out: *mut f32) -> aleph_hip_status;
    pub fn aleph_hip_model_gradient(ctx: *mut aleph_hip_ctx, m: *const aleph_hip_model, X: *const f32, y: *const f32, rows: u32, grads: *mut f32, loss: *mut f32) -> aleph_hip_status;
    pub fn aleph_hip_model_train(ctx: *mut aleph_hip_ctx, m: *mut aleph_hip_model, X: *const f32, y: *const f32, rows: u32, epochs: u32, batch: u32, lr: f32, seed: u64, loss: *mut f32) -> aleph_hip_status;
    pub fn aleph_hip_model_free(m: *mut aleph_hip_model);
    pub fn aleph_hip_ctx_set_cost_model(ctx: *mut aleph_hip_ctx, m: *const aleph_hip_model) -> aleph_hip_status;
    pub fn aleph_hip_plan_features(plan: *const aleph_hip_plan, out: *mut f32, launch: *mut u32, max_rows: u32, rows: *mut u32) -> aleph_hip_status;
    pub fn aleph_hip_batch_create(ctx: *mut aleph_hip_ctx, n_instances: u32, num_qubits: *const u32, precision: u32, words: *const u32, word_offsets: *const u64, reals: *const f64, real_offsets: *const u64, opts: *const aleph_hip_batch_opts, out: *mut *mut aleph_hip_batch) -> aleph_hip_status;
    pub fn aleph_hip_batch_create_qaoa(ctx: *mut aleph_hip_ctx, n_instances: u32, num_qubits: *const u32, edge_offsets: *const u64, edges: *const u32, p: u32, gammas: *const f64, betas: *const f64, precision: u32, opts: *const aleph_hip_batch_opts, out: *mut *mut aleph_hip_batch) -> aleph_hip_status;
    pub fn aleph_hip_batch_set_qaoa_angles(ctx: *mut aleph_hip_ctx, b: *mut aleph_hip_batch, gammas: *const f64, betas: *const f64) -> aleph_hip_status;
    pub fn aleph_hip_batch_run(ctx: *mut aleph_hip_ctx, b: *mut aleph_hip_batch) -> aleph_hip_status;
    pub fn aleph_hip_batch_run_timed(ctx: *mut aleph_hip_ctx, b: *mut aleph_hip_batch, ms: *mut f64) -> aleph_hip_status;
    pub fn aleph_hip_batch_expect_maxcut(ctx: *mut aleph_hip_ctx, b: *mut aleph_hip_batch, edge_offsets: *const u64, edges: *const u32, out: *mut f64) -> aleph_hip_status;
    pub fn aleph_hip_batch_read_state(ctx: *mut aleph_hip_ctx, b: *mut aleph_hip_batch, instance: u32, out: *mut f64) -> aleph_hip_status;
    pub fn aleph_hip_batch_stats_get(b: *const aleph_hip_batch, out: *mut aleph_hip_batch_stats) -> aleph_hip_status;
    pub fn aleph_hip_batch_autotune(ctx: *mut aleph_hip_ctx, b: *mut aleph_hip_batch, reps: u32, results: *mut aleph_hip_batch_tune, max_results: u32, n_results: *mut u32) -> aleph_hip_status;
    pub fn aleph_hip_batch_free(b: *mut aleph_hip_batch);
    pub fn aleph_hip_batch_build_host_qaoa(n_instances: u32, num_qubits: *const u32, edge_offsets: *const u64, edges: *const u32, p: u32, gammas: *const f64, betas: *const f64, precision: u32, opts: *const aleph_hip_batch_opts, out: *mut *mut aleph_hip_batch, err: *mut aleph_hip_error) -> aleph_hip_status;
    pub fn aleph_hip_batch_view_get(b: *const aleph_hip_batch, out: *mut aleph_hip_batch_view) -> aleph_hip_status;

    pub fn aleph_hip_state_normalize(ctx: *mut aleph_hip_ctx, st: *mut aleph_hip_state, norm_before: *mut f64)
        -> aleph_hip_status;
    pub fn aleph_hip_state_compare(
        ctx: *mut aleph_hip_ctx, a: *mut aleph_hip_state, b: *mut aleph_hip_state, out: *mut f64
    ) -> aleph_hip_status;
    pub fn aleph_hip_dist_alloc_rank(
        ctx: *mut aleph_hip_ctx, m: u32, precision: u32, rank: u32, out: *mut *mut aleph_hip_state
    ) -> aleph_hip_status;
    pub fn aleph_hip_gate_matrix(g: *const aleph_hip_gate, num_qubits: u32, out: *mut f64, dim: *mut u32)
        -> aleph_hip_status;
    pub fn aleph_hip_dist_exchange(
        ctx: *mut aleph_hip_ctx, ranks: *const *mut aleph_hip_state, n_ranks: u32, global_bits: *const u32, k: u32
    ) -> aleph_hip_status;
    pub fn aleph_hip_run_noisy(
        ctx: *mut aleph_hip_ctx, num_qubits: u32, words: *const u32, n_words: usize, reals: *const f64,
        n_reals: usize, channels: *const aleph_hip_channel, n_channels: u32, readout: *const f64,
        readout_mask: u64, shots: u32, seed: u64, shot_offset: u64, budget_bytes: usize, outcomes: *mut u64
    ) -> aleph_hip_status;
}

#[cfg(test)]
mod tests {
    use super::*;

    #[test]
    fn layouts() {
        assert_eq!(std::mem::size_of::<aleph_hip_error>(), 336);
        assert_eq!(std::mem::size_of::<aleph_hip_gate>(), 64);
        assert_eq!(std::mem::size_of::<aleph_hip_plan_opts>(), 32);
        assert_eq!(std::mem::size_of::<aleph_hip_batch_opts>(), 32);
        assert_eq!(std::mem::size_of::<aleph_hip_batch_stats>(), 80);
        assert_eq!(std::mem::size_of::<aleph_hip_plan_stats>(), 88);
        assert_eq!(std::mem::size_of::<aleph_hip_channel>(), 48);
    }
}
