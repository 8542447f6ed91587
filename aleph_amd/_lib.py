"""ctypes binding of `libaleph_hip.so` (declarations in include/aleph_hip.h).

Loading fails loudly when the library has not been built: there is no
CPU fallback behind this module.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libaleph_hip.so"

u32, u64, i32, f64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_double
P = C.c_void_p
PU32, PU64, PI32, PF64, PU8 = (C.POINTER(u32), C.POINTER(u64), C.POINTER(i32), C.POINTER(f64),
                               C.POINTER(C.c_uint8))

# status codes (BackendError variants)
OK = 0
ERR_DEVICE_FAILURE = 1
ERR_QUBIT_OUT_OF_RANGE = 2
ERR_DUPLICATE_QUBIT = 3
ERR_ARITY_MISMATCH = 4
ERR_UNSUPPORTED_GATE = 5
ERR_UNSUPPORTED_INSTRUCTION = 6
ERR_SYMBOLIC_PARAM = 7
ERR_NON_FINITE_PARAM = 8
ERR_NON_UNITARY_MATRIX = 9
ERR_EMPTY_CIRCUIT = 10
ERR_DEGENERATE_MEASUREMENT = 11
ERR_TOO_MANY_QUBITS = 12
ERR_INVALID_PAULI_STRING = 13
ERR_INVALID_STATE = 14
ERR_OPTIMIZATION = 15
ERR_NO_DEVICE = 16
ERR_INVALID_ARGUMENT = 17
ERR_OUT_OF_MEMORY = 18

F64, F32 = 64, 32
PLANNER_UNFUSED, PLANNER_RULE, PLANNER_LEARNED = 0, 1, 2
PLAN_NONTEMPORAL = 1
PLAN_NO_FUSE_2Q = 2
PLAN_NO_FUSE_DIAG = 4
PLAN_JIT = 8
PLAN_RENORMALIZE = 16
PLAN_GENERIC_PHASE = 32
INSTR_DIAGONAL_PHASE, INSTR_BARRIER = 64, 65


class Error(C.Structure):
    _fields_ = [("code", i32), ("qubit", u32), ("num_qubits", u32), ("expected", u32), ("got", u32),
                ("value", f64), ("kind", C.c_char * 48), ("message", C.c_char * 256)]


class Gate(C.Structure):
    _fields_ = [("kind", u32), ("n_qubits", u32), ("n_controls", u32), ("k", u32), ("qubits", PU32),
                ("controls", PU32), ("params", f64 * 3), ("matrix", PF64)]


class PlanOpts(C.Structure):
    _fields_ = [("planner", u32), ("local_qubits", u32), ("max_ops_per_pass", u32), ("coalesce_bits", u32),
                ("fuse_1q", u32), ("reg_bits", u32), ("flags", u32), ("kq_width", u32)]


class PlanStats(C.Structure):
    _fields_ = [("plan_ms", f64), ("input_gates", u32), ("n_passes", u32), ("n_stages", u32), ("n_ops", u32),
                ("n_launches", u32), ("planner_used", u32), ("sweep_bytes", f64), ("chosen_local_qubits", u32),
                ("chosen_max_ops", u32), ("chosen_flags", u32), ("n_candidates", u32), ("predicted_us", f64),
                ("chosen_kq_width", u32), ("chosen_rule", u32), ("rule_predicted_us", f64),
                ("cancelled_gates", u32), ("reserved", u32)]


MODEL_INPUTS = 16
MODEL_PARAMS = 9473
PF32 = C.POINTER(C.c_float)


BATCH_PERSISTENT, BATCH_ROUNDS = 0, 1
BATCH_NONTEMPORAL = 1
BATCH_UNREACHABLE_DEPS = 2


class BatchOpts(C.Structure):
    _fields_ = [("mode", u32), ("local_qubits", u32), ("chunks_per_item", u32), ("groups_per_cu", u32),
                ("max_ops_per_pass", u32), ("spin_limit", u32), ("flags", u32), ("solo_qubits", u32)]


class BatchStats(C.Structure):
    _fields_ = [("plan_ms", f64), ("sweep_bytes", f64), ("pool_bytes", u64), ("n_instances", u32),
                ("n_passes", u32), ("n_items", u32), ("n_rounds", u32), ("grid", u32), ("threads", u32),
                ("lds_bytes", u32), ("mode", u32), ("local_qubits", u32), ("chunks_per_item", u32),
                ("groups_per_cu", u32), ("solo_instances", u32), ("prepared_instances", u32),
                ("prepared_qubits", u32)]


class BatchTune(C.Structure):
    _fields_ = [("mode", u32), ("local_qubits", u32), ("chunks_per_item", u32), ("groups_per_cu", u32),
                ("ms", f64)]


class BatchView(C.Structure):
    _fields_ = [("items", P), ("n_items", u32), ("item_bytes", u32),
                ("passes", P), ("n_passes", u32), ("pass_bytes", u32),
                ("stages", P), ("n_stages", u32), ("stage_bytes", u32),
                ("ops", P), ("n_ops", u32), ("op_bytes", u32),
                ("terms", P), ("n_terms", u32), ("term_bytes", u32),
                ("offsets", PU64), ("n_instances", u32), ("n_rounds", u32),
                ("round_begin", PU32), ("threads", u32), ("lds_bytes", u32), ("total_amps", u64),
                ("plus_masks", PU64)]


CHANNEL_PAULI, CHANNEL_KRAUS = 0, 1


class Channel(C.Structure):
    _fields_ = [("after", u32), ("kind", u32), ("n_qubits", u32), ("n_branches", u32), ("qubits", u32 * 2),
                ("weights", PF64), ("paulis", PU8), ("kraus", PF64)]


# name -> (restype, argtypes); every symbol include/aleph_hip.h declares.
SIGNATURES = {
    "aleph_hip_device_count": (i32, [PI32]),
    "aleph_hip_ctx_create": (i32, [i32, u64, C.POINTER(P)]),
    "aleph_hip_ctx_destroy": (None, [P]),
    "aleph_hip_ctx_set_qubit_cap": (i32, [P, u32]),
    "aleph_hip_ctx_reseed": (i32, [P, u64]),
    "aleph_hip_synchronize": (i32, [P]),
    "aleph_hip_last_error": (i32, [P, C.POINTER(Error)]),
    "aleph_hip_ctx_stream": (i32, [P, C.POINTER(P)]),
    "aleph_hip_state_alloc": (i32, [P, u32, u32, C.POINTER(P)]),
    "aleph_hip_state_free": (None, [P]),
    "aleph_hip_state_reset": (i32, [P, P]),
    "aleph_hip_state_load": (i32, [P, P, PF64]),
    "aleph_hip_state_read": (i32, [P, P, PF64]),
    "aleph_hip_state_load_device": (i32, [P, P, P]),
    "aleph_hip_state_read_device": (i32, [P, P, P]),
    "aleph_hip_state_info": (i32, [P, PU32, PU32, C.POINTER(P)]),
    "aleph_hip_apply_gate": (i32, [P, P, C.POINTER(Gate)]),
    "aleph_hip_apply_diagonal_phase": (i32, [P, P, u32, PF64, PU64, PU32]),
    "aleph_hip_unpermute": (i32, [P, P, PU32]),
    "aleph_hip_apply_1q": (i32, [P, P, u32, u64, PF64]),
    "aleph_hip_apply_kq": (i32, [P, P, u32, PU32, u64, PF64]),
    "aleph_hip_apply_diag": (i32, [P, P, u32, PU32, u64, PF64]),
    "aleph_hip_apply_cnot": (i32, [P, P, u32, u32]),
    "aleph_hip_apply_swap": (i32, [P, P, u32, u32]),
    "aleph_hip_plan_create": (i32, [P, u32, u32, PU32, C.c_size_t, PF64, C.c_size_t, C.POINTER(PlanOpts),
                                    C.POINTER(P)]),
    "aleph_hip_plan_execute": (i32, [P, P, P]),
    "aleph_hip_plan_execute_timed": (i32, [P, P, P, PF64, PU32]),
    "aleph_hip_plan_stats_get": (i32, [P, C.POINTER(PlanStats)]),
    "aleph_hip_plan_free": (None, [P]),
    "aleph_hip_plan_build_host": (i32, [u32, u32, PU32, C.c_size_t, PF64, C.c_size_t, C.POINTER(PlanOpts),
                                        C.POINTER(P), C.POINTER(Error)]),
    "aleph_hip_plan_view_get": (i32, [P, P]),
    "aleph_hip_run_program": (i32, [P, P, PU32, C.c_size_t, PF64, C.c_size_t, C.POINTER(PlanOpts),
                                    C.POINTER(PlanStats)]),
    "aleph_hip_norm_sqr": (i32, [P, P, PF64]),
    "aleph_hip_validate": (i32, [P, P]),
    "aleph_hip_measure": (i32, [P, P, u32, PI32]),
    "aleph_hip_measure_with_uniform": (i32, [P, P, u32, f64, PI32, PI32]),
    "aleph_hip_sample": (i32, [P, P, u32, PU64]),
    "aleph_hip_sample_with_uniforms": (i32, [P, P, PF64, u32, PU64]),
    "aleph_hip_sample_with_uniforms_noalloc": (i32, [P, P, PF64, u32, PU64]),
    "aleph_hip_expectation": (i32, [P, P, f64, PU32, PU8, u32, PF64]),
    "aleph_hip_expectation_sum": (i32, [P, P, u32, PF64, PU32, PU32, PU8, PF64]),
    "aleph_hip_linear_xeb": (i32, [P, P, PF64]),
    "aleph_hip_state_normalize": (i32, [P, P, PF64]),
    "aleph_hip_state_compare": (i32, [P, P, P, PF64]),
    "aleph_hip_probabilities": (i32, [P, P, PU32, u32, PF64]),
    "aleph_hip_batch_create": (i32, [P, u32, PU32, u32, PU32, PU64, PF64, PU64, C.POINTER(BatchOpts),
                                     C.POINTER(P)]),
    "aleph_hip_batch_create_qaoa": (i32, [P, u32, PU32, PU64, PU32, u32, PF64, PF64, u32, C.POINTER(BatchOpts),
                                          C.POINTER(P)]),
    "aleph_hip_batch_set_qaoa_angles": (i32, [P, P, PF64, PF64]),
    "aleph_hip_batch_run": (i32, [P, P]),
    "aleph_hip_batch_run_timed": (i32, [P, P, PF64]),
    "aleph_hip_batch_expect_maxcut": (i32, [P, P, PU64, PU32, PF64]),
    "aleph_hip_batch_read_state": (i32, [P, P, u32, PF64]),
    "aleph_hip_batch_stats_get": (i32, [P, C.POINTER(BatchStats)]),
    "aleph_hip_batch_autotune": (i32, [P, P, u32, C.POINTER(BatchTune), u32, PU32]),
    "aleph_hip_batch_free": (None, [P]),
    "aleph_hip_model_create": (i32, [P, PF32, u64, C.POINTER(P)]),
    "aleph_hip_model_get_params": (i32, [P, P, PF32]),
    "aleph_hip_model_set_params": (i32, [P, P, PF32]),
    "aleph_hip_model_predict": (i32, [P, P, PF32, u32, PF32]),
    "aleph_hip_model_gradient": (i32, [P, P, PF32, PF32, u32, PF32, PF32]),
    "aleph_hip_model_train": (i32, [P, P, PF32, PF32, u32, u32, u32, C.c_float, u64, PF32]),
    "aleph_hip_model_free": (None, [P]),
    "aleph_hip_ctx_set_cost_model": (i32, [P, P]),
    "aleph_hip_plan_features": (i32, [P, PF32, PU32, u32, PU32]),
    "aleph_hip_plan_jit_source": (i32, [P, u32, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "aleph_hip_batch_build_host_qaoa": (i32, [u32, PU32, PU64, PU32, u32, PF64, PF64, u32, C.POINTER(BatchOpts),
                                              C.POINTER(P), C.POINTER(Error)]),
    "aleph_hip_batch_view_get": (i32, [P, C.POINTER(BatchView)]),
    "aleph_hip_dist_alloc_rank": (i32, [P, u32, u32, u32, C.POINTER(P)]),
    "aleph_hip_gate_matrix": (i32, [C.POINTER(Gate), u32, PF64, PU32]),
    "aleph_hip_dist_exchange": (i32, [P, C.POINTER(P), u32, PU32, u32]),
    "aleph_hip_run_noisy": (i32, [P, u32, PU32, C.c_size_t, PF64, C.c_size_t, C.POINTER(Channel), u32, PF64, u64,
                                  u32, u64, u64, C.c_size_t, PU64]),
}

_lib = None


def load() -> C.CDLL:
    """Load the built library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m aleph_amd.build` (or __graft_entry__.build())")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def device_count() -> int:
    lib = load()
    c = i32(0)
    lib.aleph_hip_device_count(C.byref(c))
    return int(c.value)
